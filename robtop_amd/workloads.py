"""Synthetic problem set-ups: voxel geometry, supports, loads and densities.

The cantilever is the reference's bits-input set-up (HierarchyGrid::genFromMesh
with h = 1, Grid.cpp:455-461) with a fixed x = 0 face and a -z load spread over
the x = max face (the reference's ``const`` load field assigns the full vector
per node, config_parser.cpp:286-292; here it is divided by the node count so
the total load is 1).
"""
import numpy as np


def cube_bits(active_xyz):
    """LSB-first uint32 words of a boolean [R, R, R] array indexed [x, y, z]."""
    return _pack_lsb(np.ascontiguousarray(active_xyz.transpose(2, 1, 0)).reshape(-1))


def _pack_lsb(flat):
    b = np.packbits(flat.astype(np.uint8), bitorder='little')
    pad = (-b.size) % 4
    if pad:
        b = np.concatenate([b, np.zeros(pad, np.uint8)])
    return b.view('<u4').astype(np.uint32)


def box_bits(R, ex, ey, ez):
    """Solid box [0, ex) x [0, ey) x [0, ez) in a cube of side R."""
    a = np.zeros((R, R, R), bool)
    a[:ex, :ey, :ez] = True
    return _pack_lsb(a.transpose(2, 1, 0).reshape(-1))


def cube_side(ex, ey, ez):
    """Smallest power of two >= the longest side (the reference pads to 2^k > gridreso)."""
    m = max(ex, ey, ez)
    R = 1
    while R < m:
        R <<= 1
    return R


def cantilever_bc(hg):
    """Support flags and load vector (box layout of the finest level) for the
    cantilever: fixed nodes at x = 0, F_z = -1/N_L on the existing nodes of the
    last x plane of the solid."""
    g0 = hg[0]
    _, vact = g0.masks()
    nx, ny, nz = g0.nodes
    v = vact.reshape(nz, ny, nx)
    ox = g0.origin[0]
    xs = np.arange(nx) + ox
    fixed = np.zeros((nz, ny, nx), bool)
    fixed[:, :, xs == 0] = True
    fixed &= v
    xmax = np.max(np.nonzero(v.any(axis=(0, 1)))[0])
    load = np.zeros((nz, ny, nx), bool)
    load[:, :, xmax] = v[:, :, xmax]
    load &= ~fixed
    nl = int(load.sum())
    f = np.zeros((3, nz, ny, nx))
    f[2][load] = -1.0 / nl
    return fixed.reshape(-1), f.reshape(3, -1), nl


def smoothed_random_density(eact_box, shape_xyz, seed, sigma=4.0, mean=0.4, scale=0.2, rho_min=1e-3):
    """Seeded Gaussian-smoothed random field mapped to [rho_min, 1] with the given
    mean over the solid elements. ``eact_box`` is the element mask in box layout
    (x fastest) of dimensions ``shape_xyz``."""
    from scipy.ndimage import gaussian_filter
    ex, ey, ez = shape_xyz
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((ez, ey, ex), dtype=np.float32)
    z = gaussian_filter(z, sigma=sigma, mode="nearest", truncate=3.0)
    act = eact_box.reshape(ez, ey, ex)
    zz = z[act]
    z = (z - zz.mean()) / max(zz.std(), 1e-12)
    base = mean + scale * z

    def avg(shift):
        return np.clip(base[act] + shift, rho_min, 1.0).astype(np.float64).mean()

    lo, hi = -1.0, 1.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if avg(mid) < mean:
            lo = mid
        else:
            hi = mid
        if hi - lo < 1e-9:
            break
    rho = np.clip(base + 0.5 * (lo + hi), rho_min, 1.0).astype(np.float32)
    rho[~act] = 0
    return rho.reshape(-1)


def _torus_sdf(px, py, pz, n, R_major, r_minor):
    """Signed distance to a torus about the z axis through the centre of the
    cube [0, n]^3; radii are fractions of the edge n."""
    c = n / 2.0
    q = np.hypot(px - c, py - c) - R_major * n
    return np.hypot(q, pz - c) - r_minor * n


def torus_bits(n, R_major=0.35, r_minor=0.12):
    """Analytic torus voxelized at n^3 by the sign of its distance at voxel
    centres, placed at the origin of the smallest power-of-two cube R >= n.
    Returns (bits, R, solid [n, n, n] indexed [x, y, z])."""
    c = np.arange(n) + 0.5
    px, py, pz = np.meshgrid(c, c, c, indexing="ij")
    solid = _torus_sdf(px, py, pz, n, R_major, r_minor) < 0
    R = cube_side(n, n, n)
    a = np.zeros((R, R, R), bool)
    a[:n, :n, :n] = solid
    return cube_bits(a), R, solid


def torus_worst_case_bc(hg, n, R_major=0.35, r_minor=0.12, fixed_fraction=0.05):
    """Supports, load nodes and load normals (finest node box) for the
    worst-case torus: nodes of the solid within the bottom `fixed_fraction` of
    its height are fixed; every other surface node is a load node whose
    admissible load is along the outward normal, the normalised gradient of the
    signed distance at the node."""
    g0 = hg[0]
    _, vact = g0.masks()
    surf = g0.surface_nodes()
    nx, ny, nz = g0.nodes
    ox, oy, oz = g0.origin
    zs = np.arange(nz) + oz
    zsolid = zs[np.nonzero(vact.reshape(nz, ny, nx).any(axis=(1, 2)))[0]]
    zcut = zsolid.min() + fixed_fraction * (zsolid.max() - zsolid.min())
    Z = np.broadcast_to(zs[:, None, None], (nz, ny, nx)).reshape(-1)
    fixed = vact & (Z <= zcut)
    load = surf & ~fixed
    X = np.broadcast_to((np.arange(nx) + ox)[None, None, :], (nz, ny, nx)).reshape(-1).astype(np.float64)
    Y = np.broadcast_to((np.arange(ny) + oy)[None, :, None], (nz, ny, nx)).reshape(-1).astype(np.float64)
    Zf = Z.astype(np.float64)
    h = 1e-3
    grad = np.stack([
        _torus_sdf(X + h, Y, Zf, n, R_major, r_minor) - _torus_sdf(X - h, Y, Zf, n, R_major, r_minor),
        _torus_sdf(X, Y + h, Zf, n, R_major, r_minor) - _torus_sdf(X, Y - h, Zf, n, R_major, r_minor),
        _torus_sdf(X, Y, Zf + h, n, R_major, r_minor) - _torus_sdf(X, Y, Zf - h, n, R_major, r_minor)])
    normals = grad / np.maximum(np.linalg.norm(grad, axis=0), 1e-30)[None]
    normals[:, ~load] = 0.0
    return fixed, load, normals
