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
