"""The reference's command-line front end: a triangle mesh (OBJ), a JSON file of
boundary-condition regions and force fields, the padded-cube voxelization and
the worst-case SIMP loop (robtop.cpp:18-48, config_parser.cpp:48-497,
boundary/BoundaryCondition.h, Grid.cpp:68-137 and :1291-1423).

Everything geometric runs on the host in FP64 with numpy; the mesh is
voxelized on the GPU (rt_voxelize_mesh) and the hierarchy, supports, loads and
the optimisation go through the library.

JSON (config_parser.cpp:385-451):
    {"model": "...obj",
     "fixedarea": [region, ...],     union of regions: supports
     "loadarea":  [region, ...],     union of regions: load nodes
     "force":     [{"type": ..., "region": region, ...}, ...]}   summed fields
Regions (BoundaryCondition.h): "sphere" {center, radius}, "halfspace" {p1, p2,
p3} (the side the normal (p2 - p1) x (p3 - p1) points to), "box" {p, q} (the
open box with diagonal p, q), "cylinder" {bottomCenter, heightVector, radius};
optional "closed" (boundary included, 1e-8 tolerance for spheres) and
"complement" (not for cylinders). Fields: "const" {norm, direction},
"RadialNormConstDir" {center, amplifier, dampRatio, direction},
"ConstNormRadialDir" {center, norm}, "RadialNormRadialDir" {NormCenter,
DirCenter, amplifier, dampRatio}, "Twist" {AxisStart, AxisEnd, amplifier}.

Nodes are placed at box_min + id * h, h the voxel size (Grid.cpp:1293-1323): a
surface node inside the fixed area is a support, otherwise a surface node
inside the load area is a load node with force = field(position). Outward
normals (for the constrained-direction worst case) and shell elements are
measured from the mesh itself: the closest point on the mesh comes from a
k-d tree over triangle centroids and an exact point-triangle distance over the
nearest candidates, instead of CGAL's AABB tree (Grid.cpp:1405-1423, :68-137).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

SMALL = 1e-8       # BoundaryCondition.h:23, :381


# ---------------------------------------------------------------------------
# mesh
# ---------------------------------------------------------------------------
def read_obj(path):
    """Vertices (n, 3) float32 and triangles (m, 3) int32 of an OBJ file;
    polygons are split into fans, negative indices count from the end."""
    vs, fs = [], []
    with open(path, "r", errors="replace") as fp:
        for line in fp:
            if line.startswith("v "):
                p = line.split()
                vs.append((float(p[1]), float(p[2]), float(p[3])))
            elif line.startswith("f "):
                idx = []
                for tok in line.split()[1:]:
                    i = int(tok.split("/")[0])
                    idx.append(i - 1 if i > 0 else len(vs) + i)
                for k in range(1, len(idx) - 1):
                    fs.append((idx[0], idx[k], idx[k + 1]))
    V = np.array(vs, np.float32).reshape(-1, 3)
    F = np.array(fs, np.int32).reshape(-1, 3)
    if F.size and (F.min() < 0 or F.max() >= len(V)):
        raise ValueError(f"{path}: face index out of range")
    return V, F


# ---------------------------------------------------------------------------
# regions (BoundaryCondition.h) on arrays of points P[n, 3]
# ---------------------------------------------------------------------------
class Sphere:
    def __init__(self, center, radius, closed=False, complement=False):
        self.c = np.asarray(center, np.float64)
        self.r2 = float(radius) ** 2
        self.closed, self.complement = closed, complement

    def contains(self, P):
        d2 = ((P - self.c) ** 2).sum(axis=1)
        inside = (d2 < self.r2) | (self.closed & ((d2 - self.r2) < SMALL))
        return inside ^ self.complement


class HalfSpace:
    """open half space on the side the normal (p2 - p1) x (p3 - p1) points to"""

    def __init__(self, p1, p2, p3, closed=False, complement=False):
        p1, p2, p3 = (np.asarray(p, np.float64) for p in (p1, p2, p3))
        self.n = np.cross(p2 - p1, p3 - p1)
        self.d = -self.n @ p1
        self.closed, self.complement = closed, complement

    def contains(self, P):
        s = P @ self.n + self.d
        return ((s > 0) | (self.closed & (s == 0))) ^ self.complement


class Box:
    """open axis-aligned box with diagonal corners p, q"""

    def __init__(self, p, q, closed=False, complement=False):
        p, q = np.asarray(p, np.float64), np.asarray(q, np.float64)
        self.lo, self.hi = np.minimum(p, q), np.maximum(p, q)
        self.closed, self.complement = closed, complement

    def contains(self, P):
        strict = np.all((P > self.lo) & (P < self.hi), axis=1)
        if self.closed:
            strict |= np.all((P >= self.lo) & (P <= self.hi), axis=1)
        return strict ^ self.complement


class Cylinder:
    """0 <= (p - b).z <= |h|, distance from the axis < radius (z = h / |h|)"""

    def __init__(self, bottom, radius, height):
        self.b = np.asarray(bottom, np.float64)
        h = np.asarray(height, np.float64)
        self.height = float(np.linalg.norm(h))
        self.z = h / self.height
        self.r = float(radius)

    def contains(self, P):
        d = P - self.b
        proj = d @ self.z
        devi = np.linalg.norm(np.cross(d, self.z), axis=1)
        return (proj <= self.height) & (proj >= 0) & (devi < self.r)


def region(desc):
    t = desc.get("type")
    closed = bool(desc.get("closed", False))
    comp = bool(desc.get("complement", False))
    if t == "sphere":
        return Sphere(desc.get("center", [0, 0, 0]), desc["radius"], closed, comp)
    if t == "halfspace":
        return HalfSpace(desc["p1"], desc["p2"], desc["p3"], closed, comp)
    if t == "box":
        return Box(desc["p"], desc["q"], closed, comp)
    if t == "cylinder":
        return Cylinder(desc["bottomCenter"], desc["radius"], desc["heightVector"])
    raise ValueError(f"unknown region type {t!r}")


class RegionUnion:
    def __init__(self, regions=()):
        self.regions = list(regions)

    def contains(self, P):
        P = np.atleast_2d(np.asarray(P, np.float64))
        out = np.zeros(len(P), bool)
        for r in self.regions:
            out |= r.contains(P)
        return out


# ---------------------------------------------------------------------------
# force fields (BoundaryCondition.h: norm generator x direction generator)
# ---------------------------------------------------------------------------
def _unit(v):
    n = np.linalg.norm(v, axis=-1, keepdims=True)
    return v / n


def _radial_norm(P, c, ampl, damp):
    r = np.linalg.norm(c - P, axis=1)
    return ampl * np.exp(-np.power(r, damp * 2))


class Field:
    def __init__(self, desc):
        self.t = desc["type"]
        self.region = region(desc["region"])
        self.d = desc

    def at(self, P):
        d, t = self.d, self.t
        n = len(P)
        if t == "const":
            v = np.broadcast_to(float(d["norm"]) * _unit(np.asarray(d["direction"], np.float64)), (n, 3))
        elif t == "RadialNormConstDir":
            v = (_radial_norm(P, np.asarray(d["center"], np.float64), float(d.get("amplifier", 1)),
                              float(d.get("dampRatio", 1)))[:, None]
                 * _unit(np.asarray(d["direction"], np.float64)))
        elif t == "ConstNormRadialDir":
            v = float(d["norm"]) * _unit(np.asarray(d["center"], np.float64) - P)
        elif t == "RadialNormRadialDir":
            v = (_radial_norm(P, np.asarray(d["NormCenter"], np.float64), float(d.get("amplifier", 1)),
                              float(d.get("dampRatio", 1)))[:, None]
                 * _unit(np.asarray(d["DirCenter"], np.float64) - P))
        elif t == "Twist":
            a, b = np.asarray(d["AxisStart"], np.float64), np.asarray(d["AxisEnd"], np.float64)
            nrm = np.linalg.norm(np.cross(P - a, _unit(b - a)), axis=1) * float(d["amplifier"])
            v = nrm[:, None] * _unit(np.cross(P - a, b - a))
        elif t == "gradient":                      # parsed and ignored by the reference
            v = np.zeros((n, 3))
        else:
            raise ValueError(f"unknown force field type {t!r}")
        return np.where(self.region.contains(P)[:, None], v, 0.0)


class ForceField:
    def __init__(self, fields=()):
        self.fields = list(fields)

    def at(self, P):
        P = np.atleast_2d(np.asarray(P, np.float64))
        out = np.zeros((len(P), 3))
        for f in self.fields:
            out += f.at(P)
        return out


class BoundaryConfig:
    """parse_benchmark_configuration (config_parser.cpp:385-451)"""

    def __init__(self, doc):
        self.model = doc.get("model")
        self.fixed = RegionUnion(region(r) for r in doc.get("fixedarea", []))
        self.load = RegionUnion(region(r) for r in doc.get("loadarea", []))
        self.force = ForceField(Field(f) for f in doc.get("force", []))

    @classmethod
    def from_file(cls, path):
        with open(path) as fp:
            return cls(json.load(fp))


# ---------------------------------------------------------------------------
# distances to the mesh (replaces CGAL's AABB tree queries)
# ---------------------------------------------------------------------------
def _closest_on_triangles(P, A, B, C):
    """closest points on triangles (A, B, C) to points P, all (n, 3) (the
    Voronoi-region walk of Ericson, Real-Time Collision Detection, 5.1.5)"""
    ab, ac, ap = B - A, C - A, P - A
    d1, d2 = (ab * ap).sum(1), (ac * ap).sum(1)
    bp = P - B
    d3, d4 = (ab * bp).sum(1), (ac * bp).sum(1)
    cp = P - C
    d5, d6 = (ab * cp).sum(1), (ac * cp).sum(1)
    vc = d1 * d4 - d3 * d2
    vb = d5 * d2 - d1 * d6
    va = d3 * d6 - d5 * d4
    out = np.empty_like(P)
    done = np.zeros(len(P), bool)

    def put(m, val):
        m = m & ~done
        out[m] = val[m] if val.ndim == 2 else val
        done[:] |= m

    put((d1 <= 0) & (d2 <= 0), A)
    put((d3 >= 0) & (d4 <= d3), B)
    with np.errstate(divide="ignore", invalid="ignore"):
        v = d1 / (d1 - d3)
        put((vc <= 0) & (d1 >= 0) & (d3 <= 0), A + v[:, None] * ab)
        put((d6 >= 0) & (d5 <= d6), C)
        w = d2 / (d2 - d6)
        put((vb <= 0) & (d2 >= 0) & (d6 <= 0), A + w[:, None] * ac)
        w2 = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        put((va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0), B + w2[:, None] * (C - B))
        den = 1.0 / (va + vb + vc)
        v2, w3 = vb * den, vc * den
        put(np.ones(len(P), bool), A + v2[:, None] * ab + w3[:, None] * ac)
    return out


class MeshDistance:
    """closest point, distance and face of a triangle mesh for many points"""

    def __init__(self, V, F, k=24):
        from scipy.spatial import cKDTree
        self.T = np.asarray(V, np.float64)[np.asarray(F)]          # (m, 3, 3)
        self.tree = cKDTree(self.T.mean(axis=1))
        self.k = min(k, len(F))
        self.N = np.cross(self.T[:, 1] - self.T[:, 0], self.T[:, 2] - self.T[:, 0])   # face orientation

    def closest(self, P, chunk=65536):
        P = np.atleast_2d(np.asarray(P, np.float64))
        best_d = np.full(len(P), np.inf)
        best_c = np.zeros_like(P)
        best_f = np.zeros(len(P), np.int64)
        for s in range(0, len(P), chunk):
            Q = P[s:s + chunk]
            _, cand = self.tree.query(Q, k=self.k)
            cand = cand.reshape(len(Q), -1)
            for j in range(cand.shape[1]):
                f = cand[:, j]
                T = self.T[f]
                c = _closest_on_triangles(Q, T[:, 0], T[:, 1], T[:, 2])
                d = ((c - Q) ** 2).sum(1)
                better = d < best_d[s:s + chunk]
                best_d[s:s + chunk][better] = d[better]
                best_c[s:s + chunk][better] = c[better]
                best_f[s:s + chunk][better] = f[better]
        return best_c, np.sqrt(best_d), best_f

    def outward_normals(self, P):
        """Grid::outwardNormal (Grid.cpp:1405-1423): the unit vector from the
        point to its closest point, oriented along that face's normal; the
        face's own normal when the point lies on the mesh (within 1e-6)"""
        c, d, f = self.closest(P)
        fn = self.N[f]
        v = c - np.atleast_2d(P)
        on = d < 1e-6
        sgn = np.where((v * fn).sum(1) < 0, -1.0, 1.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            n = np.where(on[:, None], _unit(fn), v / (d * sgn)[:, None])
        return n


# ---------------------------------------------------------------------------
# problem set-up from a mesh and a configuration
# ---------------------------------------------------------------------------
class MeshProblem:
    """voxelize (buildGrids, optimization.cpp:14-30) and set supports and
    loads (Grid::computeProjectionMatrix, Grid.cpp:1291-1403)"""

    def __init__(self, V, F, config, gridreso, youngs=1e6, poisson=0.4, power=3.0, shell_width=0, mode="wsff"):
        from .grid import HierarchyGrid, voxelize_mesh
        if mode not in ("wsff", "wscf", "nsff", "nscf"):
            raise ValueError("work modes: wsff / wscf (with supports), nsff / nscf (without); ff = free load "
                             "directions, cf = normal loads")
        self.V, self.F, self.cfg, self.mode = V, F, config, mode
        t0 = time.perf_counter()
        bits, R, box = voxelize_mesh(V, F, gridreso)
        self.R, self.box = R, box.astype(np.float64)
        self.h = (self.box[1, 0] - self.box[0, 0]) / R
        self.hg = HierarchyGrid.genFromMesh(bits, R)
        self.hg.set_material(youngs, poisson, power)
        self.t_voxelize = time.perf_counter() - t0
        g0 = self.hg[0]
        nx, ny, nz = g0.nodes
        ox, oy, oz = g0.origin
        idx = np.stack(np.meshgrid(np.arange(nx) + ox, np.arange(ny) + oy, np.arange(nz) + oz, indexing="ij"), -1)
        self.node_pos = self.box[0] + idx.transpose(2, 1, 0, 3).reshape(-1, 3) * self.h   # x fastest
        surf = g0.surface_nodes()
        fixed = np.zeros_like(surf)
        if mode.startswith("ws"):        # the no-support modes leave the body free
            fixed[surf] = config.fixed.contains(self.node_pos[surf])
        load = np.zeros_like(surf)
        rest = surf & ~fixed
        load[rest] = config.load.contains(self.node_pos[rest])
        self.fixed, self.load = fixed, load
        f = np.zeros((3, surf.size))
        if load.any():
            f[:, load] = config.force.at(self.node_pos[load]).T
        self.force = f
        self.hg.set_support(fixed)
        self.hg.set_force(f)
        self.normals = None
        self.mesh = MeshDistance(V, F) if (mode.endswith("cf") or shell_width > 0) else None
        if load.any():
            if mode.endswith("cf"):
                nrm = np.zeros((3, surf.size))
                nrm[:, load] = self.mesh.outward_normals(self.node_pos[load]).T
                self.normals = nrm
                self.n_load = self.hg.set_load_nodes(load.astype(np.uint8), nrm)
            else:
                self.n_load = self.hg.set_load_nodes(load.astype(np.uint8))
        else:
            self.n_load = 0
        self.shell = None
        if shell_width > 0:
            self.shell = self.shell_elements(shell_width)
            self.hg.set_shell(self.shell)

    def shell_elements(self, width):
        """setSolidShellElement (Grid.cpp:68-137): solid elements whose centre
        lies closer than width * h to the mesh"""
        from scipy import ndimage
        g0 = self.hg[0]
        eact, _ = g0.masks()
        ex, ey, ez = g0.elements
        act = eact.reshape(ez, ey, ex)
        # candidates: solid elements within width + 2 voxels of a void voxel (or the box side)
        band = ndimage.distance_transform_edt(np.pad(act, 1)) [1:-1, 1:-1, 1:-1] <= width + 2
        cand = act & band
        zz, yy, xx = np.nonzero(cand)
        ox, oy, oz = g0.origin
        cen = self.box[0] + (np.stack([xx + ox, yy + oy, zz + oz], 1) + 0.5) * self.h
        shell = np.zeros(act.shape, bool)
        if len(cen):
            _, d, _ = self.mesh.closest(cen)
            shell[zz, yy, xx] = d * d < (width * self.h) ** 2
        return shell.reshape(-1)


def main(argv=None):
    """robtop.cpp:18-48 with the reference's flag names (config_parser.cpp)"""
    ap = argparse.ArgumentParser(description="worst-case topology optimisation of a triangle mesh")
    ap.add_argument("--jsonfile", required=True)
    ap.add_argument("--meshfile", default=None, help="OBJ mesh (default: the JSON's model next to it)")
    ap.add_argument("--outdir", default="./out")
    ap.add_argument("--power_penalty", type=float, default=3.0)
    ap.add_argument("--volume_ratio", type=float, default=0.3)
    ap.add_argument("--youngs_modulus", type=float, default=1e6)
    ap.add_argument("--poisson_ratio", type=float, default=0.4)
    ap.add_argument("--damp_ratio", type=float, default=0.5)
    ap.add_argument("--design_step", type=float, default=0.03)
    ap.add_argument("--vol_reduction", type=float, default=0.05)
    ap.add_argument("--filter_radius", type=int, default=2)
    ap.add_argument("--gridreso", type=int, default=200)
    ap.add_argument("--max_itn", type=int, default=100)
    ap.add_argument("--min_density", type=float, default=1e-3)
    ap.add_argument("--shell_width", type=float, default=0)
    ap.add_argument("--workmode", default="wscf", choices=["wsff", "wscf", "nsff", "nscf"])
    ap.add_argument("--method", default="lanczos", choices=["pm", "lanczos"],
                    help="worst-load search: the reference's modified power method or block Lanczos")
    a = ap.parse_args(argv)
    if a.workmode.startswith("ns") and a.method == "lanczos":
        a.method = "pm"                  # block Lanczos solves need supports
    from .optimization import Parameter, worst_case_topopt
    cfg = BoundaryConfig.from_file(a.jsonfile)
    mesh = a.meshfile or os.path.join(os.path.dirname(os.path.abspath(a.jsonfile)), cfg.model)
    V, F = read_obj(mesh)
    prob = MeshProblem(V, F, cfg, a.gridreso, a.youngs_modulus, a.poisson_ratio, a.power_penalty, a.shell_width,
                       a.workmode)
    ne, nn = prob.hg.num_active(0)
    print(f"[robtop_amd] {mesh}: R = {prob.R}, {ne} elements, {nn} nodes, {int(prob.fixed.sum())} supports, "
          f"{prob.n_load} load nodes", flush=True)
    params = Parameter(volume_ratio=a.volume_ratio, volume_decrease=a.vol_reduction, design_step=a.design_step,
                       damp_ratio=a.damp_ratio, power_penalty=a.power_penalty, min_rho=a.min_density,
                       filter_radius=a.filter_radius)
    os.makedirs(a.outdir, exist_ok=True)
    recs = worst_case_topopt(prob.hg, params, n_iter=a.max_itn, method=a.method,
                             lanczos=dict(k=8, max_blocks=20, tol=1e-6, solve_tol=1e-6, pre=2, post=2),
                             log=lambda r: print(f"iter {r['iter']}: worst compliance {r['worst_compliance']:.6e}, "
                                                 f"{r['seconds']:.2f} s", flush=True))
    np.save(os.path.join(a.outdir, "rho.npy"), prob.hg.get_density())
    with open(os.path.join(a.outdir, "history.json"), "w") as fp:
        json.dump(recs, fp, default=float)
    prob.hg.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
