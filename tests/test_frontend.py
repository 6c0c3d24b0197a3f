"""The JSON + OBJ front end (robtop_amd/frontend.py; config_parser.cpp:48-497,
boundary/BoundaryCondition.h, Grid.cpp:68-137 and :1291-1423).

CPU: the OBJ reader, every region and force-field type on hand-worked points
(boundary points for open / closed regions, complements, the half space's
orientation from its three points), the closest-point search against brute
force over all triangles, and outward normals of a sphere against the radial
direction. GPU: a 24 x 12 x 24 box mesh with the reference's cube24x24x12
configuration (supports on the half space x < 0.45, loads in a sphere of
radius 1.5 around (24, 6, 12)) is voxelized, its supports and loads equal a
direct evaluation at the surface nodes, the shell elements equal a brute-force
distance test, and the command-line driver runs two worst-case iterations.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from robtop_amd import frontend as fe

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def box_mesh(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    V = np.array([[x0, y0, z0], [x1, y0, z0], [x0, y1, z0], [x1, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x0, y1, z1], [x1, y1, z1]], np.float32)
    F = np.array([[0, 2, 1], [1, 2, 3], [4, 5, 6], [5, 7, 6], [0, 1, 4], [1, 5, 4],
                  [2, 6, 3], [3, 6, 7], [0, 4, 2], [2, 4, 6], [1, 3, 5], [3, 7, 5]], np.int32)
    return V, F


def sphere_mesh(c, r, nu=48, nv=24):
    th = np.linspace(0, np.pi, nv + 1)[1:-1]
    ph = np.linspace(0, 2 * np.pi, nu, endpoint=False)
    T, P = np.meshgrid(th, ph, indexing="ij")
    V = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    V = np.vstack([[0, 0, 1], V, [0, 0, -1]]) * r + np.asarray(c)
    F = []
    ring = lambda i, j: 1 + i * nu + (j % nu)
    for j in range(nu):
        F.append([0, ring(0, j), ring(0, j + 1)])
    for i in range(nv - 2):
        for j in range(nu):
            a, b, cc, d = ring(i, j), ring(i, j + 1), ring(i + 1, j), ring(i + 1, j + 1)
            F += [[a, cc, b], [b, cc, d]]
    last = len(V) - 1
    for j in range(nu):
        F.append([ring(nv - 2, j), last, ring(nv - 2, j + 1)])
    return V.astype(np.float32), np.array(F, np.int32)


def write_obj(path, V, F):
    with open(path, "w") as fp:
        fp.write("# box\n")
        for v in V:
            fp.write(f"v {v[0]:.9g} {v[1]:.9g} {v[2]:.9g}\n")
        for f in F:
            fp.write(f"f {f[0] + 1} {f[1] + 1} {f[2] + 1}\n")


def test_read_obj_polygons_and_index_forms(tmp_path):
    p = tmp_path / "q.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\nvt 0 0\n"
                 "f 1/1/1 2/1/1 3/1/1 4/1/1\nf -4 -2 -1\n")
    V, F = fe.read_obj(p)
    assert V.shape == (4, 3) and F.tolist() == [[0, 1, 2], [0, 2, 3], [0, 2, 3]]


def test_regions_hand_worked():
    P = np.array([[1.0, 0, 0], [0.5, 0, 0], [2.0, 0, 0], [0, 0, 0]])
    s_open, s_closed = fe.Sphere([0, 0, 0], 1.0), fe.Sphere([0, 0, 0], 1.0, closed=True)
    assert s_open.contains(P).tolist() == [False, True, False, True]
    assert s_closed.contains(P).tolist() == [True, True, False, True]
    assert fe.Sphere([0, 0, 0], 1.0, complement=True).contains(P).tolist() == [True, False, True, False]
    # plane through p1, p2, p3 with normal (p2 - p1) x (p3 - p1) = +z: the side z > 36.8
    h = fe.HalfSpace([0, 0, 36.8], [1, 0, 36.8], [0, 1, 36.8])
    Q = np.array([[5, 5, 40.0], [5, 5, 36.8], [5, 5, 30.0]])
    assert h.contains(Q).tolist() == [True, False, False]
    assert fe.HalfSpace([0, 0, 36.8], [1, 0, 36.8], [0, 1, 36.8], closed=True).contains(Q).tolist() == [True, True, False]
    # reversed orientation flips the side (the cube24x24x12 support: x < 0.45)
    hx = fe.HalfSpace([0.45, 0, 0], [0.45, 0, 1], [0.45, 1, 0])
    assert hx.contains(np.array([[0.0, 3, 3], [1.0, 3, 3]])).tolist() == [True, False]
    b = fe.Box([2, 0, 0], [0, 1, 1])
    B = np.array([[1, 0.5, 0.5], [2, 0.5, 0.5], [3, 0.5, 0.5]])
    assert b.contains(B).tolist() == [True, False, False]
    assert fe.Box([2, 0, 0], [0, 1, 1], closed=True).contains(B).tolist() == [True, True, False]
    c = fe.Cylinder([0, 0, 0], 1.0, [0, 0, 2])
    C = np.array([[0.5, 0, 1], [0.5, 0, 2], [0.5, 0, 2.1], [1.0, 0, 1], [0, 0, -0.1]])
    assert c.contains(C).tolist() == [True, True, False, False, False]
    u = fe.RegionUnion([fe.Sphere([0, 0, 0], 1), fe.Sphere([3, 0, 0], 1)])
    assert u.contains(np.array([[0, 0, 0.5], [3, 0.5, 0], [1.5, 0, 0]])).tolist() == [True, True, False]


def test_force_fields_hand_worked():
    everywhere = {"type": "sphere", "center": [0, 0, 0], "radius": 1e9}
    P = np.array([[1.0, 0, 0], [0, 2.0, 0]])
    f = fe.Field({"type": "const", "norm": 2, "direction": [0, 0, -3], "region": everywhere})
    assert np.allclose(f.at(P), [[0, 0, -2], [0, 0, -2]])
    f = fe.Field({"type": "ConstNormRadialDir", "norm": 3, "center": [0, 0, 0], "region": everywhere})
    assert np.allclose(f.at(P), [[-3, 0, 0], [0, -3, 0]])
    f = fe.Field({"type": "RadialNormConstDir", "center": [0, 0, 0], "amplifier": 2, "dampRatio": 1,
                  "direction": [1, 0, 0], "region": everywhere})
    assert np.allclose(f.at(P), [[2 * np.exp(-1), 0, 0], [2 * np.exp(-4), 0, 0]])
    f = fe.Field({"type": "RadialNormRadialDir", "NormCenter": [0, 0, 0], "DirCenter": [0, 0, 1], "amplifier": 1,
                  "dampRatio": 0.5, "region": everywhere})
    assert np.allclose(f.at(P)[0], np.exp(-1) * np.array([-1, 0, 1]) / np.sqrt(2))
    # twist about the z axis: |p x z| amplifier along (p - a) x (b - a)
    f = fe.Field({"type": "Twist", "AxisStart": [0, 0, 0], "AxisEnd": [0, 0, 2], "amplifier": 3, "region": everywhere})
    assert np.allclose(f.at(P), [[0, -3, 0], [6, 0, 0]])
    ff = fe.ForceField([fe.Field({"type": "const", "norm": 1, "direction": [1, 0, 0],
                                  "region": {"type": "sphere", "center": [0, 0, 0], "radius": 1.5}}),
                        fe.Field({"type": "const", "norm": 1, "direction": [0, 1, 0],
                                  "region": {"type": "box", "p": [-1, -1, -1], "q": [3, 3, 1]}})])
    assert np.allclose(ff.at(P), [[1, 1, 0], [0, 1, 0]])


def test_closest_points_match_brute_force():
    rng = np.random.default_rng(0)
    V, F = sphere_mesh((0.2, -0.1, 0.3), 1.3, 16, 8)
    V = V + rng.normal(0, 0.05, V.shape).astype(np.float32)
    md = fe.MeshDistance(V, F, k=24)
    P = rng.uniform(-2, 2, (300, 3))
    c, d, f = md.closest(P)
    T = V.astype(np.float64)[F]
    best = np.full(len(P), np.inf)
    for t in range(len(F)):
        q = fe._closest_on_triangles(P, np.broadcast_to(T[t, 0], P.shape), np.broadcast_to(T[t, 1], P.shape),
                                     np.broadcast_to(T[t, 2], P.shape))
        best = np.minimum(best, np.linalg.norm(q - P, axis=1))
    assert np.allclose(d, best, rtol=1e-12, atol=1e-12)
    assert np.allclose(np.linalg.norm(c - P, axis=1), d)
    # the closest point on one triangle against a dense sample of it
    A, B, C = T[0]
    w = rng.dirichlet(np.ones(3), 20000)
    S = w @ np.stack([A, B, C])
    for p in P[:20]:
        q = fe._closest_on_triangles(p[None], A[None], B[None], C[None])[0]
        assert np.linalg.norm(q - p) <= np.min(np.linalg.norm(S - p, axis=1)) + 1e-12


def test_outward_normals_of_a_sphere():
    V, F = sphere_mesh((0, 0, 0), 10.0, 96, 48)
    md = fe.MeshDistance(V, F)
    rng = np.random.default_rng(1)
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for r in (9.5, 10.5):
        n = md.outward_normals(d * r)
        assert np.all(np.sum(n * d, axis=1) > 0.98), r


def cube_config():
    """the reference's bench/cube24x24x12 configuration (JSON content)"""
    return {"model": "cube24x24x12.obj",
            "fixedarea": [{"type": "halfspace", "p1": [0.45, 0.0, 0.0], "p2": [0.45, 0.0, 1.0], "p3": [0.45, 1.0, 0.0]}],
            "force": [],
            "loadarea": [{"type": "sphere", "radius": 1.5, "center": [24.0, 6.0, 12.0]}]}


@pytest.mark.gpu
def test_mesh_problem_supports_loads_and_shell(tmp_path):
    V, F = box_mesh((0, 0, 0), (24, 12, 24))
    cfg = fe.BoundaryConfig(cube_config())
    print("[frontend] box 24x12x24, gridreso 47, wscf, shell width 2", flush=True)
    prob = fe.MeshProblem(V, F, cfg, 47, shell_width=2, mode="wscf")
    hg = prob.hg
    assert prob.R == 64
    g0 = hg[0]
    surf = g0.surface_nodes()
    P = prob.node_pos
    fixed = surf & (P[:, 0] < 0.45)
    load = surf & ~fixed & (((P - [24, 6, 12]) ** 2).sum(1) < 2.25)
    assert np.array_equal(prob.fixed, fixed) and fixed.sum() > 0
    assert np.array_equal(prob.load, load) and prob.n_load == int(load.sum()) > 0
    # normals of the load nodes on the x = 24 face point along +x
    nrm = prob.normals[:, load]
    onface = np.abs(P[load, 0] - 24) < 0.6
    assert np.all(nrm[0, onface] > 0.7)
    # shell: solid elements whose centres are closer than 2 h to the box surface
    eact, _ = g0.masks()
    ex, ey, ez = g0.elements
    ox, oy, oz = g0.origin
    zz, yy, xx = np.meshgrid(np.arange(ez), np.arange(ey), np.arange(ex), indexing="ij")
    cen = prob.box[0] + (np.stack([xx + ox, yy + oy, zz + oz], -1).reshape(-1, 3) + 0.5) * prob.h
    dist = np.min(np.stack([cen[:, 0], 24 - cen[:, 0], cen[:, 1], 12 - cen[:, 1], cen[:, 2], 24 - cen[:, 2]]), axis=0)
    want = eact & (dist * dist < (2 * prob.h) ** 2)
    assert np.array_equal(prob.shell, want) and want.sum() > 0
    hg.init_rho(0.5)
    hg.update_stencil()
    st = hg.solve(tol=1e-6, maxit=500)
    assert st.converged
    hg.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["wscf", "nsff"])
def test_command_line_driver(tmp_path, mode):
    """The front end end to end: with supports and normal loads (block
    Lanczos), and as a free body with balanced free loads (the modified power
    method, rigid motion removed)."""
    V, F = box_mesh((0, 0, 0), (24, 12, 24))
    write_obj(tmp_path / "cube24x24x12.obj", V, F)
    with open(tmp_path / "config1.json", "w") as fp:
        json.dump(cube_config(), fp)
    print(f"[frontend] python -m robtop_amd.frontend, {mode}, 2 iterations", flush=True)
    p = subprocess.run([sys.executable, "-m", "robtop_amd.frontend", "--jsonfile", str(tmp_path / "config1.json"),
                        "--gridreso", "47", "--max_itn", "2", "--workmode", mode, "--outdir", str(tmp_path / "out")],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert p.returncode == 0, p.stderr[-2000:]
    hist = json.load(open(tmp_path / "out" / "history.json"))
    assert len(hist) == 2 and all(np.isfinite(h["worst_compliance"]) and h["worst_compliance"] > 0 for h in hist)
    assert np.load(tmp_path / "out" / "rho.npy").size > 0
