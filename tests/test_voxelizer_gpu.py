"""GPU solid voxelizer (k_voxel.hip, rt_voxelize_mesh) against the oracle's
restatement of voxelize_solid.cu:78-150 / voxelize_main.cpp:312-404, bit for
bit, on closed meshes at cube sides 64, 128 and 256:

* box meshes (faces normal to y and z have n.x = 0 and are skipped);
* a triangulated sphere and torus in general position;
* lattice meshes whose vertices sit on voxel centres, so that triangle edges
  and vertices pass exactly through voxel centres and the top-left rule
  decides every coincidence (a mesh spanning [0, reso]^3 is padded into a cube
  with bmin = -0.5 and unit 1, so voxel centre i lies at the integer i);
* the C++ facade's buildGrids(coords, trifaces) (optimization.cpp:14-30), whose
  finest element mask must equal the hierarchy built from the oracle's bits.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE2 = os.path.join(ROOT, "tests", "cpp", "facade_worst_mma")


def box_mesh(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    V = np.array([[x0, y0, z0], [x1, y0, z0], [x0, y1, z0], [x1, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x0, y1, z1], [x1, y1, z1]], np.float32)
    F = np.array([[0, 2, 1], [1, 2, 3], [4, 5, 6], [5, 7, 6], [0, 1, 4], [1, 5, 4],
                  [2, 6, 3], [3, 6, 7], [0, 4, 2], [2, 4, 6], [1, 3, 5], [3, 7, 5]], np.int32)
    return V, F


def sphere_mesh(c, r, nu=40, nv=20):
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


def torus_mesh(c, R, r, nu=48, nv=24):
    u = np.linspace(0, 2 * np.pi, nu, endpoint=False)
    v = np.linspace(0, 2 * np.pi, nv, endpoint=False)
    U, W = np.meshgrid(u, v, indexing="ij")
    V = np.stack([(R + r * np.cos(W)) * np.cos(U), (R + r * np.cos(W)) * np.sin(U), r * np.sin(W)], -1)
    V = V.reshape(-1, 3) + np.asarray(c)
    idx = lambda i, j: (i % nu) * nv + (j % nv)
    F = []
    for i in range(nu):
        for j in range(nv):
            F += [[idx(i, j), idx(i + 1, j), idx(i, j + 1)], [idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)]]
    return V.astype(np.float32), np.array(F, np.int32)


def octahedron_lattice(n):
    """integer vertices spanning [0, n]^3 (n odd): every edge runs through voxel centres"""
    h = (n - 1) // 2
    V = np.array([[0, h, h], [n, h, h], [h, 0, h], [h, n, h], [h, h, 0], [h, h, n]], np.float32)
    F = np.array([[0, 2, 4], [0, 4, 3], [0, 3, 5], [0, 5, 2], [1, 4, 2], [1, 3, 4], [1, 5, 3], [1, 2, 5]], np.int32)
    return V, F


def snapped_sphere(n, nu=24, nv=12):
    """sphere through [0, n]^3 with vertices rounded to integers (slivers and
    zero-area faces included)"""
    V, F = sphere_mesh((n / 2, n / 2, n / 2), n / 2, nu, nv)
    V = np.rint(V).astype(np.float32)
    V[np.argmin(V[:, 0]), 0] = 0
    V[np.argmax(V[:, 0]), 0] = n
    return V, F


def meshes(reso):
    s = float(reso)
    yield "box", box_mesh((0.1 * s, 0.2 * s, 0.15 * s), (0.9 * s, 0.55 * s, 0.6 * s))
    yield "sphere", sphere_mesh((0.5 * s, 0.45 * s, 0.52 * s), 0.37 * s)
    yield "torus", torus_mesh((0.5 * s, 0.5 * s, 0.5 * s), 0.3 * s, 0.1 * s)
    yield "lattice box", box_mesh((0, 0, 0), (reso, reso // 2, reso // 3))
    yield "lattice octahedron", octahedron_lattice(reso)
    yield "snapped sphere", snapped_sphere(reso)


@pytest.mark.gpu
@pytest.mark.parametrize("reso", [63, 127, 255])
def test_gpu_voxelizer_bit_exact(reso):
    import robtop_amd as rt
    for name, (V, F) in meshes(reso):
        print(f"[vox] {name} reso {reso}", flush=True)
        bo, Ro, boxo = oracle.voxelize_mesh(V, F, reso)
        bg, Rg, boxg = rt.voxelize_mesh(V, F, reso)
        assert Rg == Ro and np.array_equal(boxg, boxo), name
        assert np.array_equal(bg, bo), (name, int(np.count_nonzero(bg != bo)))
        assert np.count_nonzero(bo) > 0, name


def _fnv(mask):
    h = 1469598103934665603
    for v in mask.astype(np.uint8).tobytes():
        h = ((h ^ v) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"


@pytest.mark.gpu
@pytest.mark.parametrize("reso", [63, 127])
def test_facade_build_grids_from_mesh(tmp_path, reso):
    import robtop_amd as rt
    V, F = torus_mesh((0.5 * reso, 0.5 * reso, 0.5 * reso), 0.3 * reso, 0.1 * reso)
    path = tmp_path / "mesh.bin"
    with open(path, "wb") as fp:
        fp.write(np.array([len(V), len(F)], np.int32).tobytes())
        fp.write(V.astype(np.float32).tobytes())
        fp.write(F.astype(np.int32).tobytes())
    print(f"[cpp] facade_worst_mma vox torus reso {reso}", flush=True)
    p = subprocess.run([EXE2, "vox", str(path), str(reso)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    r = json.loads([l for l in p.stdout.splitlines() if l.startswith("{")][-1])
    bits, R, box = oracle.voxelize_mesh(V, F, reso)
    assert r["R"] == R
    assert np.array_equal(np.array(r["box"], np.float32).reshape(2, 3), box)   # %.9g round-trips a float
    hg = rt.HierarchyGrid.genFromMesh(bits, R)
    eact, _ = hg[0].masks()
    assert list(hg[0].elements) == r["elements"]
    assert r["active"] == int(eact.sum())
    assert r["hash"] == _fnv(eact)
    hg.close()
