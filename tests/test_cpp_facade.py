"""The C++ facade (include/robtop_amd.hpp) driven by a C++ SIMP program
(tests/cpp/cantilever_simp.cpp) the way the reference's driver calls its host
API (example.cpp:38-67, test_utils.cpp:589-660).

CPU: the program is built and starts (usage line, no GPU call).
GPU: the 64 x 32 x 32 cantilever (uniform start 0.4, filter r = 2, OC step 0.03,
damp 0.5, volume 0.4) runs 10 SIMP iterations; its supports and loads match the
cantilever definition, its final displacement solves the oracle's FP64
operator to the FP64 evaluation floor, compliance = f.u, and the OC update holds the volume.
"""
import os
import subprocess
import json

import numpy as np
import pytest

from problems import bits_of, rel_err

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "cantilever_simp")


def test_program_built_and_starts():
    assert os.path.exists(EXE), "run python -m robtop_amd.build"
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and p.stdout.startswith("usage:")


@pytest.mark.gpu
def test_cantilever_simp_through_cpp_facade(tmp_path):
    import oracle
    ex, ey, ez, iters = 64, 32, 32, 10
    print(f"[cpp] cantilever_simp {ex} {ey} {ez} {iters}", flush=True)
    p = subprocess.run([EXE, str(ex), str(ey), str(ez), str(iters), str(tmp_path)], capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stderr
    recs = [json.loads(l) for l in p.stdout.splitlines() if l.startswith("{")]
    assert 1 <= len(recs) <= iters
    for r in recs:
        assert np.isfinite(r["compliance"]) and r["compliance"] > 0 and r["rel_res"] <= 2e-6

    info = open(tmp_path / "info.txt").read().split()
    R, nx, ny, nz = int(info[0]), int(info[4]), int(info[5]), int(info[6])
    c_final = float(info[7])
    assert (R, nx, ny, nz) == (64, ex + 1, ey + 1, ez + 1)
    nn = nx * ny * nz
    rho = np.fromfile(tmp_path / "rho.f32", np.float32)
    u = np.fromfile(tmp_path / "u.f64").reshape(3, nn)
    f = np.fromfile(tmp_path / "f.f64").reshape(3, nn)
    fixed = np.fromfile(tmp_path / "fixed.u8", np.uint8).astype(bool)

    # boundary conditions: the x = 0 node plane fixed, -z load 1/N on the x = ex plane
    xs = np.broadcast_to(np.arange(nx), (nz, ny, nx)).reshape(-1)
    assert np.array_equal(fixed, xs == 0)
    fz = np.where(xs == ex, -1.0 / ((ey + 1) * (ez + 1)), 0.0)
    assert np.array_equal(f[2], fz) and not f[0].any() and not f[1].any()

    # the final displacement solves the oracle's FP64 operator
    a = np.zeros((R, R, R), bool)
    a[:ex, :ey, :ez] = True
    o = oracle.Hierarchy(R, bits_of(a))
    o.set_material(1e6, 0.4, 3.0)
    o.set_fixed(fixed.astype(np.uint8))
    o.set_rho(rho)
    o.set_rhop(np.fromfile(tmp_path / "rhop.f32", np.float32))   # the device's rho^p
    uz = u.copy()
    uz[:, fixed] = 0
    r = f - o.apply(uz)
    r[:, fixed] = 0
    # after 10 OC steps rho^p spans ~1e-9..1: FP64 evaluations of f - K u in
    # different summation orders differ by ~1e-6 relative at the solution, so
    # this bound is the evaluation floor, far below any error of supports,
    # loads or operator
    assert np.linalg.norm(r) / np.linalg.norm(f) < 5e-6
    assert rel_err(c_final, float(np.sum(f * u))) < 1e-12

    # OC held the volume at the goal over the solid elements (all of the box)
    assert abs(float(rho.astype(np.float64).mean()) - 0.4) < 2e-3
    assert rho.min() >= 1e-3 - 1e-9 and rho.max() <= 1.0


EXE2 = os.path.join(ROOT, "tests", "cpp", "facade_worst_mma")


def test_worst_mma_program_built_and_starts():
    assert os.path.exists(EXE2), "run python -m robtop_amd.build"
    p = subprocess.run([EXE2], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and p.stdout.startswith("usage:")


@pytest.mark.gpu
def test_mma_t_through_cpp_facade_matches_oracle():
    """MMA::mma_t(2, 4), init(-1, 1), update(df, dg rows, g) on the reference's
    2-D test problem (mma_t.cpp:208-300): every update's design equals the
    oracle's restatement, updated from the same designs."""
    import oracle
    print("[cpp] facade_worst_mma mma", flush=True)
    p = subprocess.run([EXE2, "mma"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    recs = [json.loads(l) for l in p.stdout.splitlines() if l.startswith("{")]
    assert recs and recs[-1]["converged"] == 1 and len(recs) < 200
    O = oracle.Mma(2, 4, -1, 1)
    for r in recs:
        x = O.x().astype(np.float64)
        a, b = x
        e = np.exp(10 * (a - b) ** 2)
        df = np.array([0.5 + 20 * (a - b) * e, np.sqrt(3) / 2 + 20 * (b - a) * e])
        g = np.array([a * a + b * b - 1, a * a - 0.75, b * b - 0.75, (a + b) ** 2 - 1])
        dg = np.array([[2 * a, 2 * b], [2 * a, 0], [0, 2 * b], [2 * (a + b), 2 * (a + b)]])
        O.update(df, dg, g)
        assert np.max(np.abs(O.x() - np.array(r["x"]))) < 1e-5, r


@pytest.mark.gpu
def test_modified_pm_and_void_solve_fem_through_cpp_facade():
    """double modifiedPM(void) and void solveFEM(void) from C++ on a 32 x 16 x
    16 box (supports x = 0, load area the top face): the worst compliance equals
    the same library calls made from Python with the same seed, and re-solving
    the worst load to 1e-4 keeps its compliance."""
    import robtop_amd as rt
    from robtop_amd.workloads import box_bits
    ex, ey, ez = 32, 16, 16
    print(f"[cpp] facade_worst_mma pm {ex} {ey} {ez}", flush=True)
    p = subprocess.run([EXE2, "pm", str(ex), str(ey), str(ez)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    r = json.loads([l for l in p.stdout.splitlines() if l.startswith("{")][-1])
    hg = rt.HierarchyGrid.genFromMesh(box_bits(32, ex, ey, ez), 32)
    hg.set_material(1e6, 0.4, 3.0)
    nx, ny, nz = hg[0].nodes
    X = np.broadcast_to(np.arange(nx), (nz, ny, nx)).reshape(-1)
    Z = np.broadcast_to(np.arange(nz)[:, None, None], (nz, ny, nx)).reshape(-1)
    surf = hg[0].surface_nodes()
    fixed = surf & (X == 0)
    load = surf & ~fixed & (Z == ez)
    f = np.zeros((3, nx * ny * nz))
    f[2, load] = -1.0
    hg.set_support(fixed)
    hg.set_force(f)
    assert hg.set_load_nodes(load.astype(np.uint8)) == r["load_nodes"]
    hg.init_rho(0.5)
    hg.update_stencil()
    hg.rand_force(1)
    c, steps = hg.modified_pm(500)
    hg.close()
    # setParameters takes floats (optimization.h:32-35), so the facade's
    # Poisson ratio is 0.4f, 1.5e-8 relative away from the 0.4 used here
    assert abs(c - r["worst_compliance"]) <= 1e-7 * abs(c)
    assert abs(r["compliance_after_solveFEM"] - c) <= 2e-2 * c
