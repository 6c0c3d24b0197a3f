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
