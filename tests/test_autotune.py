"""Autotuner (rt_autotune): it picks per-level launch shapes by timing; the
solves after it must give the same solution (the finest-level alternatives
are bit-identical, the coarse ones differ only in summation order, which the
FP64 solve's tolerance absorbs) and the same PCG iteration count to within a
couple of iterations."""
import numpy as np
import pytest

import robtop_amd as rt
from problems import rel_err

pytestmark = pytest.mark.gpu


def test_autotune_keeps_solution():
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from bench import build_problem
    print("[autotune] cantilever 128x64x64", flush=True)
    hg, bits, R, fixed, f, rho = build_problem(128, 64, 64, 3)
    hg.update_stencil()
    res = {}
    for tuned in (False, True):
        if tuned:
            choice = hg.autotune(reps=3)
            print(f"[autotune] {choice}", flush=True)
            assert choice["fine_zc"] in (8, 12, 16, 24, 32, 48)
            assert choice["fine_smoother"] in (0, 1)
        out = {}
        for name, mode, tol in (("fp64", rt.RT_SOLVE_FP64, 1e-11), ("reliable", rt.RT_SOLVE_RELIABLE, 1e-8)):
            hg.set_displacement(np.zeros_like(f))
            st = hg.solve(tol=tol, maxit=2000, mode=mode, pre=2, post=2)
            out[name] = (st.iterations, st.rel_res, hg.get_displacement())
        res[tuned] = out
    for name in ("fp64", "reliable"):
        (i0, r0, u0), (i1, r1, u1) = res[False][name], res[True][name]
        print(f"[autotune] {name}: iterations {i0} -> {i1}, residual {r0:.2e} -> {r1:.2e}", flush=True)
        assert abs(i1 - i0) <= 2
        assert r1 < 3 * max(r0, 1e-12) + 1e-11
    assert rel_err(res[True]["fp64"][2], res[False]["fp64"][2]) < 1e-9
    assert rel_err(res[True]["reliable"][2], res[False]["fp64"][2]) < 1e-6
    hg.close()
