"""Mixed-precision solves against an FP64 solution (the smoothed-random-density
cantilever of the benchmark at 128 x 64 x 64 elements).

The FP64 MG-PCG (FP64 V-cycle and vectors) solved to a true relative residual
below 1e-10 is the reference; its kernels are pinned against the oracle operator by
operator in test_gpu_parity.py. The FP32-preconditioned solvers (reliable
updates, iterative refinement), with FP32 or bfloat16 coarse stencils, must
reach a true FP64 residual of 1e-8 and agree with it in displacement and
compliance. The preconditioner's precision changes the iteration count only,
not what the solve converges to.
"""
import numpy as np
import pytest

import robtop_amd as rt
from problems import rel_err

pytestmark = pytest.mark.gpu

SIZE = (128, 64, 64)


@pytest.fixture(scope="module")
def cantilever():
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from bench import build_problem
    print(f"[mixed] cantilever {SIZE}", flush=True)
    hg, bits, R, fixed, f, rho = build_problem(*SIZE, 0)
    hg.update_stencil()
    hg.set_displacement(np.zeros_like(f))
    st = hg.solve(tol=1e-12, maxit=3000, mode=rt.RT_SOLVE_FP64, pre=2, post=2)
    assert st.rel_res < 1e-10, st.rel_res
    u_ref = hg.get_displacement()
    c_ref = hg.compliance()
    yield hg, f, u_ref, c_ref
    hg.close()


@pytest.mark.parametrize("bits", [32, 16])
@pytest.mark.parametrize("mode", ["reliable", "fp32_ir"])
def test_fp32_preconditioned_solve_matches_fp64(cantilever, bits, mode):
    hg, f, u_ref, c_ref = cantilever
    print(f"[mixed] coarse stencil bits={bits} mode={mode}", flush=True)
    m = rt.RT_SOLVE_RELIABLE if mode == "reliable" else rt.RT_SOLVE_FP32_IR
    hg.set_coarse_stencil_bits(bits)
    try:
        hg.set_displacement(np.zeros_like(f))
        st = hg.solve(tol=1e-8, maxit=2000, mode=m, pre=2, post=2)
    finally:
        hg.set_coarse_stencil_bits(32)
    assert st.rel_res < 1.5e-8, st.rel_res
    assert rel_err(hg.get_displacement(), u_ref) < 1e-5
    assert abs(hg.compliance() - c_ref) / c_ref < 1e-7


def test_bf16_coarse_stencils_keep_iteration_count(cantilever):
    hg, f, u_ref, c_ref = cantilever
    its = {}
    for bits in (32, 16):
        hg.set_coarse_stencil_bits(bits)
        hg.set_displacement(np.zeros_like(f))
        its[bits] = hg.solve(tol=1e-8, maxit=2000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2).iterations
    hg.set_coarse_stencil_bits(32)
    print(f"[mixed] PCG iterations to 1e-8: {its}", flush=True)
    assert its[16] <= 1.25 * its[32] + 2, its


def test_subspace_warm_start(cantilever):
    """A sequence of solves under slowly changing densities: starting each from
    the energy-optimal combination of the last three solutions reaches the same
    tolerance and the same solution as starting from the previous one."""
    hg, f, u_ref, c_ref = cantilever
    rho0 = hg.get_density()
    its, sols = {}, {}
    for m in (0, 3):
        hg.set_warm_start(m)
        hg.set_displacement(np.zeros_like(f))
        its[m] = []
        for step in range(4):
            hg.set_density(np.clip(rho0 * (1 + 0.03 * step), 1e-3, 1.0).astype(np.float32))
            hg.update_stencil()
            st = hg.solve(tol=1e-8, maxit=2000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)
            assert st.rel_res < 1.5e-8, (m, step, st.rel_res)
            its[m].append(st.iterations)
        sols[m] = hg.get_displacement()
    hg.set_warm_start(0)
    hg.set_density(rho0)
    hg.update_stencil()
    print(f"[mixed] PCG iterations per step, warm start 0 / 3: {its}", flush=True)
    assert rel_err(sols[3], sols[0]) < 1e-5



def test_cg_ritz_values(cantilever):
    """The Ritz values the library reports (Sturm bisection on the Lanczos
    tridiagonal of the CG coefficients) equal numpy's eigenvalues of the same
    tridiagonal, and lie in (0, ~1]: a symmetric V(2,2) cycle is a
    contraction-based preconditioner whose preconditioned operator has its
    spectrum near (0, 1]."""
    hg, f, u_ref, c_ref = cantilever
    hg.set_displacement(np.zeros_like(f))
    st = hg.solve(tol=1e-8, maxit=2000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)
    a, b = hg.cg_coefficients()
    n = len(a)
    assert n == st.iterations
    m = n if n == len(b) and b[-1] != 0 else n   # beta_j forms p_{j+1}; the last may be absent
    d = np.array([1 / a[j] + (b[j - 1] / a[j - 1] if j else 0.0) for j in range(m)])
    e = np.array([np.sqrt(b[j]) / a[j] for j in range(m - 1)])
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    ev = np.linalg.eigvalsh(T)
    lo, hi = hg.cg_ritz_values(5)
    print(f"[ritz] {n} iterations, smallest {lo}, largest {hi}", flush=True)
    assert np.allclose(lo, ev[:5], rtol=1e-9, atol=1e-12)
    assert np.allclose(hi, ev[::-1][:5], rtol=1e-9, atol=1e-12)
    assert 0 < lo[0] and hi[0] < 1.5
