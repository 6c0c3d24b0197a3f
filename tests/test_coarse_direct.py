"""The preconditioner's early exact solve (rt_set_coarse_direct_dofs; no
reference counterpart): the PCG preconditioner stops one level above the
coarsest and solves it with a dense FP64 inverse when that level is small.

* the library takes the early stop exactly when the level above the coarsest
  has at most the set number of unknowns;
* the reference-semantics V-cycle (rt_vcycle) is unaffected: the same bits
  with and without the early stop;
* MG-PCG with and without it reaches the tolerance, with solutions that agree
  to the tolerance and no more iterations (an exact solve of a level is at
  least as good a coarse correction as a cycle over it);
* batched and single solves stay bit-identical with it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cantilever(size=(64, 32, 32), seed=0):
    from bench import build_problem
    hg, bits, R, fixed, f, rho = build_problem(*size, seed)
    hg.update_stencil()
    return hg, f, fixed


def test_early_exact_solve_engages_by_level_size():
    import robtop_amd as rt
    hg, f, _ = _cantilever()
    n = len(hg)
    _, nodes = hg.num_active(n - 2)
    print(f"[direct] {n} levels, level {n - 2} has {3 * nodes} unknowns", flush=True)
    assert n >= 3
    hg.set_coarse_direct_dofs(3 * nodes)
    assert hg.preconditioner_levels() == n - 1
    hg.set_coarse_direct_dofs(3 * nodes - 1)
    assert hg.preconditioner_levels() == n
    hg.set_coarse_direct_dofs(0)
    assert hg.preconditioner_levels() == n
    for bad in (-1, 30001):
        with pytest.raises(rt.RobtopError):
            hg.set_coarse_direct_dofs(bad)
    assert hg.preconditioner_levels() == n
    hg.close()


def test_vcycle_unchanged_and_pcg_agrees():
    import robtop_amd as rt
    hg, f, _ = _cantilever(seed=2)
    n = len(hg)
    _, nodes = hg.num_active(n - 2)
    out = {}
    for dofs in (0, 3 * nodes):
        hg.set_coarse_direct_dofs(dofs)
        hg.set_displacement(np.zeros_like(f))
        rels = [hg.v_cycle(2, 2) for _ in range(3)]
        u_v = hg.get_displacement()
        hg.set_displacement(np.zeros_like(f))
        st = hg.solve(tol=1e-8, maxit=1000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)
        out[dofs] = (rels, u_v, st.iterations, st.rel_res, hg.get_displacement())
        print(f"[direct] max dofs {dofs}: preconditioner levels {hg.preconditioner_levels()}, "
              f"PCG {st.iterations} iterations, rel_res {st.rel_res:.2e}", flush=True)
    (r0, v0, i0, q0, u0), (r1, v1, i1, q1, u1) = out[0], out[3 * nodes]
    assert r0 == r1 and np.array_equal(v0, v1)
    assert q0 < 1.5e-8 and q1 < 1.5e-8
    assert i1 <= i0
    assert np.linalg.norm(u1 - u0) / np.linalg.norm(u0) < 1e-6
    hg.close()


def test_batch_bit_identical_with_early_solve():
    hg, f, fixed = _cantilever(seed=4)
    n = len(hg)
    _, nodes = hg.num_active(n - 2)
    hg.set_coarse_direct_dofs(3 * nodes)
    assert hg.preconditioner_levels() == n - 1
    rng = np.random.default_rng(1)
    _, vact = hg[0].masks()
    F = np.stack([f] + [rng.standard_normal(f.shape) * (vact & ~fixed.astype(bool))[None] for _ in range(3)])
    U4, it4, rel4 = hg.solve_batch(F, tol=1e-8, maxit=1000)
    assert np.all(rel4 < 1.5e-8)
    for b in (0, 2):
        U1, it1, _ = hg.solve_batch(F[b:b + 1], tol=1e-8, maxit=1000)
        assert it1[0] == it4[b] and np.array_equal(U1[0], U4[b])
    hg.close()


def test_level_sweeps_override_the_multiplier():
    """rt_set_level_sweeps: overriding every coarse level with multiplier m is
    the same preconditioner, bit for bit, as rt_set_preconditioner(m); a
    single override changes only its level (different iterates, same
    tolerance reached)."""
    import robtop_amd as rt
    hg, f, _ = _cantilever(seed=5)
    hg.set_coarse_direct_dofs(0)     # every coarse level smoothed
    n = len(hg)
    assert hg.preconditioner_levels() == n

    def solve():
        hg.set_displacement(np.zeros_like(f))
        st = hg.solve(tol=1e-8, maxit=1000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)
        return st.iterations, st.rel_res, hg.get_displacement()

    hg.set_preconditioner(2, 1)
    i2, _, u2 = solve()
    hg.set_preconditioner(1, 1)
    for lv in range(1, n):
        hg.set_level_sweeps(lv, 2)
    io, _, uo = solve()
    assert io == i2 and np.array_equal(uo, u2)
    for lv in range(1, n):
        hg.set_level_sweeps(lv, 0)
    i1, _, u1 = solve()
    hg.set_level_sweeps(1, 2)
    im, qm, um = solve()
    print(f"[sweeps] x1 {i1}, x2 {i2}, level 1 only x2 {im} iterations", flush=True)
    assert qm < 1.5e-8
    assert not np.array_equal(um, u1)
    with pytest.raises(Exception):
        hg.set_level_sweeps(0, 2)
    hg.close()
