"""Batched solves (rt_solve_batch, and block Lanczos's block solves; no
reference counterpart): nb load cases pushed through one batched V-cycle.

* no cross-talk: a load case solved alone and inside a batch of eight gives
  bit for bit the same displacement and iteration count (the batched kernels
  apply each vector's arithmetic in the single kernels' order);
* the batched solve agrees with the one-at-a-time MG-PCG (rt_solve, FP64
  Krylov vectors) to the solve tolerance, in about the same iterations;
* block Lanczos with batched block solves (FP64 Krylov vectors, or FP32 ones
  with reliable updates) finds the same worst compliance as with one solve per
  right-hand side.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cantilever(seed=0, size=(64, 32, 32)):
    from bench import build_problem
    hg, bits, R, fixed, f, rho = build_problem(*size, seed)
    hg.update_stencil()
    return hg, f, fixed


def _loads(hg, f, fixed, nb, seed):
    rng = np.random.default_rng(seed)
    _, vact = hg[0].masks()
    out = [f]
    for _ in range(nb - 1):
        g = rng.standard_normal(f.shape) * (vact & ~fixed.astype(bool))[None]
        out.append(g / np.linalg.norm(g))
    return np.stack(out)


def test_batch_has_no_cross_talk():
    hg, f, fixed = _cantilever()
    F = _loads(hg, f, fixed, 8, 1)
    print("[batch] 8 load cases vs each alone, 64x32x32", flush=True)
    U8, it8, rel8 = hg.solve_batch(F, tol=1e-8, maxit=1000)
    assert np.all(rel8 < 1.5e-8), rel8
    for b in (0, 3, 7):
        U1, it1, rel1 = hg.solve_batch(F[b:b + 1], tol=1e-8, maxit=1000)
        assert it1[0] == it8[b]
        assert np.array_equal(U1[0], U8[b]), b
    hg.close()


def test_batch_matches_single_solves():
    import robtop_amd as rt
    hg, f, fixed = _cantilever(seed=3)
    F = _loads(hg, f, fixed, 5, 2)
    print("[batch] 5 load cases vs rt_solve (mixed), 64x32x32", flush=True)
    U, it, rel = hg.solve_batch(F, tol=1e-8, maxit=1000)
    for b in range(len(F)):
        hg.set_force(F[b])
        hg.set_displacement(np.zeros_like(F[b]))
        st = hg.solve(tol=1e-8, maxit=1000, mode=rt.RT_SOLVE_MIXED, pre=2, post=2)
        u = hg.get_displacement()
        err = np.linalg.norm(U[b] - u) / np.linalg.norm(u)
        print(f"[batch] case {b}: batched {it[b]} iterations, single {st.iterations}, difference {err:.2e}", flush=True)
        assert rel[b] < 1.5e-8 and st.rel_res < 1.5e-8
        assert err < 1e-5
        assert abs(int(it[b]) - st.iterations) <= 3
    hg.close()


@pytest.mark.parametrize("mode", ["mixed", "reliable"])
def test_block_lanczos_batched_equals_sequential(mode):
    import robtop_amd as rt
    from robtop_amd.workloads import torus_bits, torus_worst_case_bc
    n = 64
    print(f"[batch] block Lanczos on the torus n={n}, batched and sequential block solves", flush=True)
    bits, R, solid = torus_bits(n)
    g = rt.HierarchyGrid.genFromMesh(bits, R)
    g.set_material(1e6, 0.4, 3.0)
    fixed, load, normals = torus_worst_case_bc(g, n)
    g.set_support(fixed)
    g.set_load_nodes(load.astype(np.uint8), normals)
    eact, _ = g[0].masks()
    rng = np.random.default_rng(0)
    g.set_density(np.where(eact, rng.uniform(0.2, 1.0, eact.size), 0).astype(np.float32))
    g.update_stencil()
    res = {}
    for batched in (True, False):
        g.set_block_solves(batched)
        m = rt.RT_SOLVE_MIXED if mode == "mixed" else rt.RT_SOLVE_RELIABLE
        c, st = g.block_lanczos(k=8, max_blocks=30, tol=1e-8, solve_tol=1e-9, seed=5, mode=m)
        res[batched] = (c, st)
        print(f"[batch] batched={batched}: worst compliance {c:.12e}, {st}", flush=True)
    (cb, sb), (cs, ss) = res[True], res[False]
    # mixed: the same arithmetic per vector; reliable: the same up to the
    # summation order of |f| (which can move a residual replacement)
    assert abs(cb - cs) <= (1e-12 if mode == "mixed" else 1e-9) * abs(cs)
    assert sb["blocks"] == ss["blocks"] and sb["solves"] == ss["solves"]
    if mode == "mixed":
        assert sb["pcg_iterations"] == ss["pcg_iterations"]
    g.set_block_solves(True)
    g.close()
