"""The worst-case and learned-start workloads at their full sizes (the kernels
are pinned against the oracle at small sizes; here the checks are the
size-independent ones).

* Torus worst case at 384^3 (5,634,640 elements): two worst-case SIMP
  iterations of optimization() (optimization.cpp:583-673) with the worst load
  found by block Lanczos, k = 8 right-hand sides per block and FP32-
  preconditioned solves; then, at the final densities, the worst compliance
  from FP32-preconditioned solves against the one from FP64 solves to 1e-11
  (bound 1e-5 relative).
* Learned warm start at 511 x 256 x 256 (101.45M DOF): a network trained on
  triples of a 255 x 128 x 128 SIMP run is deployed for three SIMP steps; every
  solve reaches 1e-6 whether the learned start is taken or not, the solution
  equals the plain solve's, and an untrained network falls back.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_torus_384_worst_case_block_lanczos():
    import robtop_amd as rt
    from robtop_amd.workloads import torus_bits, torus_worst_case_bc
    n = 384
    print(f"[config 3] torus {n}^3", flush=True)
    bits, R, solid = torus_bits(n)
    hg = rt.HierarchyGrid.genFromMesh(bits, R)
    assert hg.num_active(0) == (5634640, 5812456)
    hg.set_material(1e6, 0.4, 3.0)
    fixed, load, normals = torus_worst_case_bc(hg, n)
    hg.set_support(fixed)
    nl = hg.set_load_nodes(load.astype(np.uint8), normals)
    assert nl > 100000
    hg.set_precision(32)
    params = rt.Parameter(volume_ratio=0.3, volume_decrease=0.05)
    lz = dict(k=8, max_blocks=20, tol=1e-6, solve_tol=1e-6, pre=2, post=2)
    recs = rt.worst_case_topopt(hg, params, n_iter=2, stop=False, method="lanczos", lanczos=lz,
                                log=lambda r: print(f"[config 3] iter {r['iter']}: c {r['worst_compliance']:.6e}, "
                                                    f"{r['blocks']} blocks, {r['solves']} solves, "
                                                    f"{r['seconds']:.2f} s", flush=True))
    assert len(recs) == 2
    for r in recs:
        assert np.isfinite(r["worst_compliance"]) and r["worst_compliance"] > 0
        assert r["blocks"] <= 20
    hg.update_stencil()
    c32, _ = hg.block_lanczos(seed=99, **lz)
    c64, st64 = hg.block_lanczos(seed=99, **dict(lz, tol=1e-8, solve_tol=1e-11, mode=rt.RT_SOLVE_FP64))
    print(f"[config 3] worst compliance FP32-preconditioned {c32:.12e}, FP64 {c64:.12e}, "
          f"rel {abs(c32 - c64) / c64:.2e}", flush=True)
    assert abs(c32 - c64) <= 1e-5 * abs(c64)
    hg.close()


def test_learned_start_deployed_at_511x256x256():
    import robtop_amd as rt
    from robtop_amd.learned import WarmStartNet, cantilever_hierarchy, collect_simp_samples
    net = WarmStartNet(3)
    seeds = collect_simp_samples(net, size=(255, 128, 128), triples=12, iters_per_run=12,
                                 log=lambda s: print(f"[config 4] {s}", flush=True))
    assert net.num_samples() == 12 and len(seeds) == 1
    l0 = net.train(1, batch=4, lr=0.0, seed=0)
    l1 = net.train(200, batch=4, lr=2e-3, seed=1)
    print(f"[config 4] training loss {l0:.4f} -> {l1:.4f}", flush=True)
    assert l1 < l0
    hg, f = cantilever_hierarchy(511, 256, 256, 0)
    plain, _ = cantilever_hierarchy(511, 256, 256, 0)
    hg.set_learned_warm_start(net)
    taken = 0
    for step in range(3):
        for h in (hg, plain):
            h.update_stencil()
        st = hg.solve(tol=1e-6, maxit=5000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)
        sp = plain.solve(tol=1e-6, maxit=5000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)
        info = hg.learned_warm_start_info()
        taken += int(info["taken"])
        print(f"[config 4] step {step}: learned start taken {info['taken']}, {st.iterations} PCG iterations "
              f"(plain {sp.iterations}), rel_res {st.rel_res:.2e}", flush=True)
        assert st.converged and sp.converged and st.rel_res <= 1e-6
        u, up = hg.get_displacement(), plain.get_displacement()
        du = np.linalg.norm(u - up) / np.linalg.norm(up)
        print(f"[config 4] step {step}: displacement difference {du:.2e}", flush=True)
        assert du <= 1e-3
        c, cp = hg.compliance(), plain.compliance()
        assert abs(c - cp) <= 1e-5 * abs(cp)
        # the two loops stay on identical densities: the same sensitivities from the same solutions
        for h in (hg, plain):
            h.set_displacement(up)
            h.compute_sensitivity()
            h.filter_sensitivity(2)
            h.update_densities(0.4, 0.03, 0.5, 1e-3)
    print(f"[config 4] learned start taken in {taken} of 3 steps", flush=True)
    raw = WarmStartNet(11)
    hg.set_learned_warm_start(raw)
    hg.update_stencil()
    st = hg.solve(tol=1e-6, maxit=5000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)
    info = hg.learned_warm_start_info()
    print(f"[config 4] untrained network: {info}", flush=True)
    assert st.converged and (not info["taken"] or info["e_learned"] <= info["e_plain"])
    hg.set_learned_warm_start(None)
    for x in (raw, net, hg, plain):
        x.close()
