"""The network as a coarse-grid correction (no reference counterpart), on
late-SIMP densities of the 128x64x64 cantilever.

The network is trained on the hierarchy's own triples up to SIMP step 12:
(density, load, displacement) after each step's solve, plus (density,
residual, error) pairs, the error being the exact solve K e = r of residuals
met by the step-12 solve, which are what a coarse-grid correction is asked to
invert. It is then used at step 13, whose densities and residuals it has not
seen, inside flexible CG (one-term orthogonalisation, Notay 2000, which
allows a preconditioner that varies from one application to the next):

* V: one V(2,2) cycle of the library (rt_vcycle) per application;
* V + net: the prolonged prediction from the restricted residual, scaled to
  minimise the energy of the error along it, then V on the residual left;
* two-grid + net: pre-smoothing, the prediction as the coarse solve (scaled
  the same way), post-smoothing.

Printed: the iterations of each against the library's own MG-PCG. What is
asserted is what holds for the trained network: both V-based variants
converge, and the network cannot replace the multigrid coarse solve (the
two-grid cycle with it needs more iterations than V; measured: V 93, V + net
125, two-grid + net stalled at 2.4e-2 after 600, the library's MG-PCG 105).
"""
import numpy as np
import pytest

import robtop_amd as rt
from robtop_amd.learned import WarmStartNet

pytestmark = pytest.mark.gpu

SIZE = (128, 64, 64)


def _fcg(apply_k, precond, f, tol, maxit, record=None):
    """Flexible CG with one-term orthogonalisation of the directions."""
    x = np.zeros_like(f)
    r = f.copy()
    nf = np.linalg.norm(f)
    z = precond(r)
    p = z.copy()
    for it in range(1, maxit + 1):
        q = apply_k(p)
        pq = float(np.vdot(p, q))
        a = float(np.vdot(z, r)) / pq
        x += a * p
        r -= a * q
        rel = np.linalg.norm(r) / nf
        if record is not None and it in record:
            record[it] = r.copy()
        if rel < tol:
            return x, it, rel
        z = precond(r)
        b = -float(np.vdot(z, q)) / pq
        p = z + b * p
    return x, maxit, rel


def _energy_scaled(c, r, apply_k):
    kc = apply_k(c)
    ckc = float(np.vdot(c, kc))
    return c * (float(np.vdot(c, r)) / ckc) if ckc > 0 else np.zeros_like(c)


def test_learned_coarse_correction_late_simp():
    from bench import build_problem
    hg, bits, R, fixed, f, rho = build_problem(*SIZE, 0)
    hg.set_precision(64)
    g0, g1 = hg[0], hg[1]
    net = WarmStartNet(3)

    def simp_step():
        hg.compute_sensitivity()
        hg.filter_sensitivity(2)
        hg.update_densities(0.4, 0.03, 0.5, 1e-3)
        hg.update_stencil()

    # SIMP steps 0..12, a (density, load, displacement) sample after each solve
    hg.update_stencil()
    for step in range(13):
        if step:
            simp_step()
        hg.set_force(f)
        st = hg.solve(tol=1e-6, maxit=3000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)
        assert st.rel_res < 1.5e-6
        net.add_sample(hg)
    print(f"[learned-cc] SIMP step 12: {st.iterations} PCG iterations", flush=True)
    u12 = hg.get_displacement()

    def apply_k(v):
        return g0.applyK(v)

    zero = np.zeros_like(f)

    def vcycle(r):
        # rt_vcycle runs on the hierarchy's force and displacement
        hg.set_force(r)
        hg.set_displacement(zero)
        hg.v_cycle(2, 2)
        return hg.get_displacement()

    # (density, residual, error) samples from the step-12 solve's residuals
    rec = {k: None for k in (4, 8, 12, 16, 20, 24, 28, 32)}
    _fcg(apply_k, vcycle, f, 0.0, max(rec), record=rec)
    for r in rec.values():
        hg.set_force(r)
        hg.set_displacement(np.zeros_like(r))
        st = hg.solve(tol=1e-8, maxit=3000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)
        assert st.rel_res < 1.5e-8
        net.add_sample(hg)
    hg.set_force(f)
    hg.set_displacement(u12)
    assert net.num_samples() == 13 + len(rec)
    l0 = net.train(1, batch=4, lr=0.0, seed=0)
    l1 = net.train(400, batch=4, lr=2e-3, seed=1)
    print(f"[learned-cc] {net.num_samples()} samples, training loss {l0:.4f} -> {l1:.4f}", flush=True)
    assert l1 < 0.5 * l0

    # step 13: unseen densities
    simp_step()
    hg.set_force(f)
    hg.set_displacement(np.zeros_like(f))
    plain = hg.solve(tol=1e-6, maxit=3000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)

    sup = fixed.astype(bool)

    def net_correction(r):
        g0.setResidual(r)
        g1.restrict_residual()
        net.predict_coarse(hg)
        g0.reset_displacement()
        g0.prolongate_correction()
        c = g0.getDisplacement()
        c[:, sup] = 0.0          # supports carry no displacement
        return c

    calls = [0]

    def v_net(r):
        # the network's correction, energy-scaled, then the cycle on what is
        # left of the residual
        c = _energy_scaled(net_correction(r), r, apply_k)
        r1 = r - apply_k(c)
        if calls[0] < 3:
            print(f"[learned-cc] V + net application {calls[0]}: |r - K c| / |r| = "
                  f"{np.linalg.norm(r1) / np.linalg.norm(r):.3e}", flush=True)
        calls[0] += 1
        return c + vcycle(r1)

    def two_grid_net(r):
        g0.setForce(r)
        g0.reset_displacement()
        g0.gs_relax(2)
        g0.update_residual()
        r0 = g0.getResidual()
        u_pre = g0.getDisplacement()
        c = _energy_scaled(net_correction(r0), r0, apply_k)
        g0.setDisplacement(u_pre + c)
        g0.gs_relax(2, reverse=True)
        return g0.getDisplacement()

    maxit = 600
    res = {}
    for name, m in (("V", vcycle), ("V + net", v_net), ("two-grid + net", two_grid_net)):
        x, it, rel = _fcg(apply_k, m, f, 1e-6, maxit)
        res[name] = (it, rel)
        print(f"[learned-cc] step 13, flexible CG with {name}: {it} iterations, rel {rel:.2e}", flush=True)
    print(f"[learned-cc] step 13, library MG-PCG: {plain.iterations} iterations", flush=True)
    assert plain.rel_res < 1.5e-6
    assert res["V"][1] < 1e-6 and res["V + net"][1] < 1e-6
    assert res["two-grid + net"][0] > res["V"][0]
    hg.close()
    net.close()
