"""Learned warm start: the network's GPU kernels against the oracle
(oracle/net.py, pinned on the CPU by tests/test_net_oracle.py), training on
solved hierarchies, and what the trained network gains.

Forward and gradients: the GPU network (bfloat16 weights and activations,
FP32 accumulation on the matrix cores) against the oracle with the same
bfloat16 roundings (emulate=True); the remaining difference is FP32
accumulation order and the rare bfloat16 tie it flips: relative L2 error
below 1e-2 for outputs and 3e-2 per layer for gradients.
"""
import os
import sys

import numpy as np
import pytest

import robtop_amd as rt
from robtop_amd.learned import WarmStartNet
from oracle import net as on

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


# (32, 32, 64) and (21, 37, 70) are large enough for the LDS-staged
# convolution at the full resolution (the second with ragged 16 x 16 tiles) and
# the four-row-tile kernel at the half resolution; the others exercise the
# one-row kernels and ragged rows
@pytest.mark.parametrize("shape", [(7, 9, 21), (2, 3, 4), (5, 8, 33), (32, 32, 64), (21, 37, 70)])
def test_gpu_net_forward_matches_oracle(shape):
    print(f"[learned] forward {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    theta = on.random_params(11)
    net = WarmStartNet(0)
    net.set_params(theta.astype(np.float32))
    x = rng.standard_normal(shape + (4,)).astype(np.float32)
    y = net.forward(x)
    yo, _ = on.forward(theta.astype(np.float32).astype(np.float64), x.astype(np.float64), emulate=True)
    assert rel(y, yo) < 1e-2, rel(y, yo)
    net.close()


@pytest.mark.parametrize("shape", [(6, 10, 19), (3, 2, 5), (32, 32, 64)])
def test_gpu_net_gradient_matches_oracle(shape):
    print(f"[learned] gradient {shape}", flush=True)
    rng = np.random.default_rng(7 + sum(shape))
    theta = on.random_params(12).astype(np.float32)
    net = WarmStartNet(0)
    net.set_params(theta)
    x = rng.standard_normal(shape + (4,)).astype(np.float32)
    t = rng.standard_normal(shape + (3,)).astype(np.float32)
    loss, g = net.gradient(x, t)
    lo, go = on.loss_and_grad(theta.astype(np.float64), x.astype(np.float64), t.astype(np.float64), emulate=True)
    assert abs(loss - lo) / lo < 1e-3, (loss, lo)
    pg, po = on.unpack(g), on.unpack(go)
    for name in on.LAYERS:
        for k in (0, 1):
            e = rel(pg[name][k], po[name][k])
            assert e < 3e-2, (name, k, e)
    # padding entries stay without gradient
    assert np.all(pg["E0A"][0][:, :, 4:] == 0) and np.all(pg["HEAD"][0][:, 3:, :] == 0)
    net.close()


SIZE = (128, 64, 64)


def _problem(seed):
    from bench import build_problem
    hg, bits, R, fixed, f, rho = build_problem(*SIZE, seed)
    hg.update_stencil()
    return hg, f, rho


def _solve_from_zero(hg, f):
    hg.set_displacement(np.zeros_like(f))
    return hg.solve(tol=1e-8, maxit=2000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)


@pytest.fixture(scope="module")
def trained():
    """A network trained on 16 solved cantilevers with seeded density fields."""
    net = WarmStartNet(1)
    hg, f, _ = _problem(100)
    from robtop_amd.workloads import smoothed_random_density
    g0 = hg[0]
    eact, _ = g0.masks()
    for seed in range(1, 17):
        hg.set_density(smoothed_random_density(eact, tuple(g0.elements), seed))
        hg.update_stencil()
        st = _solve_from_zero(hg, f)
        assert st.rel_res < 1.5e-8
        net.add_sample(hg)
    assert net.num_samples() == 16
    l0 = net.train(1, batch=4, lr=0.0, seed=0)
    l1 = net.train(400, batch=4, lr=2e-3, seed=1)
    print(f"[learned] training loss {l0:.4f} -> {l1:.4f}", flush=True)
    yield net, hg, f, l0, l1
    hg.set_learned_warm_start(None)
    hg.close()
    net.close()


def test_gpu_training_reduces_loss(trained):
    net, hg, f, l0, l1 = trained
    assert l1 < 0.5 * l0, (l0, l1)


def _coarse_part(u, nodes):
    """u (3, nodes box) sampled at every 4th node and trilinearly interpolated
    back: the best any prediction on the x4 level's nodes can give."""
    nx, ny, nz = nodes
    v = u.reshape(3, nz, ny, nx)
    out = np.empty_like(v)
    for c in range(3):
        a = v[c][::4, ::4, ::4]
        for axis, n in ((2, nx), (1, ny), (0, nz)):
            xc = np.arange(a.shape[axis]) * 4.0
            a = np.apply_along_axis(lambda r: np.interp(np.arange(n), xc, r), axis, a)
        out[c] = a
    return out.reshape(3, -1)


def test_gpu_learned_warm_start_gain(trained):
    """What the trained network gains, on a density field not in the training
    set: its prediction of the unit-RMS displacement on the x4 nodes has a mean
    squared error well below the untrained network's, and the solve from the
    learned start reaches the same solution in at most as many PCG iterations
    as from a zero displacement (plus 2 for rounding). Printed beside it: the
    iterations from the best start any x4-level prediction could give (the
    solution sampled on the x4 nodes and interpolated). With the stop on the
    residual norm they equal the zero start's: the interpolation error leaves a
    residual as large as |f|, so a displacement prediction on that level
    cannot save PCG iterations."""
    net, hg, f, _, _ = trained
    from robtop_amd.workloads import smoothed_random_density
    g0 = hg[0]
    eact, _ = g0.masks()
    hg.set_density(smoothed_random_density(eact, tuple(g0.elements), 999))
    hg.update_stencil()
    hg.set_learned_warm_start(None)
    plain = _solve_from_zero(hg, f)
    u_plain = hg.get_displacement()
    e_star = -0.5 * hg.compliance()         # E(u*) = u*.K u* / 2 - f.u* = -f.u* / 2
    raw = WarmStartNet(5)
    res = {}
    for name, n in (("untrained", raw), ("trained", net)):
        hg.set_displacement(u_plain)
        loss = n.evaluate(hg)
        hg.set_learned_warm_start(n)
        st = _solve_from_zero(hg, f)
        info = hg.learned_warm_start_info()
        hg.set_learned_warm_start(None)
        res[name] = dict(loss=loss, iterations=st.iterations, energy_fraction=info["e_learned"] / e_star,
                         taken=info["taken"])
        assert st.rel_res < 1.5e-8
        assert rel(hg.get_displacement(), u_plain) < 1e-5
    raw.close()
    hg.set_displacement(_coarse_part(u_plain, tuple(e + 1 for e in g0.elements)))
    best = hg.solve(tol=1e-8, maxit=2000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)
    print(f"[learned] zero start: {plain.iterations} PCG iterations; {res}; "
          f"exact x4-level start: {best.iterations} iterations", flush=True)
    assert res["trained"]["loss"] < 0.5 * res["untrained"]["loss"]
    assert res["trained"]["iterations"] <= plain.iterations + 2


def test_gpu_untrained_network_falls_back_or_helps(trained):
    """An untrained network cannot make the start worse: the start keeps the
    lower energy of the two, and the solve reaches the same tolerance."""
    _, hg, f, _, _ = trained
    raw = WarmStartNet(5)
    u0 = hg.get_displacement()
    hg.set_learned_warm_start(raw)
    st = hg.solve(tol=1e-8, maxit=2000, mode=rt.RT_SOLVE_RELIABLE, pre=2, post=2)
    info = hg.learned_warm_start_info()
    hg.set_learned_warm_start(None)
    raw.close()
    print(f"[learned] untrained network from a converged start: {info}, iterations {st.iterations}", flush=True)
    assert info["e_learned"] <= info["e_plain"] or not info["taken"]
    assert st.rel_res < 1.5e-8
    assert rel(hg.get_displacement(), u0) < 1e-5


def test_gpu_adam_matches_oracle():
    """k_adam against the oracle's float32 Adam (oracle/net.py adam_step,
    pinned by tests/test_net_oracle.py) over four steps from a set state:
    random gradients over seven decades, exact zeros, a repeated gradient."""
    net = WarmStartNet(1)
    n = net.num_params
    rng = np.random.default_rng(5)
    w = rng.standard_normal(n).astype(np.float32) * 0.1
    net.set_params(w)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    g1 = (rng.standard_normal(n) * 10.0 ** rng.uniform(-4, 3, n)).astype(np.float32)
    g1[::7] = 0
    grads = [g1, (rng.standard_normal(n) * 0.01).astype(np.float32), g1, np.zeros(n, np.float32)]
    lrs = [1e-3, 5e-4, 1e-3, 1e-3]
    for t, (g, lr) in enumerate(zip(grads, lrs), start=1):
        print(f"[adam] step {t}", flush=True)
        net.adam_step(g, lr)
        w, m, v = on.adam_step(w, m, v, g, t, lr)
        wg = net.params()
        # the GPU's sqrtf and division differ from numpy's by an ulp or two
        assert np.max(np.abs(wg - w)) <= 1e-5 * lr + 2e-7 * np.max(np.abs(w)), t
    net.close()
