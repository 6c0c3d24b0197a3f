"""Pins the oracle (oracle/oracle.cpp) against independent brute force,
closed forms and the reference's own known answers. CPU only."""
import numpy as np
import pytest

import bruteforce as bf
import oracle
from problems import active_cube, bits_of, problem_inputs, rel_err


# --- KE --------------------------------------------------------------------------
def test_ke_matches_exact_integration():
    for E, nu in [(1e6, 0.4), (1.0, 0.3), (210e3, 0.25)]:
        assert rel_err(oracle.ke(E, nu), bf.ke_exact(E, nu)) < 1e-14


def test_ke_invariants():
    K = oracle.ke(1.0, 0.4)
    assert np.allclose(K, K.T, atol=1e-15)
    w = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(w) < 1e-12) == 6 and w.min() > -1e-12      # 6 rigid modes, PSD, rank 18
    # rigid translations and rotations are in the kernel
    X = np.array([[i % 2, i // 2 % 2, i // 4] for i in range(8)], float)
    for a in range(3):
        t = np.zeros(24); t[a::3] = 1
        assert np.abs(K @ t).max() < 1e-13
    for a, b in [(0, 1), (1, 2), (0, 2)]:
        r = np.zeros((8, 3)); r[:, a] = -X[:, b]; r[:, b] = X[:, a]
        assert np.abs(K @ r.reshape(-1)).max() < 1e-13
    # linear in E, invariant to h (3D stiffness scales with h^(3-2))
    assert rel_err(oracle.ke(7.0, 0.4), 7.0 * K) < 1e-14
    assert rel_err(oracle.ke(1.0, 0.4, h=2.0), 2.0 * K) < 1e-14


# --- BitSAT known answer (optimization.cu:236-284) -------------------------------
def test_bitsat_selftest_known_answer():
    words = np.full(10000, 0xAAAAAAAA, dtype=np.uint32)
    for i in [0, 1, 2, 3, 31, 32, 33, 12345, 319999]:
        r = oracle.bitsat_rank(words, i)
        assert r == (i // 2 if i % 2 else -1)


# --- levels ---------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["box3", "holes3", "box4", "lshape4"])
def test_levels_match_bruteforce(kind):
    a = active_cube(kind, 64 if kind.endswith("4") else 32)
    H = oracle.Hierarchy(a.shape[0], bits_of(a))
    L = bf.Levels(a)
    assert H.nlayers == L.nlayers
    assert [H.layer_active(i) for i in range(H.nlayers)] == L.counts
    for lo, lb in zip(H.levels, L.levels):
        assert lo["layer"] == lb["layer"]
        assert np.array_equal(lo["o"], lb["o"]) and np.array_equal(lo["e"], lb["e"])
        assert np.array_equal(lo["eact"], bf.to_flat(lb["eact"]))
        assert np.array_equal(lo["vact"], bf.to_flat(lb["vact"]))


def test_config_level_counts():
    """Node counts of the reference's coarsening rule on the 64x32x32 cantilever box."""
    a = active_cube("cantilever64", 64)
    H = oracle.Hierarchy(64, bits_of(a))
    assert H.nlayers == 4
    nodes = [int(l["vact"].sum()) for l in H.levels]
    assert nodes == [70785, 1377, 225]


# --- operators vs brute force ---------------------------------------------------
def _setup(kind, seed):
    a = active_cube(kind, 64 if kind.endswith("4") else 32)
    H = oracle.Hierarchy(a.shape[0], bits_of(a))
    H.set_material(1e6, 0.4, 3.0)
    lv0 = H.levels[0]
    fixed, f, rho = problem_inputs(lv0, seed)
    e = lv0["e"]
    rho3 = bf.from_flat(rho, e)
    rhop3 = (rho3.astype(np.float64) ** 3).astype(np.float32)
    fixed3 = bf.from_flat(fixed.astype(bool), lv0["n"])
    P = bf.Problem(a, rho3, fixed3, rhop=rhop3)
    H.set_fixed(fixed)
    H.set_rho(rho)
    H.set_rhop(bf.to_flat(rhop3))
    return H, P, fixed, f


def _bf(v):   # [3, nn] -> interleaved dofs
    return v.T.reshape(-1)


def _or(v):
    return v.reshape(-1, 3).T


@pytest.fixture(scope="module", params=["holes3", "lshape4"])
def setup(request):
    return _setup(request.param, 3)


def test_residual_and_gs(setup):
    H, P, fixed, f = setup
    nn = H.levels[0]["nn"]
    rng = np.random.default_rng(5)
    u = rng.standard_normal((3, nn)) * H.levels[0]["vact"]
    u[:, fixed.astype(bool)] = 0
    H.set_vec(0, 0, u); H.set_vec(0, 1, f)
    H.update_residual(0)
    assert rel_err(H.get_vec(0, 2), _or(P.residual(0, _bf(u), _bf(f)))) < 1e-13
    for rev in (False, True):
        H.set_vec(0, 0, u)
        H.gs_relax(0, 2, rev)
        assert rel_err(H.get_vec(0, 0), _or(P.gs(0, _bf(u), _bf(f), 2, rev))) < 1e-13
    # apply = K x with the WS rules
    y = H.apply(u)
    assert rel_err(y, _or(P.A_res @ _bf(u))) < 1e-13


def test_transfers_and_galerkin(setup):
    H, P, fixed, f = setup
    H.update_stencil()
    rng = np.random.default_rng(7)
    for li in range(1, H.nlevels):
        st = H.get_stencil(li)
        assert rel_err(st, P.coarse_stencil(li)) < 1e-12
        r = rng.standard_normal((3, H.levels[li - 1]["nn"])) * H.levels[li - 1]["vact"]
        H.set_vec(li - 1, 2, r)
        H.restrict_residual(li)
        assert rel_err(H.get_vec(li, 1), _or(P.P[li - 1].T @ _bf(r))) < 1e-14
        uf = rng.standard_normal((3, H.levels[li - 1]["nn"])) * H.levels[li - 1]["vact"]
        uc = rng.standard_normal((3, H.levels[li]["nn"])) * H.levels[li]["vact"]
        H.set_vec(li - 1, 0, uf); H.set_vec(li, 0, uc)
        H.prolongate_correction(li - 1)
        assert rel_err(H.get_vec(li - 1, 0), _or(_bf(uf) + P.P[li - 1] @ _bf(uc))) < 1e-14


def test_coarse_ops_and_vcycle(setup):
    H, P, fixed, f = setup
    H.update_stencil()
    rng = np.random.default_rng(9)
    for li in range(1, H.nlevels - 1):
        act = H.levels[li]["vact"]
        u = rng.standard_normal((3, H.levels[li]["nn"])) * act
        ff = rng.standard_normal((3, H.levels[li]["nn"])) * act
        H.set_vec(li, 0, u); H.set_vec(li, 1, ff)
        H.update_residual(li)
        assert rel_err(H.get_vec(li, 2), _or(P.residual(li, _bf(u), _bf(ff)))) < 1e-12
        H.gs_relax(li, 1, False)
        assert rel_err(H.get_vec(li, 0), _or(P.gs(li, _bf(u), _bf(ff), 1))) < 1e-12
    L = H.nlevels - 1
    ff = rng.standard_normal((3, H.levels[L]["nn"])) * H.levels[L]["vact"]
    H.set_vec(L, 1, ff)
    H.solve_coarsest()
    assert rel_err(H.get_vec(L, 0), _or(P.coarsest_solve(_bf(ff)))) < 1e-9
    nn = H.levels[0]["nn"]
    H.set_vec(0, 0, np.zeros((3, nn))); H.set_vec(0, 1, f)
    u = np.zeros(3 * nn)
    for sym in (False, True):
        rr = H.v_cycle(1, 1, sym)
        u, rb = P.vcycle(u, _bf(f), 1, 1, symmetric=sym)
        assert abs(rr - rb) / rb < 1e-10
        assert rel_err(H.get_vec(0, 0), _or(u)) < 1e-10


def test_vcycle_converges_to_direct_solve(setup):
    import scipy.sparse.linalg as spla
    H, P, fixed, f = setup
    H.update_stencil()
    nn = H.levels[0]["nn"]
    H.set_vec(0, 0, np.zeros((3, nn))); H.set_vec(0, 1, f)
    for _ in range(100):
        if H.v_cycle(2, 2) < 1e-11:
            break
    act = np.nonzero(P.active[0] & ~np.repeat(fixed.astype(bool), 3))[0]
    A = P.A_res[act][:, act].tocsc()
    x = np.zeros(3 * nn)
    x[act] = spla.spsolve(A, _bf(f)[act])
    assert rel_err(H.get_vec(0, 0), _or(x)) < 1e-8
    assert abs(H.compliance() - _bf(f) @ x) / abs(_bf(f) @ x) < 1e-9


def test_sensitivity_filter_oc(setup):
    H, P, fixed, f = setup
    lv0 = H.levels[0]
    rng = np.random.default_rng(11)
    u = rng.standard_normal((3, lv0["nn"])) * lv0["vact"]
    H.set_vec(0, 0, u)
    g = H.sensitivity()
    rho = H.get_rho()
    gb = bf.sensitivity(P.L.levels[0], P.KE, bf.from_flat(rho, lv0["e"]), 3.0, _bf(u))
    assert rel_err(g, bf.to_flat(gb)) < 1e-6
    gf = H.filter(2.0, g)
    gfb = bf.filter_sens(P.L.levels[0], bf.from_flat(g.astype(np.float64), lv0["e"]), 2.0)
    assert rel_err(gf, bf.to_flat(gfb)) < 1e-6
    thr, nt = H.oc_update(0.4, 0.03, 0.5, 1e-3, gf)
    rb, thrb = bf.oc_update(lv0["eact"], rho, gf, 0.4, 0.03, 0.5, 1e-3)
    assert abs(thr - thrb) <= 1e-6 * abs(thrb)
    assert rel_err(H.get_rho(), rb) < 1e-6


# --- voxelizer ------------------------------------------------------------------
def box_mesh(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    V = np.array([[x0, y0, z0], [x1, y0, z0], [x0, y1, z0], [x1, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x0, y1, z1], [x1, y1, z1]], np.float32)
    Fc = np.array([[0, 2, 1], [1, 2, 3], [4, 5, 6], [5, 7, 6], [0, 1, 4], [1, 5, 4],
                   [2, 6, 3], [3, 6, 7], [0, 4, 2], [2, 4, 6], [1, 3, 5], [3, 7, 5]], np.int32)
    return V, Fc


def test_oc_shell_elements(setup):
    """Shell elements are set to density 1 in every bisection trial
    (optimization.cu:159), so they also count as 1 in the volume."""
    H, P, fixed, f = setup
    lv0 = H.levels[0]
    e = lv0["e"]
    rng = np.random.default_rng(12)
    rho0 = (rng.uniform(0.2, 0.6, lv0["ne"]) * lv0["eact"]).astype(np.float32)
    g = (-rng.uniform(0, 1, lv0["ne"]) * lv0["eact"]).astype(np.float32)
    act = bf.from_flat(lv0["eact"], e)
    shell = np.zeros_like(act)
    shell[0] = shell[-1] = True
    shell &= act
    sh = bf.to_flat(shell)
    assert 0 < sh.sum() < lv0["eact"].sum()
    H.set_rho(rho0)
    H.set_shell(sh)
    thr, nt = H.oc_update(0.45, 0.03, 0.5, 1e-3, g)
    rb, thrb = bf.oc_update(lv0["eact"], rho0, g, 0.45, 0.03, 0.5, 1e-3, shell=sh)
    assert abs(thr - thrb) <= 1e-6 * abs(thrb)
    r = H.get_rho()
    assert rel_err(r, rb) < 1e-6
    assert np.all(r[sh] == 1) and np.all(r[lv0["eact"] & ~sh] < 1)
    # without the shell the multiplier and densities differ
    H.set_rho(rho0)
    H.set_shell(np.zeros_like(sh))
    thr2, _ = H.oc_update(0.45, 0.03, 0.5, 1e-3, g)
    assert thr2 != thr


def test_voxelize_box_exact():
    """A box whose faces avoid voxel centres voxelizes to exactly the voxels
    whose centres it contains."""
    R = 8
    bmin = np.zeros(3, np.float32)
    unit = np.ones(3, np.float32)
    c = np.arange(R) + 0.5
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    for lo, hi in [((0.6, 0.2, 0.3), (5.3, 3.7, 2.45)), ((1.7, 2.1, 0.9), (7.2, 7.9, 6.6))]:
        tris_v, tris_f = box_mesh(lo, hi)
        bits = oracle.voxelize_triangles(tris_v[tris_f].reshape(-1), bmin, unit, R)
        inside = (X > lo[0]) & (X < hi[0]) & (Y > lo[1]) & (Y < hi[1]) & (Z > lo[2]) & (Z < hi[2])
        assert np.array_equal(bf.bits_to_cube(bits, R), inside)


def test_voxelize_truncation_quirk():
    """xmax = int(x/unit - 0.5) truncates toward zero (voxelize_solid.cu:131): a
    face whose hit lies in (0, 0.5) voxel also flips voxel 0, so a box starting
    at x = 0.1 loses its first voxel column."""
    R = 8
    tris_v, tris_f = box_mesh((0.1, 0.2, 0.3), (5.3, 3.7, 2.45))
    bits = oracle.voxelize_triangles(tris_v[tris_f].reshape(-1), np.zeros(3, np.float32), np.ones(3, np.float32), R)
    act = bf.bits_to_cube(bits, R)
    assert not act[0].any() and act[1:5, :4, :2].all() and act.sum() == 32


def test_voxelize_mesh_padding():
    V, Fc = box_mesh((0.0, 0.0, 0.0), (2.0, 1.0, 1.0))
    bits, R, box = oracle.voxelize_mesh(V, Fc, 20)
    assert R == 32
    side = 2.0 * 32 / 20
    assert np.allclose(box[1] - box[0], side, rtol=1e-6)
    assert np.allclose((box[0] + box[1]) / 2, [1.0, 0.5, 0.5], atol=1e-6)
    act = bf.bits_to_cube(bits, R)
    assert 0 < act.sum() < R ** 3
