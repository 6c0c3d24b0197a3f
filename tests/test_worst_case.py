"""Worst-case load search: forceProject and modifiedPM
(projection.cu:165-231, projection.cpp:289-360, optimization.cpp:123-261).

CPU: the oracle is pinned against what the mathematics fixes.
  - forceProject equals f - Nd^T Nd f with Nd built from an orthonormal
    tangent pair per load node (projection.cpp:176-197), computed here by an
    SVD null space rather than the normal-component shortcut.
  - modifiedPM converges to the largest eigenvalue of the aggregated
    compliance operator C = N^T K^-1 N (the reference's own cross-check,
    test_utils.cpp:1249-1296), with K assembled by brute force and factored by
    scipy. Its final load is the top eigenvector. The reference's stopping
    rule (load change 1e-4, residual 1e-2) bounds the agreement.
GPU: the HIP path reproduces the oracle step by step (FP64 V-cycle) from the
same start: the same number of V-cycles and the same worst compliance, load
and displacement to 1e-10.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import bruteforce as bf
import oracle
from problems import active_cube, bits_of, problem_inputs, rel_err


def _problem(kind="holes3", seed=3):
    a = active_cube(kind, 32)
    H = oracle.Hierarchy(32, bits_of(a))
    H.set_material(1e6, 0.4, 3.0)
    lv0 = H.levels[0]
    fixed, _, rho = problem_inputs(lv0, seed)
    rhop = (rho.astype(np.float64) ** 3).astype(np.float32)
    H.set_fixed(fixed)
    H.set_rho(rho)
    H.set_rhop(rhop)
    n = lv0["n"]
    vact = lv0["vact"].reshape(n[2], n[1], n[0])
    load = np.zeros_like(vact)
    zs = np.nonzero(vact.any(axis=(1, 2)))[0]
    load[zs[-1]] = vact[zs[-1]]                  # top face of the solid
    load = load.reshape(-1) & ~fixed.astype(bool)
    rng = np.random.default_rng(5)
    nrm = np.zeros((3, lv0["nn"]))
    nrm[2] = 1.0
    nrm += 0.3 * rng.standard_normal(nrm.shape)   # tilted normals: every component matters
    f0 = np.random.default_rng(1).uniform(-1, 1, (3, lv0["nn"])) * lv0["vact"][None]
    return a, H, fixed, rho, rhop, load, nrm, f0


@pytest.fixture(scope="module")
def prob():
    return _problem()


def test_force_project_matches_tangent_basis(prob):
    a, H, fixed, rho, rhop, load, nrm, f0 = prob
    nn = H.levels[0]["nn"]
    for normals in (None, nrm):
        H.set_load_nodes(load.astype(np.uint8), normals)
        H.set_vec(0, 1, f0)
        H.force_project()
        got = H.get_vec(0, 1)
        ref = np.zeros((3, nn))
        for v in np.nonzero(load)[0]:
            if normals is None:
                ref[:, v] = f0[:, v]
            else:
                _, _, vt = np.linalg.svd(normals[:, v][None, :])
                Nd = vt[1:]                      # orthonormal tangent pair
                ref[:, v] = f0[:, v] - Nd.T @ (Nd @ f0[:, v])
        assert rel_err(got, ref) < 1e-13


def _dense_top(prob, normals, rhop_used=None):
    """Largest eigenpair of the aggregated compliance C = N^T K^-1 N, with K
    assembled by brute force and factored by scipy (test_utils.cpp:1249-1296
    is the reference's own form of this check). Returns (lambda_max, the top
    load as an interleaved-dof vector)."""
    a, H, fixed, rho, rhop, load, nrm, f0 = prob
    if rhop_used is not None:
        rhop = rhop_used
    nn = H.levels[0]["nn"]
    P = bf.Problem(a, bf.from_flat(rho, H.levels[0]["e"]), bf.from_flat(fixed.astype(bool), H.levels[0]["n"]),
                   rhop=bf.from_flat(rhop, H.levels[0]["e"]))
    act = np.nonzero(P.active[0] & ~np.repeat(fixed.astype(bool), 3))[0]
    lu = spla.splu(P.A_res[act][:, act].tocsc())
    cols = []
    for v in np.nonzero(load)[0]:
        dirs = [np.eye(3)[k] for k in range(3)] if normals is None else [normals[:, v] / np.linalg.norm(normals[:, v])]
        for d in dirs:
            x = np.zeros(3 * nn)
            x[3 * v:3 * v + 3] = d
            cols.append(x)
    N = np.array(cols).T
    KN = np.zeros_like(N)
    KN[act] = lu.solve(N[act])
    w, V = np.linalg.eigh(N.T @ KN)
    return w[-1], N @ V[:, -1]


@pytest.mark.parametrize("mode", ["wsff", "wscf"])
def test_modified_pm_finds_top_eigenpair(prob, mode):
    a, H, fixed, rho, rhop, load, nrm, f0 = prob
    normals = nrm if mode == "wscf" else None
    H.update_stencil()
    H.set_load_nodes(load.astype(np.uint8), normals)
    H.set_vec(0, 1, f0)
    c, steps = H.modified_pm(500)
    assert 1 < steps < 500
    lam, ftop = _dense_top(prob, normals)
    assert abs(c - lam) / lam < 5e-4
    f = H.get_vec(0, 1).T.reshape(-1)            # interleaved dofs
    assert abs(abs(f @ ftop) - 1) < 1e-3
    assert abs(np.linalg.norm(f) - 1) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["wsff", "wscf"])
def test_gpu_modified_pm_matches_oracle(prob, mode):
    import robtop_amd as rt
    a, H, fixed, rho, rhop, load, nrm, f0 = prob
    normals = nrm if mode == "wscf" else None
    print(f"[pm] holes3 {mode}", flush=True)
    g = rt.HierarchyGrid.genFromMesh(bits_of(a), 32)
    g.set_material(1e6, 0.4, 3.0)
    g.set_support(fixed)
    g.set_density(rho)
    H = _problem()[1]                              # a fresh oracle with the device's rho^p
    H.set_rhop(g.get_rhop())
    g.set_precision(64)
    g.update_stencil()
    nl = g.set_load_nodes(load.astype(np.uint8), normals)
    assert nl == int(load.sum())
    f0z = f0.copy()
    f0z[:, fixed.astype(bool)] = 0                 # set_force clears supports; they are not load nodes
    g.set_force(f0z)
    cg, sg = g.modified_pm(500)

    H.update_stencil()
    H.set_load_nodes(load.astype(np.uint8), normals)
    H.set_vec(0, 1, f0z)
    co, so = H.modified_pm(500)
    assert sg == so
    assert abs(cg - co) / abs(co) < 1e-10
    assert rel_err(g.get_force(), H.get_vec(0, 1)) < 1e-10
    assert rel_err(g.get_displacement(), H.get_vec(0, 0)) < 1e-10
    g.close()


@pytest.mark.gpu
def test_gpu_rand_force_and_project():
    import robtop_amd as rt
    a, H, fixed, rho, rhop, load, nrm, f0 = _problem()
    g = rt.HierarchyGrid.genFromMesh(bits_of(a), 32)
    g.set_material(1e6, 0.4, 3.0)
    g.set_support(fixed)
    g.set_load_nodes(load.astype(np.uint8), nrm)
    g.rand_force(7)
    f = g.get_force()
    vact = H.levels[0]["vact"].astype(bool)
    assert np.all(np.abs(f[:, vact]) <= 1) and not f[:, ~vact].any()
    assert abs(f[:, vact].mean()) < 0.05 and 0.30 < f[:, vact].var() < 0.37   # uniform(-1, 1): var 1/3
    g.rand_force(7)
    assert np.array_equal(g.get_force(), f)
    g.force_project()
    H.set_load_nodes(load.astype(np.uint8), nrm)
    H.set_vec(0, 1, f)
    H.force_project()
    assert rel_err(g.get_force(), H.get_vec(0, 1)) < 1e-14
    g.close()


@pytest.mark.gpu
def test_gpu_torus_worst_case_loop():
    """The worst-case SIMP loop (optimization.cpp:583-673) on a voxelized torus
    with normal-direction loads: three iterations run, and at the final
    densities the worst compliance bounds the compliance of random admissible
    unit loads (it is the largest eigenvalue of N^T K^-1 N)."""
    import robtop_amd as rt
    from robtop_amd.workloads import torus_bits, torus_worst_case_bc
    n = 48
    print(f"[pm] torus n={n}", flush=True)
    bits, R, solid = torus_bits(n)
    g = rt.HierarchyGrid.genFromMesh(bits, R)
    g.set_material(1e6, 0.4, 3.0)
    fixed, load, normals = torus_worst_case_bc(g, n)
    assert fixed.any() and load.any()
    g.set_support(fixed)
    nl = g.set_load_nodes(load.astype(np.uint8), normals)
    assert nl == int(load.sum())
    g.set_precision(32)
    params = rt.Parameter(volume_ratio=0.3, volume_decrease=0.05)
    recs = rt.worst_case_topopt(g, params, n_iter=3, stop=False)
    assert len(recs) == 3 and all(np.isfinite(r["worst_compliance"]) and r["worst_compliance"] > 0 for r in recs)
    g.update_stencil()
    g.rand_force(99)
    cw, _ = g.modified_pm(500)
    for seed in range(3):
        g.rand_force(1000 + seed)
        g.force_project()
        f = g.get_force()
        f /= np.linalg.norm(f)
        g.set_force(f)
        g.set_displacement(np.zeros_like(f))
        st = g.solve(tol=1e-8, maxit=500, mode=rt.RT_SOLVE_MIXED, pre=2, post=2)
        assert st.rel_res < 1e-7
        assert g.compliance() <= cw * (1 + 2e-3)
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["wsff", "wscf"])
def test_gpu_block_lanczos_top_eigenpair(prob, mode):
    """Block Lanczos (k = 8, no reference counterpart) reaches the largest
    eigenvalue of the dense C to 1e-7 with FP64 solves and 1e-5 with the
    reliable-update FP32 solver, returns its eigenvector as a unit load, and
    agrees with the modified power method within the latter's stopping rule."""
    import robtop_amd as rt
    a, H, fixed, rho, rhop, load, nrm, f0 = prob
    normals = nrm if mode == "wscf" else None
    print(f"[lanczos] holes3 {mode}", flush=True)
    g = rt.HierarchyGrid.genFromMesh(bits_of(a), 32)
    g.set_material(1e6, 0.4, 3.0)
    g.set_support(fixed)
    g.set_density(rho)
    g.update_stencil()
    g.set_load_nodes(load.astype(np.uint8), normals)
    lam, ftop = _dense_top(prob, normals, g.get_rhop())     # the device's rho^p (FP32 powf)
    c, st = g.block_lanczos(k=8, max_blocks=40, tol=1e-9, solve_tol=1e-11, mode=rt.RT_SOLVE_FP64, seed=3)
    assert abs(c - lam) / lam < 1e-7, (c, lam, st)
    assert 1 < st["solves"] <= 8 * st["blocks"] + 1
    f = g.get_force().T.reshape(-1)
    assert abs(np.linalg.norm(f) - 1) < 1e-12
    assert abs(abs(f @ ftop) - 1) < 1e-6
    c32, _ = g.block_lanczos(k=8, max_blocks=40, tol=1e-7, solve_tol=1e-8, seed=4)
    assert abs(c32 - lam) / lam < 1e-5, (c32, lam)
    g.rand_force(7)
    cpm, _ = g.modified_pm(500)
    assert abs(cpm - lam) / lam < 5e-4
    g.close()


@pytest.mark.gpu
def test_gpu_torus_lanczos_matches_pm():
    """On the voxelized torus with normal-direction loads, the worst-case loop
    runs with block Lanczos, and at the final densities block Lanczos and the
    modified power method find the same worst compliance (the power method to
    its stopping rule)."""
    import robtop_amd as rt
    from robtop_amd.workloads import torus_bits, torus_worst_case_bc
    n = 48
    print(f"[lanczos] torus n={n}", flush=True)
    bits, R, solid = torus_bits(n)
    g = rt.HierarchyGrid.genFromMesh(bits, R)
    g.set_material(1e6, 0.4, 3.0)
    fixed, load, normals = torus_worst_case_bc(g, n)
    g.set_support(fixed)
    g.set_load_nodes(load.astype(np.uint8), normals)
    g.set_precision(32)
    params = rt.Parameter(volume_ratio=0.3, volume_decrease=0.05)
    lz = dict(k=8, max_blocks=12, tol=1e-6, solve_tol=1e-8)
    recs = rt.worst_case_topopt(g, params, n_iter=2, stop=False, method="lanczos", lanczos=lz)
    assert len(recs) == 2 and all(r["worst_compliance"] > 0 and r["solves"] > 8 for r in recs)
    g.update_stencil()
    cl, st = g.block_lanczos(seed=11, **lz)
    g.rand_force(5)
    cp, _ = g.modified_pm(500)
    assert abs(cl - cp) / cl < 2e-3, (cl, cp, st)
    g.close()


# --------------------------------------------------------------------------
# no-support modes (nsff, nscf): balanced loads and rigid-motion removal
# (projection.cpp:46-112, :208-257, :322-335, :374-394; optimization.cpp:189-194)
# --------------------------------------------------------------------------
def _ns_problem():
    a, H, fixed, rho, rhop, load, nrm, f0 = _problem()
    H.set_fixed(np.zeros_like(fixed))              # a free body
    return a, H, rho, rhop, load, nrm, f0


def _rigid(lv0, mask):
    """orthonormal rigid motions of the nodes in mask (interleaved dofs of the
    whole node box), by QR"""
    n = lv0["n"]
    nn = lv0["nn"]
    idx = np.nonzero(mask)[0]
    x, y, z = idx % n[0], idx // n[0] % n[1], idx // (n[0] * n[1])
    R = np.zeros((3 * nn, 6))
    for c in range(3):
        R[3 * idx + c, c] = 1.0
    # p x e_k columns: rotation k maps p to e_k x p
    R[3 * idx + 1, 3], R[3 * idx + 2, 3] = -z, y        # e_x x p = (0, -z, y)
    R[3 * idx + 0, 4], R[3 * idx + 2, 4] = z, -x        # e_y x p = (z, 0, -x)
    R[3 * idx + 0, 5], R[3 * idx + 1, 5] = -y, x        # e_z x p = (-y, x, 0)
    q, _ = np.linalg.qr(R)
    return q


def _admissible(lv0, load, normals):
    """orthonormal basis (interleaved dofs) of the balanced admissible loads:
    f on the load nodes, along the normals if given, with no net force or
    moment (R^T f = 0 for every rigid motion R of the whole body)"""
    nn = lv0["nn"]
    cols = []
    for v in np.nonzero(load)[0]:
        dirs = [np.eye(3)[k] for k in range(3)] if normals is None else [normals[:, v] / np.linalg.norm(normals[:, v])]
        for d in dirs:
            x = np.zeros(3 * nn)
            x[3 * v:3 * v + 3] = d
            cols.append(x)
    B = np.array(cols).T
    Rf = _rigid(lv0, lv0["vact"].astype(bool))
    _, s, vt = np.linalg.svd(Rf.T @ B)
    Z = vt[int(np.sum(s > 1e-10 * s[0])):].T
    return B @ Z


@pytest.mark.parametrize("mode", ["nsff", "nscf"])
def test_ns_force_project_is_the_balanced_projector(mode):
    a, H, rho, rhop, load, nrm, f0 = _ns_problem()
    normals = nrm if mode == "nscf" else None
    lv0 = H.levels[0]
    H.set_load_nodes(load.astype(np.uint8), normals)
    H.set_vec(0, 1, f0)
    H.force_project()
    got = H.get_vec(0, 1).T.reshape(-1)
    S = _admissible(lv0, load, normals)
    ref = S @ (S.T @ f0.T.reshape(-1))
    assert rel_err(got, ref) < 1e-10
    # no net force or moment remains
    Rf = _rigid(lv0, lv0["vact"].astype(bool))
    assert np.abs(Rf.T @ got).max() < 1e-10 * np.abs(got).max()


def _ns_dense_top(a, H, rho, rhop, load, normals):
    """largest eigenpair of the compliance on the balanced admissible loads,
    S^T K^+ S, with K^+ b the minimum-norm solution: a statically determinate
    3-2-1 support pins the rigid motion for a sparse solve, then the rigid part
    is projected out"""
    lv0 = H.levels[0]
    nn = lv0["nn"]
    fixed = np.zeros(nn, bool)
    P = bf.Problem(a, bf.from_flat(rho, lv0["e"]), bf.from_flat(fixed, lv0["n"]), rhop=bf.from_flat(rhop, lv0["e"]))
    K = P.K0.tocsr()
    act_nodes = np.nonzero(lv0["vact"])[0]
    n = lv0["n"]
    xs, ys, zs = act_nodes % n[0], act_nodes // n[0] % n[1], act_nodes // (n[0] * n[1])
    A = act_nodes[0]
    B_ = act_nodes[np.argmax(xs - xs[0] - 1e-3 * np.abs(ys - ys[0]) - 1e-3 * np.abs(zs - zs[0]))]
    C = act_nodes[np.argmax(ys - ys[0] - 1e-3 * np.abs(zs - zs[0]))]
    pinned = {3 * A, 3 * A + 1, 3 * A + 2, 3 * B_ + 1, 3 * B_ + 2, 3 * C + 2}
    act = np.array([d for d in np.nonzero(np.repeat(lv0["vact"].astype(bool), 3))[0] if d not in pinned])
    lu = spla.splu(K[act][:, act].tocsc())
    Rf = _rigid(lv0, lv0["vact"].astype(bool))
    S = _admissible(lv0, load, normals)
    X = np.zeros_like(S)
    X[act] = lu.solve(S[act])
    X -= Rf @ (Rf.T @ X)
    w, V = np.linalg.eigh(S.T @ X)
    return w[-1], S @ V[:, -1]


@pytest.mark.parametrize("mode", ["nsff", "nscf"])
def test_ns_modified_pm_finds_top_eigenpair(mode):
    a, H, rho, rhop, load, nrm, f0 = _ns_problem()
    normals = nrm if mode == "nscf" else None
    H.update_stencil()
    H.set_load_nodes(load.astype(np.uint8), normals)
    H.set_vec(0, 1, f0)
    c, steps = H.modified_pm(500)
    assert 1 < steps < 500
    lam, ftop = _ns_dense_top(a, H, rho, rhop, load, normals)
    assert abs(c - lam) / lam < 5e-4
    f = H.get_vec(0, 1).T.reshape(-1)
    assert abs(abs(f @ ftop) - 1) < 1e-3
    u = H.get_vec(0, 0).T.reshape(-1)
    Rf = _rigid(H.levels[0], H.levels[0]["vact"].astype(bool))
    assert np.abs(Rf.T @ u).max() < 1e-10 * np.abs(u).max()     # no rigid motion in the displacement


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["nsff", "nscf"])
def test_gpu_ns_modified_pm_matches_oracle(mode):
    """No supports: the balanced-load projection, the singular coarsest solve
    (minimum norm) and the rigid-motion removal after each V-cycle, step by
    step against the oracle (FP64 V-cycle)."""
    import robtop_amd as rt
    a, H, rho, rhop, load, nrm, f0 = _ns_problem()
    normals = nrm if mode == "nscf" else None
    print(f"[pm] holes3 {mode}, no supports", flush=True)
    g = rt.HierarchyGrid.genFromMesh(bits_of(a), 32)
    g.set_material(1e6, 0.4, 3.0)
    g.set_density(rho)
    H.set_rhop(g.get_rhop())
    g.set_precision(64)
    g.update_stencil()
    nl = g.set_load_nodes(load.astype(np.uint8), normals)
    assert nl == int(load.sum())
    g.set_force(f0)
    g.force_project()
    H.set_load_nodes(load.astype(np.uint8), normals)
    H.set_vec(0, 1, f0)
    H.force_project()
    assert rel_err(g.get_force(), H.get_vec(0, 1)) < 1e-12
    g.set_force(f0)
    cg, sg = g.modified_pm(500)
    H.update_stencil()
    H.set_vec(0, 1, f0)
    co, so = H.modified_pm(500)
    print(f"[pm] {mode}: GPU {sg} cycles c {cg:.12e}, oracle {so} cycles c {co:.12e}", flush=True)
    assert sg == so
    assert abs(cg - co) / abs(co) < 1e-10
    assert rel_err(g.get_force(), H.get_vec(0, 1)) < 1e-10
    assert rel_err(g.get_displacement(), H.get_vec(0, 0)) < 1e-10
    g.close()
