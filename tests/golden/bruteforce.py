"""Independent brute-force restatement used to pin the oracle (test infrastructure).

Everything here is assembled from first principles on tiny grids with numpy /
scipy, without sharing code or loop structure with ``oracle/oracle.cpp``:

* KE from the bilinear form  a(u, v) = int lam div u div v + 2 mu eps(u):eps(v)
  integrated with 3x3x3 Gauss-Legendre (exact for the trilinear element), not
  with the reference's 2x2x2 B^T D B loop (element/templateMatrix.cpp:31-124);
* global K by sparse assembly, the WS rules of Grid.cu:922-1066 / :1501-1557 as
  matrix edits (fixed rows/columns removed, n_e I on fixed diagonals);
* point Gauss-Seidel in (colour, component) order on that matrix;
* transfers as an explicit sparse trilinear P (Grid.cu:1680-1895);
* Galerkin as P^T A~ P with the Dirichlet-penalised element matrices
  (Grid.cu:392-506 semantics: fixed rows/cols zeroed, 1e6 on fixed diagonals);
* coarsest by numpy dense solve; sensitivity, filter, OC by direct formulas.

The layout conventions (bounding box aligned to the coarsest spacing, natural
x-fastest order) are the ones the oracle and the library share, so arrays can
be compared element by element.
"""
import numpy as np
import scipy.sparse as sp


# --------------------------------------------------------------------------
# element stiffness
# --------------------------------------------------------------------------
def ke_exact(E=1.0, nu=0.4, h=1.0):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    g, w = np.polynomial.legendre.leggauss(3)
    g = (g + 1) / 2
    w = w / 2
    K = np.zeros((24, 24))
    verts = [(i % 2, i // 2 % 2, i // 4) for i in range(8)]
    for ix in range(3):
        for iy in range(3):
            for iz in range(3):
                x = (g[ix], g[iy], g[iz])
                wt = w[ix] * w[iy] * w[iz]
                grad = np.zeros((8, 3))
                for i, v in enumerate(verts):
                    f = [x[k] if v[k] else 1 - x[k] for k in range(3)]
                    d = [1.0 if v[k] else -1.0 for k in range(3)]
                    grad[i] = [d[0] * f[1] * f[2], f[0] * d[1] * f[2], f[0] * f[1] * d[2]]
                grad /= h
                for i in range(8):
                    for a in range(3):
                        for j in range(8):
                            for b in range(3):
                                val = lam * grad[i, a] * grad[j, b] + mu * grad[i, b] * grad[j, a]
                                if a == b:
                                    val += mu * grad[i] @ grad[j]
                                K[3 * i + a, 3 * j + b] += wt * val
    return K * h ** 3


# --------------------------------------------------------------------------
# hierarchy (reference rule: coarsen while > 400 active, layer 1 dummy)
# --------------------------------------------------------------------------
def bits_to_cube(bits, R):
    idx = np.arange(R ** 3, dtype=np.int64)
    b = (bits[idx >> 5] >> (idx & 31).astype(np.uint32)) & 1
    return b.reshape(R, R, R).transpose(2, 1, 0).astype(bool)  # [x, y, z]


def cube_to_bits(act):
    R = act.shape[0]
    flat = act.transpose(2, 1, 0).reshape(-1).astype(np.uint64)
    n = (R ** 3 + 31) // 32
    words = np.zeros(n, dtype=np.uint64)
    idx = np.arange(flat.size)
    np.add.at(words, idx >> 5, flat << (idx & 31).astype(np.uint64))
    return words.astype(np.uint32)


class Levels:
    def __init__(self, act0):
        R = act0.shape[0]
        acts = [act0]
        while acts[-1].sum() > 400:
            a = acts[-1]
            r = a.shape[0] // 2
            acts.append(a.reshape(r, 2, r, 2, r, 2).any(axis=(1, 3, 5)))
        self.nlayers = len(acts)
        assert self.nlayers >= 3
        K = self.nlayers - 1
        A = 1 << K
        nz = np.nonzero(act0)
        lo = np.array([c.min() for c in nz])
        hi = np.array([c.max() + 1 for c in nz])
        lo_al = (lo // A) * A
        hi_al = np.minimum(R, -(-hi // A) * A)
        self.levels = []
        for layer in [0] + list(range(2, self.nlayers)):
            o = lo_al >> layer
            e = (hi_al - lo_al) >> layer
            ea = acts[layer][o[0]:o[0] + e[0], o[1]:o[1] + e[1], o[2]:o[2] + e[2]]
            pad = np.pad(ea, 1)
            va = np.zeros(tuple(e + 1), bool)
            for k in range(8):
                dx, dy, dz = k % 2, k // 2 % 2, k // 4
                va |= pad[dx:dx + e[0] + 1, dy:dy + e[1] + 1, dz:dz + e[2] + 1]
            self.levels.append(dict(layer=layer, scale=1 << layer, o=o, e=e, n=e + 1, eact=ea, vact=va))
        self.counts = [int(a.sum()) for a in acts]


def flat_nodes(n):
    """node index (x fastest) for a [nx, ny, nz] array."""
    return np.arange(n[0] * n[1] * n[2]).reshape(n[2], n[1], n[0]).transpose(2, 1, 0)


def to_flat(a):
    """[x, y, z] array -> x-fastest flat vector."""
    return a.transpose(2, 1, 0).reshape(-1)


def from_flat(v, shape):
    return v.reshape(shape[2], shape[1], shape[0]).transpose(2, 1, 0)


# --------------------------------------------------------------------------
# assembly
# --------------------------------------------------------------------------
def element_dofs(lv):
    e = lv['e']
    nid = flat_nodes(lv['n'])
    ex, ey, ez = np.nonzero(lv['eact'])
    dofs = np.zeros((ex.size, 24), np.int64)
    for k in range(8):
        v = nid[ex + k % 2, ey + k // 2 % 2, ez + k // 4]
        for c in range(3):
            dofs[:, 3 * k + c] = 3 * v + c
    return (ex, ey, ez), dofs


def assemble(lv, elem_mats):
    """sum of per-element 24x24 matrices (array [ne_active, 24, 24])."""
    _, dofs = element_dofs(lv)
    nd = 3 * int(np.prod(lv['n']))
    rows = np.repeat(dofs[:, :, None], 24, axis=2).reshape(-1)
    cols = np.repeat(dofs[:, None, :], 24, axis=1).reshape(-1)
    return sp.csr_matrix((elem_mats.reshape(-1), (rows, cols)), shape=(nd, nd))


def rho_p(rho, p):
    """FP32 rho^p. Rounded from a float64 power; the reference's FP32 powf and
    this differ by at most a unit in the last place, so comparisons that must be
    exact pass rho^p explicitly (``rhop=``)."""
    return (rho.astype(np.float64) ** p).astype(np.float32).astype(np.float64)


def fine_K(lv, KE, rho, p, rhop=None):
    (ex, ey, ez), _ = element_dofs(lv)
    rp = rho_p(rho[ex, ey, ez], p) if rhop is None else rhop[ex, ey, ez].astype(np.float64)
    return assemble(lv, rp[:, None, None] * KE[None])


def ws_matrices(lv, K, fixed):
    """(A_res, A_gs): residual operator and GS operator under the WS rules."""
    nn = int(np.prod(lv['n']))
    fdof = np.repeat(to_flat(fixed), 3)
    D = sp.diags((~fdof).astype(float))
    A_res = (D @ K @ D).tocsr()
    # n_e: number of active incident elements per node
    pad = np.pad(lv['eact'], 1)
    n = lv['n']
    ne = np.zeros(tuple(n), int)
    for k in range(8):
        dx, dy, dz = k % 2, k // 2 % 2, k // 4
        ne += pad[dx:dx + n[0], dy:dy + n[1], dz:dz + n[2]]
    diag_fix = np.repeat(to_flat(ne * fixed), 3).astype(float)
    A_gs = (A_res + sp.diags(diag_fix)).tocsr()
    return A_res, A_gs


def colours(lv):
    n = lv['n']
    x, y, z = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing='ij')
    a = lv['o'] * 0  # box origin is even; parity of box coords = absolute parity
    return to_flat((x & 1) | ((y & 1) << 1) | ((z & 1) << 2))


def point_gs(A, u, f, colour, active, nsweep=1, reverse=False):
    u = u.copy()
    A = A.tocsr()
    diag = A.diagonal()
    order = list(range(8))
    comps = [0, 1, 2]
    if reverse:
        order = order[::-1]
        comps = comps[::-1]
    for _ in range(nsweep):
        for c in order:
            nodes = np.nonzero((colour == c) & active)[0]
            for a in comps:
                rows = 3 * nodes + a
                Ar = A[rows]
                s = Ar @ u - diag[rows] * u[rows]
                u[rows] = (f[rows] - s) / diag[rows]
    return u


def prolongation(fl, cl):
    """sparse P: coarse active nodes -> fine active nodes (trilinear, any ratio)."""
    ratio = cl['scale'] // fl['scale']
    nf = fl['n']
    fx, fy, fz = np.nonzero(fl['vact'])
    fid = flat_nodes(nf)[fx, fy, fz]
    absf = np.stack([fx, fy, fz], 1) + fl['o'][None]
    rows, cols, vals = [], [], []
    cid = flat_nodes(cl['n'])
    base = absf // ratio
    for k in range(8):
        off = np.array([k % 2, k // 2 % 2, k // 4])
        cpos = base + off[None]
        w = np.prod(np.clip(1 - np.abs(absf - ratio * cpos) / ratio, 0, None), axis=1)
        loc = cpos - cl['o'][None]
        ok = (w > 0) & np.all(loc >= 0, 1) & np.all(loc < cl['n'][None], 1)
        lc = loc[ok]
        ok2 = cl['vact'][lc[:, 0], lc[:, 1], lc[:, 2]]
        r = fid[ok][ok2]
        c = cid[lc[ok2, 0], lc[ok2, 1], lc[ok2, 2]]
        rows.append(r); cols.append(c); vals.append(w[ok][ok2])
    rows, cols, vals = map(np.concatenate, (rows, cols, vals))
    P1 = sp.csr_matrix((vals, (rows, cols)), shape=(int(np.prod(nf)), int(np.prod(cl['n']))))
    return sp.kron(P1, sp.identity(3), format='csr')


def penalised_element_mats(lv, KE, rho, p, fixed, rhop=None):
    """WS element matrices of Grid.cu:483-488: fixed rows/cols zeroed, 1e6 diag."""
    (ex, ey, ez), _ = element_dofs(lv)
    rp = rho_p(rho[ex, ey, ez], p) if rhop is None else rhop[ex, ey, ez].astype(np.float64)
    M = rp[:, None, None] * KE[None]
    for k in range(8):
        fk = fixed[ex + k % 2, ey + k // 2 % 2, ez + k // 4]
        for c in range(3):
            d = 3 * k + c
            M[fk, d, :] = 0
            M[fk, :, d] = 0
            M[fk, d, d] = 1e6
    return M


def stencil_from_matrix(lv, A):
    """[27][9][nn] stencil arrays from a sparse matrix over the node box."""
    n = lv['n']
    nn = int(np.prod(n))
    A = A.tocsr()
    st = np.zeros((27, 9, nn))
    x, y, z = [to_flat(a) for a in np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing='ij')]
    v = np.arange(nn)
    for nb in range(27):
        dx, dy, dz = nb % 3 - 1, nb // 3 % 3 - 1, nb // 9 - 1
        jx, jy, jz = x + dx, y + dy, z + dz
        ok = (jx >= 0) & (jy >= 0) & (jz >= 0) & (jx < n[0]) & (jy < n[1]) & (jz < n[2])
        j = jx + n[0] * (jy + n[1] * jz)
        for r in range(3):
            for q in range(3):
                vals = np.asarray(A[3 * v[ok] + r, 3 * j[ok] + q]).reshape(-1)
                st[nb, 3 * r + q, v[ok]] = vals
    return st


class Problem:
    """All operators of one hierarchy, assembled by brute force."""

    def __init__(self, act0, rho, fixed, E=1e6, nu=0.4, p=3.0, rhop=None):
        self.L = Levels(act0)
        self.KE = ke_exact(E, nu)
        self.p = p
        lv0 = self.L.levels[0]
        self.rho = rho          # [ex, ey, ez] on the level-0 box
        self.fixed = fixed & lv0['vact']
        K0 = fine_K(lv0, self.KE, rho, p, rhop)
        self.K0 = K0
        self.A_res, self.A_gs = ws_matrices(lv0, K0, self.fixed)
        Ms = penalised_element_mats(lv0, self.KE, rho, p, self.fixed, rhop)
        At = assemble(lv0, Ms)
        self.P = []
        self.A = [self.A_gs]
        A = At
        for i in range(1, len(self.L.levels)):
            P = prolongation(self.L.levels[i - 1], self.L.levels[i])
            self.P.append(P)
            A = (P.T @ A @ P).tocsr()
            self.A.append(A)
        self.active = [np.repeat(to_flat(lv['vact']), 3) for lv in self.L.levels]

    def coarse_stencil(self, i):
        return stencil_from_matrix(self.L.levels[i], self.A[i])

    def coarsest_solve(self, f):
        act = self.active[-1]
        idx = np.nonzero(act)[0]
        A = self.A[-1][idx][:, idx].toarray()
        u = np.zeros_like(f)
        u[idx] = np.linalg.solve(A, f[idx])
        return u

    def gs(self, i, u, f, n, reverse=False):
        lv = self.L.levels[i]
        col = np.repeat(colours(lv), 3)[::3]
        act = to_flat(lv['vact'])
        return point_gs(self.A[i], u, f, col, act, n, reverse)

    def residual(self, i, u, f):
        if i == 0:
            return f - self.A_res @ u
        return (f - self.A[i] @ u) * self.active[i]

    def vcycle(self, u0, f0, pre=1, post=1, symmetric=False):
        n = len(self.L.levels)
        U = [None] * n
        F = [None] * n
        U[0], F[0] = u0.copy(), f0.copy()
        for i in range(n):
            if i > 0:
                r = self.residual(i - 1, U[i - 1], F[i - 1])
                F[i] = self.P[i - 1].T @ r
                U[i] = np.zeros_like(F[i])
            if i < n - 1:
                U[i] = self.gs(i, U[i], F[i], pre)
            else:
                U[i] = self.coarsest_solve(F[i])
        for i in range(n - 2, -1, -1):
            U[i] = U[i] + (self.P[i] @ U[i + 1]) * self.active[i]
            U[i] = self.gs(i, U[i], F[i], post, reverse=symmetric)
        r = self.residual(0, U[0], F[0])
        return U[0], np.linalg.norm(r) / np.linalg.norm(F[0])


# --------------------------------------------------------------------------
# SIMP kernels
# --------------------------------------------------------------------------
def sensitivity(lv, KE, rho, p, u):
    (ex, ey, ez), dofs = element_dofs(lv)
    ue = u[dofs]
    energy = np.einsum('ei,ij,ej->e', ue, KE, ue)
    pen = (np.float32(p) * np.power(rho[ex, ey, ez].astype(np.float32), np.float32(p - 1))).astype(np.float64)
    g = np.zeros(lv['e'])
    g[ex, ey, ez] = -pen * energy
    return g


def filter_sens(lv, g, R):
    eact = lv['eact']
    Ri = int(R + 0.5)
    num = np.zeros(eact.shape)
    den = np.zeros(eact.shape)
    pad = Ri
    gp = np.pad(np.where(eact, g, 0.0), pad)
    ap = np.pad(eact, pad)
    e = eact.shape
    for dx in range(-Ri, Ri + 1):
        for dy in range(-Ri, Ri + 1):
            for dz in range(-Ri, Ri + 1):
                r2 = dx * dx + dy * dy + dz * dz
                if r2 > R * R:
                    continue
                r = np.sqrt(r2 / (R * R))
                w = 1 - 6 * r ** 2 + 8 * r ** 3 - 3 * r ** 4
                sl = (slice(pad + dx, pad + dx + e[0]), slice(pad + dy, pad + dy + e[1]), slice(pad + dz, pad + dz + e[2]))
                num += w * gp[sl]
                den += w * ap[sl]
    out = np.where(eact, num / np.where(den > 0, den, 1), 0)
    return out


def oc_update(eact, rho, g, Vgoal, step, damp, rhomin, shell=None):
    """OC bisection; shell elements (mask like eact) take density 1 in every
    trial, so they count as 1 in the volume"""
    act = eact
    gg = np.abs(np.minimum(g[act], 0)).astype(np.float32)
    ro = rho[act].astype(np.float32)
    gmax = np.float32(np.max(np.abs(g[act])))
    sh = None if shell is None else np.asarray(shell, bool)[act]

    def trial(thr):
        rn = ro * np.power(gg / np.float32(thr), np.float32(damp))
        rn = np.clip(rn, ro - np.float32(step), ro + np.float32(step))
        rn = np.clip(rn, np.float32(rhomin), np.float32(1)).astype(np.float32)
        if sh is not None:
            rn[sh] = 1
        return rn

    lo, up = np.float32(0), gmax
    itn = 0
    while True:
        thr = np.float32((lo + up) / np.float32(2))
        V = np.float32(np.sum(trial(thr).astype(np.float64)) / act.sum())
        if V > Vgoal:
            lo = thr
        elif V < Vgoal:
            up = thr
        if not (abs(V - np.float32(Vgoal)) > 1e-4 and itn < 30):
            break
        itn += 1
    out = rho.copy()
    out[act] = trial(thr)
    return out, thr
