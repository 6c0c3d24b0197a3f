// ORACLE — test infrastructure only (see oracle.cpp).
//
// CPU restatement of the reference's MMA update (optimization/mma_t.h:159-266,
// mma_t.cpp:18-51, mma_t.cu): Svanberg's method of moving asymptotes with the
// subproblem solved by a primal-dual interior-point iteration on the reduced
// (m+1) x (m+1) Newton system, backtracking on the 2-norm of the KKT residual,
// and epsilon lowered by 10 down to 1e-6.
//
// Vectors are FP32 as in the reference (gv::Scalar = float); sums and dot
// products accumulate in FP64 and are rounded to FP32 (the reference reduces
// in FP32 trees of unspecified order). (u - x)^3 is formed as a product of
// three factors (the reference calls powf(., 3)).
//
// Pinning: tests/test_mma.py checks the converged designs of the reference's
// own test problems (mma_t.cpp:167-311) against scipy's SLSQP and a KKT check,
// and a SIMP compliance step against a bisection of the same separable
// subproblem.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace {

typedef float S;

struct Mma {
    int n = 0, m = 0;
    // n-vectors
    std::vector<S> x, xmin, xmax, al, au, alpha, beta, lastdx, xi, eta;
    std::vector<std::vector<S>> p, q;   // [m + 1][n]
    // m-vectors
    std::vector<S> y, lambda, mu, s, a, c, d, b, lam_pq;
    S z = 1, zeta = 1, a0 = 1, eps = 1;
    int stop_counter = 0;
    // diagnostics of the last update
    int sub_iters = 0;
    bool overflowed = false;
};

// clamp_asymptotes (mma_t.cu:803-816) after set_bound (mma_t.cu:842-866)
void clamp_asym(Mma& M) {
    for (int j = 0; j < M.n; j++) {
        const S st = M.xmax[j] - M.xmin[j], x = M.x[j];
        M.al[j] = std::min(std::max(M.al[j], x - 10.f * st), x - 0.01f * st);
        M.au[j] = std::min(std::max(M.au[j], x + 0.01f * st), x + 10.f * st);
        M.alpha[j] = std::max(M.xmin[j], std::max(M.al[j] + 0.1f * (x - M.al[j]), x - 0.5f * st));
        M.beta[j] = std::min(M.xmax[j], std::min(M.au[j] - 0.1f * (M.au[j] - x), x + 0.5f * st));
    }
}

// mma_t constructor (mma_t.h:202-213) + init (:218-221)
void init(Mma& M, int n, int m, const S* lo, const S* hi) {
    M.n = n; M.m = m;
    M.xmin.assign(lo, lo + n); M.xmax.assign(hi, hi + n);
    M.alpha = M.xmin; M.beta = M.xmax; M.al = M.alpha; M.au = M.beta;
    M.x.resize(n);
    for (int j = 0; j < n; j++) M.x[j] = (M.alpha[j] + M.beta[j]) / 2.f;
    clamp_asym(M);
    M.lastdx.assign(n, 0.f);
    M.y.assign(m, 1.f); M.lambda.assign(m, 1.f); M.mu.assign(m, 0.f); M.s.assign(m, 1.f);
    M.a.assign(m, 0.f); M.c.assign(m, 10.f); M.d.assign(m, 1.f); M.b.assign(m, 0.f);
    M.z = 1; M.zeta = 1; M.a0 = 1; M.eps = 1;
    // init_subproblem_variable (mma_t.cu:830-840)
    M.xi.resize(n); M.eta.resize(n);
    for (int j = 0; j < n; j++) {
        M.x[j] = (M.alpha[j] + M.beta[j]) / 2.f;
        M.xi[j] = std::max(1.f / (M.x[j] - M.alpha[j]), 1.f);
        M.eta[j] = std::max(1.f / (M.beta[j] - M.x[j]), 1.f);
    }
    for (int i = 0; i < m; i++) M.mu[i] = std::max(M.c[i] / 2.f, 1.f);
    M.p.assign(m + 1, std::vector<S>(n)); M.q.assign(m + 1, std::vector<S>(n));
}

// p = p0 + sum lam_i p_{i+1} (update_pqlambda, mma_t.cu:760-772), one element
inline S plam(const Mma& M, const std::vector<S>& lam, int j) {
    S v = M.p[0][j];
    for (int i = 0; i < M.m; i++) v += lam[i] * M.p[i + 1][j];
    return v;
}
inline S qlam(const Mma& M, const std::vector<S>& lam, int j) {
    S v = M.q[0][j];
    for (int i = 0; i < M.m; i++) v += lam[i] * M.q[i + 1][j];
    return v;
}

struct Trial {
    std::vector<S> x, xi, eta, y, lambda, mu, s;
    S z, zeta;
};

// kkt_err (mma_t.cu:595-700): (2-norm, inf-norm) of the perturbed KKT residual;
// the relaxed-constraint block is divided by lambda_damp
std::pair<S, S> kkt_err(Mma& M, const Trial& t, S damp) {
    M.lam_pq = t.lambda;
    const int n = M.n, m = M.m;
    double s2 = 0;
    S mx = -10.f;
    std::vector<double> g(m, 0.0);
    double rx2 = 0, rxs2 = 0, re2 = 0;
    S rxm = 0, rxsm = 0, rem = 0;
    for (int j = 0; j < n; j++) {
        const S ux = M.au[j] - t.x[j], xl = t.x[j] - M.al[j];
        const S rex = plam(M, M.lam_pq, j) / (ux * ux) - qlam(M, M.lam_pq, j) / (xl * xl) - t.xi[j] + t.eta[j];
        rx2 += (double)rex * rex; rxm = std::max(rxm, std::fabs(rex));
        const S rxsi = t.xi[j] * (t.x[j] - M.alpha[j]) - M.eps;
        rxs2 += (double)rxsi * rxsi; rxsm = std::max(rxsm, std::fabs(rxsi));
        const S reeta = t.eta[j] * (M.beta[j] - t.x[j]) - M.eps;
        re2 += (double)reeta * reeta; rem = std::max(rem, std::fabs(reeta));
        for (int i = 0; i < m; i++) g[i] += (double)(M.p[i + 1][j] / ux + M.q[i + 1][j] / xl);
    }
    s2 += (double)(S)rx2; mx = std::max(mx, rxm);
    double ry2 = 0; S rym = 0;
    for (int i = 0; i < m; i++) {
        const S rey = M.c[i] + M.d[i] * t.y[i] - t.lambda[i] - t.mu[i];
        ry2 += (double)rey * rey; rym = std::max(rym, std::fabs(rey));
    }
    s2 += (double)(S)ry2; mx = std::max(mx, rym);
    double la = 0;
    for (int i = 0; i < m; i++) la += (double)t.lambda[i] * M.a[i];
    const S rez = M.a0 - t.zeta - (S)la;
    s2 += (double)(rez * rez); mx = std::max(mx, std::fabs(rez));
    double rl2 = 0; S rlm = 0;
    for (int i = 0; i < m; i++) {
        const S relam = (S)g[i] - M.a[i] * t.z - t.y[i] + t.s[i] - M.b[i];
        rl2 += (double)relam * relam; rlm = std::max(rlm, std::fabs(relam));
    }
    s2 += (double)((S)rl2 / (damp * damp)); mx = std::max(mx, rlm / damp);
    s2 += (double)(S)rxs2; mx = std::max(mx, rxsm);
    s2 += (double)(S)re2; mx = std::max(mx, rem);
    double rm2 = 0; S rmm = 0;
    for (int i = 0; i < m; i++) {
        const S remu = t.mu[i] * t.y[i] - M.eps;
        rm2 += (double)remu * remu; rmm = std::max(rmm, std::fabs(remu));
    }
    s2 += (double)(S)rm2; mx = std::max(mx, rmm);
    const S rezet = t.zeta * t.z - M.eps;
    s2 += (double)(rezet * rezet); mx = std::max(mx, std::fabs(rezet));
    double rs2 = 0; S rsm = 0;
    for (int i = 0; i < m; i++) {
        const S res = t.lambda[i] * t.s[i] - M.eps;
        rs2 += (double)res * res; rsm = std::max(rsm, std::fabs(res));
    }
    s2 += (double)(S)rs2; mx = std::max(mx, rsm);
    return {(S)std::sqrt(s2), mx};
}

// dense solve of the (m+1) system, Gaussian elimination with partial pivoting
// in FP64 (the reference uses a column-pivoted QR, mma_t.cpp:94-128)
std::vector<double> dense_solve(std::vector<double> A, std::vector<double> b, int k) {
    for (int c = 0; c < k; c++) {
        int piv = c;
        for (int r = c + 1; r < k; r++) if (std::fabs(A[r * k + c]) > std::fabs(A[piv * k + c])) piv = r;
        if (A[piv * k + c] == 0) throw std::runtime_error("singular MMA Newton system");
        if (piv != c) { for (int j = 0; j < k; j++) std::swap(A[c * k + j], A[piv * k + j]); std::swap(b[c], b[piv]); }
        for (int r = c + 1; r < k; r++) {
            const double f = A[r * k + c] / A[c * k + c];
            for (int j = c; j < k; j++) A[r * k + j] -= f * A[c * k + j];
            b[r] -= f * b[c];
        }
    }
    std::vector<double> x(k);
    for (int r = k - 1; r >= 0; r--) {
        double t = b[r];
        for (int j = r + 1; j < k; j++) t -= A[r * k + j] * x[j];
        x[r] = t / A[r * k + r];
    }
    return x;
}

// mma_subproblem_t::solve (mma_t.cu:330-527) + mma_t::update (mma_t.cpp:18-51)
bool update(Mma& M, const S* df, const S* const* dg, const S* gval) {
    const int n = M.n, m = M.m;
    const std::vector<S> curx = M.x;
    // p, q and the proxy's constant (mma_t.cu:69-90), at the current design
    for (int i = 0; i <= m; i++) {
        const S* D = i == 0 ? df : dg[i - 1];
        for (int j = 0; j < n; j++) {
            const S ux = M.au[j] - M.x[j], xl = M.x[j] - M.al[j];
            const S dp = std::max(D[j], 0.f), dn = std::max(-D[j], 0.f), r = 1e-5f / (M.xmax[j] - M.xmin[j]);
            M.p[i][j] = (ux * ux) * (1.001f * dp + 0.001f * dn + r);
            M.q[i][j] = (xl * xl) * (0.001f * dp + 1.001f * dn + r);
        }
    }
    for (int i = 0; i < m; i++) {
        double a1 = 0, a2 = 0;
        for (int j = 0; j < n; j++) {
            a1 += (double)(M.p[i + 1][j] * (1.f / (M.au[j] - M.x[j])));
            a2 += (double)(M.q[i + 1][j] * (1.f / (M.x[j] - M.al[j])));
        }
        M.b[i] = ((S)a1 + (S)a2) - gval[i];
    }
    // reset (mma_t.cu:775-786)
    for (int j = 0; j < n; j++) {
        M.x[j] = (M.alpha[j] + M.beta[j]) / 2.f;
        M.xi[j] = std::max(1.f / (M.x[j] - M.alpha[j]), 1.f);
        M.eta[j] = std::max(1.f / (M.beta[j] - M.x[j]), 1.f);
    }
    M.eps = 1; M.zeta = 1; M.z = 1;
    for (int i = 0; i < m; i++) { M.y[i] = 1; M.lambda[i] = 1; M.s[i] = 1; M.mu[i] = std::max(M.c[i] / 2.f, 1.f); }
    M.lam_pq = M.lambda;

    S err = 1e30f, lasterr = 0;
    int nsub = 0;
    bool over = false;
    std::vector<S> Dx(n), hdx(n), dx(n);
    std::vector<std::vector<S>> G(m, std::vector<S>(n));
    while (M.eps > 1e-6f) {
        err = 1e30f;
        S damp = 1;
        int stays = 0;
        while (err > 0.9f * M.eps) {
            // Newton system (mma_t.cu:371-395, solveLinearSystem :232-328)
            std::vector<double> gp(m, 0.0);
            for (int j = 0; j < n; j++) {
                const S ux = M.au[j] - M.x[j], xl = M.x[j] - M.al[j];
                const S pl = plam(M, M.lam_pq, j), ql = qlam(M, M.lam_pq, j);
                const S psi = 2.f * pl / (ux * ux * ux) + 2.f * ql / (xl * xl * xl);
                Dx[j] = psi + M.xi[j] / (M.x[j] - M.alpha[j]) + M.eta[j] / (M.beta[j] - M.x[j]);
                const S varphi = pl / (ux * ux) - ql / (xl * xl);
                hdx[j] = -(M.eps / (M.x[j] - M.alpha[j]) - M.eps / (M.beta[j] - M.x[j]) - varphi);
                for (int i = 0; i < m; i++) {
                    G[i][j] = M.p[i + 1][j] / (ux * ux) - M.q[i + 1][j] / (xl * xl);
                    gp[i] += (double)(M.p[i + 1][j] * (1.f / ux));
                }
            }
            std::vector<S> gproxy(m);
            for (int i = 0; i < m; i++) {
                double a2 = 0;
                for (int j = 0; j < n; j++) a2 += (double)(M.q[i + 1][j] * (1.f / (M.x[j] - M.al[j])));
                gproxy[i] = (S)gp[i] + (S)a2;
            }
            std::vector<S> Dy(m), hdy(m), hdl(m), Dly(m), hdly(m);
            double la = 0;
            for (int i = 0; i < m; i++) la += (double)M.lambda[i] * M.a[i];
            const S hdz = -((S)la + M.eps / M.z - M.a0);
            for (int i = 0; i < m; i++) {
                Dy[i] = M.d[i] + M.mu[i] / M.y[i];
                hdy[i] = -(M.lambda[i] + M.eps / M.y[i] - M.c[i] - M.d[i] * M.y[i]);
                hdl[i] = -(M.z * M.a[i] + M.y[i] + M.b[i] - M.eps / M.lambda[i] - gproxy[i]);
                Dly[i] = M.s[i] / M.lambda[i] + 1.f / Dy[i];
                hdly[i] = hdl[i] + hdy[i] / Dy[i];
            }
            const int k = m + 1;
            std::vector<double> A(k * k, 0.0), rhs(k, 0.0);
            for (int i = 0; i < m; i++) {
                for (int i2 = 0; i2 < m; i2++) {
                    double t = 0;
                    for (int j = 0; j < n; j++) t += (double)(G[i][j] * (G[i2][j] / Dx[j]));
                    A[i * k + i2] = (double)(S)t;
                }
                A[i * k + i] += Dly[i];
                double t = 0;
                for (int j = 0; j < n; j++) t += (double)(G[i][j] * (hdx[j] / Dx[j]));
                rhs[i] = (double)(hdly[i] - (S)t);
                A[i * k + m] = M.a[i];
                A[m * k + i] = M.a[i];
            }
            A[m * k + m] = -(M.zeta / M.z);
            rhs[m] = hdz;
            const std::vector<double> sol = dense_solve(A, rhs, k);
            std::vector<S> dlam(m), dy(m);
            const S dz = (S)sol[m];
            for (int i = 0; i < m; i++) dlam[i] = (S)sol[i];
            for (int j = 0; j < n; j++) {
                S v = -hdx[j] / Dx[j];
                for (int i = 0; i < m; i++) v -= dlam[i] * G[i][j] / Dx[j];
                dx[j] = v;
            }
            for (int i = 0; i < m; i++) dy[i] = dlam[i] / Dy[i] - hdy[i] / Dy[i];
            // directions of the multipliers (mma_t.cu:311-315 and :546-550)
            std::vector<S> dmu(m), ds(m);
            for (int i = 0; i < m; i++) {
                dmu[i] = M.eps / M.y[i] - dy[i] * M.mu[i] / M.y[i] - M.mu[i];
                ds[i] = M.eps / M.lambda[i] - M.s[i] / M.lambda[i] * dlam[i] - M.s[i];
            }
            const S dzeta = M.eps / M.z - M.zeta - (M.zeta / M.z) * dz;
            // initial step: keep every slack positive with factor 0.99 (mma_t.cu:412-416)
            const S f1 = -1.f / 0.99f, f2 = 1.f / 0.99f;
            S t1 = -INFINITY, t2 = -INFINITY, t3 = -INFINITY;
            for (int j = 0; j < n; j++) {
                t1 = std::max(t1, f1 * (dx[j] / (M.x[j] - M.alpha[j])));
                t2 = std::max(t2, f2 * (dx[j] / (M.beta[j] - M.x[j])));
                const S dxi = M.eps / (M.x[j] - M.alpha[j]) - dx[j] / (M.x[j] - M.alpha[j]) * M.xi[j] - M.xi[j];
                const S deta = M.eps / (M.beta[j] - M.x[j]) + M.eta[j] / (M.beta[j] - M.x[j]) * dx[j] - M.eta[j];
                t3 = std::max(t3, std::max(f1 * dxi / M.xi[j], f1 * deta / M.eta[j]));
            }
            for (int i = 0; i < m; i++) {
                t3 = std::max(t3, f1 * dy[i] / M.y[i]);
                t3 = std::max(t3, f1 * dlam[i] / M.lambda[i]);
                t3 = std::max(t3, f1 * dmu[i] / M.mu[i]);
                t3 = std::max(t3, f1 * ds[i] / M.s[i]);
            }
            t3 = std::max(t3, f1 * dz / M.z);
            t3 = std::max(t3, f1 * dzeta / M.zeta);
            S step = 1.f / std::max(std::max(t1, t2), std::max(t3, 1.f));
            // backtracking (search_step, mma_t.cu:529-592)
            Trial cur{M.x, M.xi, M.eta, M.y, M.lambda, M.mu, M.s, M.z, M.zeta};
            const S olderr = kkt_err(M, cur, 1.f).first;
            S newerr = 1e30f, newinf = 1e30f;
            Trial t = cur;
            while (step > 1e-12f && newerr > olderr) {
                for (int j = 0; j < n; j++) {
                    const S dxi = M.eps / (M.x[j] - M.alpha[j]) - dx[j] / (M.x[j] - M.alpha[j]) * M.xi[j] - M.xi[j];
                    const S deta = M.eps / (M.beta[j] - M.x[j]) + M.eta[j] / (M.beta[j] - M.x[j]) * dx[j] - M.eta[j];
                    t.x[j] = M.x[j] + dx[j] * step;
                    t.xi[j] = M.xi[j] + dxi * step;
                    t.eta[j] = M.eta[j] + deta * step;
                }
                for (int i = 0; i < m; i++) {
                    t.y[i] = M.y[i] + dy[i] * step;
                    t.lambda[i] = M.lambda[i] + dlam[i] * step;
                    t.mu[i] = M.mu[i] + dmu[i] * step;
                    t.s[i] = M.s[i] + ds[i] * step;
                }
                t.z = M.z + dz * step;
                t.zeta = M.zeta + dzeta * step;
                const auto e = kkt_err(M, t, damp);
                newerr = e.first; newinf = e.second;
                if (newerr < olderr) {
                    M.x = t.x; M.xi = t.xi; M.eta = t.eta; M.y = t.y; M.lambda = t.lambda; M.mu = t.mu; M.s = t.s;
                    M.z = t.z; M.zeta = t.zeta;
                    break;
                }
                step /= 2.f;
            }
            err = newinf;
            if ((err - lasterr) / err < 1e-5f) stays++; else stays = 0;
            lasterr = err;
            if (stays > 6) { stays = 0; damp *= 10.f; }
            if (nsub++ > 80) { over = true; break; }
        }
        if (over) break;
        M.eps *= 0.1f;
    }
    M.sub_iters = nsub;
    M.overflowed = over;
    // adjust_asym (mma_t.cu:161-226) with this and the previous design change
    double nrm = 0;
    for (int j = 0; j < n; j++) {
        const S d1 = M.x[j] - curx[j], d2 = M.lastdx[j];
        S gamma = 1.f;
        if (d1 * d2 > 0) gamma = 1.2f;
        if (d1 * d2 < 0) gamma = 0.7f;
        const S x = M.x[j], st = M.xmax[j] - M.xmin[j];
        S nal = x - gamma * (x - d1 - M.al[j]);
        S nau = x + gamma * (M.au[j] - x + d1);
        nal = std::min(std::max(nal, x - 10.f * st), x - 0.01f * st);
        nau = std::min(std::max(nau, x + 0.01f * st), x + 10.f * st);
        M.al[j] = nal; M.au[j] = nau;
        S na = M.xmin[j];
        if (na < nal + 0.1f * (x - nal)) na = nal + 0.1f * (x - nal);
        if (na < x - 0.5f * st) na = x - 0.5f * st;
        S nb = M.xmax[j];
        if (nb > nau - 0.1f * (nau - x)) nb = nau - 0.1f * (nau - x);
        if (nb > x + 0.5f * st) nb = x + 0.5f * st;
        M.alpha[j] = na; M.beta[j] = nb;
        M.lastdx[j] = d1;
        nrm += (double)d1 * d1;
    }
    if ((S)std::sqrt(nrm) < 1e-3f) M.stop_counter++;
    return M.stop_counter > 10;
}

thread_local std::string g_mma_err;

}  // namespace

extern "C" {

void* or_mma_create(int n, int m, const float* lo, const float* hi) {
    Mma* M = new Mma();
    init(*M, n, m, lo, hi);
    return M;
}
void or_mma_destroy(void* h) { delete (Mma*)h; }
// returns 1 when converged, 0 otherwise, -1 on error
int or_mma_update(void* h, const float* df, const float* dg, const float* g) {
    Mma* M = (Mma*)h;
    std::vector<const float*> rows(M->m);
    for (int i = 0; i < M->m; i++) rows[i] = dg + (size_t)i * M->n;
    try { return update(*M, df, rows.data(), g) ? 1 : 0; }
    catch (std::exception& e) { g_mma_err = e.what(); return -1; }
}
void or_mma_get_x(void* h, float* x) { Mma* M = (Mma*)h; std::memcpy(x, M->x.data(), sizeof(float) * M->n); }
void or_mma_set_x(void* h, const float* x) { Mma* M = (Mma*)h; std::memcpy(M->x.data(), x, sizeof(float) * M->n); }
// {alpha, beta, asym_l, asym_u} (4 x n)
void or_mma_get_bounds(void* h, float* out) {
    Mma* M = (Mma*)h;
    const size_t n = M->n;
    std::memcpy(out, M->alpha.data(), 4 * n); std::memcpy(out + n, M->beta.data(), 4 * n);
    std::memcpy(out + 2 * n, M->al.data(), 4 * n); std::memcpy(out + 3 * n, M->au.data(), 4 * n);
}
// lambda (m), z, sub-iterations, overflow flag
void or_mma_stats(void* h, float* lambda, float* z, int* sub_iters, int* overflowed) {
    Mma* M = (Mma*)h;
    std::memcpy(lambda, M->lambda.data(), sizeof(float) * M->m);
    *z = M->z; *sub_iters = M->sub_iters; *overflowed = M->overflowed ? 1 : 0;
}
const char* or_mma_last_error(void) { return g_mma_err.c_str(); }

}
