// Host side of the MMA optimiser (MMA::mma_t, optimization/mma_t.h:159-266):
// the m-sized primal-dual variables and the (m+1) x (m+1) Newton system live
// on the host; every length-n pass runs on the device (k_mma.hip) and returns
// per-block FP64 partials that are combined here in block order.
#include "robtop_amd.h"
#include "rt_mma.h"

#include <cmath>
#include <cstring>
#include <memory>

using namespace rt;

struct rt_mma {
    int n = 0, m = 0;
    hipStream_t s = nullptr;
    void* mem = nullptr;       // all n-vectors, p and q, partials
    MmaVecs V{};
    float *p = nullptr, *q = nullptr, *dg_dev = nullptr, *df_dev = nullptr;
    double* part = nullptr;
    std::vector<double> hpart;
    int nb = 0;
    // m-sized state (mma_t.h:166-186)
    std::vector<float> y, lambda, mu, s_, a, c, d, b, lam_pq;
    float z = 1, zeta = 1, a0 = 1, eps = 1;
    int stop_counter = 0, sub_iters = 0;
    bool overflowed = false;
    ~rt_mma() {
        if (mem) (void)hipFree(mem);
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace {

// sums (or maxima) of slot k over the blocks of the last reduction
std::vector<double> fetch(rt_mma& M, int slots) {
    const size_t cnt = (size_t)M.nb * slots;
    RT_HIP(hipMemcpyAsync(M.hpart.data(), M.part, sizeof(double) * cnt, hipMemcpyDeviceToHost, M.s));
    RT_HIP(hipStreamSynchronize(M.s));
    return std::vector<double>(M.hpart.begin(), M.hpart.begin() + cnt);
}
double bsum(const std::vector<double>& h, int slots, int k, int nb) {
    double t = 0;
    for (int b = 0; b < nb; b++) t += h[(size_t)b * slots + k];
    return t;
}
double bmax(const std::vector<double>& h, int slots, int k, int nb) {
    double t = -INFINITY;
    for (int b = 0; b < nb; b++) t = std::max(t, h[(size_t)b * slots + k]);
    return t;
}

MmaLam lam_of(const std::vector<float>& v) {
    MmaLam L{};
    for (size_t i = 0; i < v.size(); i++) L.v[i] = v[i];
    return L;
}

std::vector<double> dense_solve(std::vector<double> A, std::vector<double> b, int k) {
    for (int c = 0; c < k; c++) {
        int piv = c;
        for (int r = c + 1; r < k; r++)
            if (std::fabs(A[r * k + c]) > std::fabs(A[piv * k + c])) piv = r;
        if (A[piv * k + c] == 0) throw std::domain_error("singular MMA Newton system");
        if (piv != c) {
            for (int j = 0; j < k; j++) std::swap(A[c * k + j], A[piv * k + j]);
            std::swap(b[c], b[piv]);
        }
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

struct SmallTrial {
    std::vector<float> y, lambda, mu, s;
    float z, zeta;
};

// kkt_err (mma_t.cu:595-700) at the trial point whose n part was just
// evaluated by k_mma_trial: (2-norm, inf-norm)
std::pair<float, float> kkt_from_trial(rt_mma& M, const std::vector<double>& h, const SmallTrial& t, float damp) {
    const int m = M.m, ns = mma_trial_slots(), nb = M.nb;
    double s2 = 0;
    float mx = -10.f;
    s2 += (double)(float)bsum(h, ns, 0, nb); mx = std::max(mx, (float)bmax(h, ns, 1, nb));
    double ry2 = 0; float rym = 0;
    for (int i = 0; i < m; i++) {
        const float rey = M.c[i] + M.d[i] * t.y[i] - t.lambda[i] - t.mu[i];
        ry2 += (double)rey * rey; rym = std::max(rym, std::fabs(rey));
    }
    s2 += (double)(float)ry2; mx = std::max(mx, rym);
    double la = 0;
    for (int i = 0; i < m; i++) la += (double)t.lambda[i] * M.a[i];
    const float rez = M.a0 - t.zeta - (float)la;
    s2 += (double)(rez * rez); mx = std::max(mx, std::fabs(rez));
    double rl2 = 0; float rlm = 0;
    for (int i = 0; i < m; i++) {
        const float relam = (float)bsum(h, ns, 6 + i, nb) - M.a[i] * t.z - t.y[i] + t.s[i] - M.b[i];
        rl2 += (double)relam * relam; rlm = std::max(rlm, std::fabs(relam));
    }
    s2 += (double)((float)rl2 / (damp * damp)); mx = std::max(mx, rlm / damp);
    s2 += (double)(float)bsum(h, ns, 2, nb); mx = std::max(mx, (float)bmax(h, ns, 3, nb));
    s2 += (double)(float)bsum(h, ns, 4, nb); mx = std::max(mx, (float)bmax(h, ns, 5, nb));
    double rm2 = 0; float rmm = 0;
    for (int i = 0; i < m; i++) {
        const float remu = t.mu[i] * t.y[i] - M.eps;
        rm2 += (double)remu * remu; rmm = std::max(rmm, std::fabs(remu));
    }
    s2 += (double)(float)rm2; mx = std::max(mx, rmm);
    const float rezet = t.zeta * t.z - M.eps;
    s2 += (double)(rezet * rezet); mx = std::max(mx, std::fabs(rezet));
    double rs2 = 0; float rsm = 0;
    for (int i = 0; i < m; i++) {
        const float res = t.lambda[i] * t.s[i] - M.eps;
        rs2 += (double)res * res; rsm = std::max(rsm, std::fabs(res));
    }
    s2 += (double)(float)rs2; mx = std::max(mx, rsm);
    return {(float)std::sqrt(s2), mx};
}

// mma_subproblem_t::solve (mma_t.cu:330-527) + mma_t::update (mma_t.cpp:18-51)
// dg: the m constraint gradients as one row-major [m][n] array, or (dg null)
// as m separate rows dg_rows[i] of n values, the reference's Scalar** form
bool mma_update(rt_mma& M, const float* df, const float* dg, const float* g, const float* const* dg_rows = nullptr) {
    const int n = M.n, m = M.m, nb = M.nb;
    hipStream_t s = M.s;
    RT_HIP(hipMemcpyAsync(M.df_dev, df, sizeof(float) * n, hipMemcpyDefault, s));
    if (dg)
        RT_HIP(hipMemcpyAsync(M.dg_dev, dg, sizeof(float) * (size_t)n * m, hipMemcpyDefault, s));
    else
        for (int i = 0; i < m; i++)
            RT_HIP(hipMemcpyAsync(M.dg_dev + (size_t)i * n, dg_rows[i], sizeof(float) * n, hipMemcpyDefault, s));
    std::vector<float> gv(m);
    RT_HIP(hipMemcpyAsync(gv.data(), g, sizeof(float) * m, hipMemcpyDefault, s));
    // p, q at the current design and the proxy constants b = proxy - g
    launch_mma_pq(M.V, n, m, M.df_dev, M.dg_dev, M.p, M.q, M.part, s);
    {
        const int ns = 2 * RT_MMA_MAX_M;
        const auto h = fetch(M, ns);
        for (int i = 0; i < m; i++)
            M.b[i] = ((float)bsum(h, ns, i, nb) + (float)bsum(h, ns, RT_MMA_MAX_M + i, nb)) - gv[i];
    }
    launch_mma_reset(M.V, n, s);
    M.eps = 1; M.zeta = 1; M.z = 1;
    for (int i = 0; i < m; i++) {
        M.y[i] = 1; M.lambda[i] = 1; M.s_[i] = 1;
        M.mu[i] = std::max(M.c[i] / 2.f, 1.f);
    }
    M.lam_pq = M.lambda;

    const float f1 = -1.f / 0.99f;
    float err = 1e30f, lasterr = 0;
    int nsub = 0;
    bool over = false;
    const int NS = mma_newton_slots(), MM = RT_MMA_MAX_M;
    while (M.eps > 1e-6f) {
        err = 1e30f;
        float damp = 1;
        int stays = 0;
        while (err > 0.9f * M.eps) {
            launch_mma_newton(M.V, M.p, M.q, lam_of(M.lam_pq), n, m, M.eps, M.part, s);
            const auto h = fetch(M, NS);
            std::vector<float> gproxy(m), Dy(m), hdy(m), Dly(m), hdly(m);
            double la = 0;
            for (int i = 0; i < m; i++) la += (double)M.lambda[i] * M.a[i];
            const float hdz = -((float)la + M.eps / M.z - M.a0);
            for (int i = 0; i < m; i++) {
                gproxy[i] = (float)bsum(h, NS, i, nb) + (float)bsum(h, NS, MM + i, nb);
                Dy[i] = M.d[i] + M.mu[i] / M.y[i];
                hdy[i] = -(M.lambda[i] + M.eps / M.y[i] - M.c[i] - M.d[i] * M.y[i]);
                const float hdl = -(M.z * M.a[i] + M.y[i] + M.b[i] - M.eps / M.lambda[i] - gproxy[i]);
                Dly[i] = M.s_[i] / M.lambda[i] + 1.f / Dy[i];
                hdly[i] = hdl + hdy[i] / Dy[i];
            }
            const int k = m + 1;
            std::vector<double> A(k * k, 0.0), rhs(k, 0.0);
            for (int i = 0; i < m; i++) {
                for (int i2 = 0; i2 < m; i2++) A[i * k + i2] = (double)(float)bsum(h, NS, 2 * MM + i * MM + i2, nb);
                A[i * k + i] += Dly[i];
                rhs[i] = (double)(hdly[i] - (float)bsum(h, NS, 2 * MM + MM * MM + i, nb));
                A[i * k + m] = M.a[i];
                A[m * k + i] = M.a[i];
            }
            A[m * k + m] = -(M.zeta / M.z);
            rhs[m] = hdz;
            const std::vector<double> sol = dense_solve(A, rhs, k);
            std::vector<float> dlam(m), dy(m), dmu(m), ds(m);
            const float dz = (float)sol[m];
            for (int i = 0; i < m; i++) dlam[i] = (float)sol[i];
            for (int i = 0; i < m; i++) {
                dy[i] = dlam[i] / Dy[i] - hdy[i] / Dy[i];
                dmu[i] = M.eps / M.y[i] - dy[i] * M.mu[i] / M.y[i] - M.mu[i];
                ds[i] = M.eps / M.lambda[i] - M.s_[i] / M.lambda[i] * dlam[i] - M.s_[i];
            }
            const float dzeta = M.eps / M.z - M.zeta - (M.zeta / M.z) * dz;
            launch_mma_dx(M.V, M.p, M.q, lam_of(M.lam_pq), lam_of(dlam), n, m, M.eps, M.part, s);
            const auto hd = fetch(M, 3);
            const float t1 = (float)bmax(hd, 3, 0, nb), t2 = (float)bmax(hd, 3, 1, nb);
            float t3 = (float)bmax(hd, 3, 2, nb);
            for (int i = 0; i < m; i++) {
                t3 = std::max(t3, f1 * dy[i] / M.y[i]);
                t3 = std::max(t3, f1 * dlam[i] / M.lambda[i]);
                t3 = std::max(t3, f1 * dmu[i] / M.mu[i]);
                t3 = std::max(t3, f1 * ds[i] / M.s_[i]);
            }
            t3 = std::max(t3, f1 * dz / M.z);
            t3 = std::max(t3, f1 * dzeta / M.zeta);
            float step = 1.f / std::max(std::max(t1, t2), std::max(t3, 1.f));
            // backtracking line search (search_step, mma_t.cu:529-592)
            const SmallTrial cur{M.y, M.lambda, M.mu, M.s_, M.z, M.zeta};
            launch_mma_trial(M.V, M.p, M.q, lam_of(M.lambda), n, m, M.eps, 0.f, M.part, s);
            M.lam_pq = M.lambda;
            const float olderr = kkt_from_trial(M, fetch(M, mma_trial_slots()), cur, 1.f).first;
            float newerr = 1e30f, newinf = 1e30f;
            SmallTrial t = cur;
            while (step > 1e-12f && newerr > olderr) {
                for (int i = 0; i < m; i++) {
                    t.y[i] = M.y[i] + dy[i] * step;
                    t.lambda[i] = M.lambda[i] + dlam[i] * step;
                    t.mu[i] = M.mu[i] + dmu[i] * step;
                    t.s[i] = M.s_[i] + ds[i] * step;
                }
                t.z = M.z + dz * step;
                t.zeta = M.zeta + dzeta * step;
                launch_mma_trial(M.V, M.p, M.q, lam_of(t.lambda), n, m, M.eps, step, M.part, s);
                M.lam_pq = t.lambda;
                const auto e = kkt_from_trial(M, fetch(M, mma_trial_slots()), t, damp);
                newerr = e.first;
                newinf = e.second;
                if (newerr < olderr) {
                    std::swap(M.V.x, M.V.nx);
                    std::swap(M.V.xi, M.V.nxi);
                    std::swap(M.V.eta, M.V.neta);
                    M.y = t.y; M.lambda = t.lambda; M.mu = t.mu; M.s_ = t.s; M.z = t.z; M.zeta = t.zeta;
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
    launch_mma_adjust(M.V, n, M.part, s);
    const auto h = fetch(M, 1);
    if ((float)std::sqrt(bsum(h, 1, 0, nb)) < 1e-3f) M.stop_counter++;
    return M.stop_counter > 10;
}

rt_mma& MM_(rt_mma* h) {
    if (!h) throw std::invalid_argument("null handle");
    return *h;
}

}  // namespace

#define RT_MMA_TRY(...)                          \
    try {                                        \
        __VA_ARGS__;                             \
        return RT_OK;                            \
    } catch (...) {                              \
        return status_of_current_exception();    \
    }

extern "C" {

int rt_mma_create(int n, int m, const float* lower, const float* upper, rt_mma** out) {
    RT_MMA_TRY({
        if (!out || !lower || !upper) throw std::invalid_argument("null pointer argument");
        if (n < 1 || m < 1 || m > RT_MMA_MAX_M) throw std::invalid_argument("need n >= 1 and 1 <= m <= 8");
        std::unique_ptr<rt_mma> M(new rt_mma());
        M->n = n; M->m = m;
        M->nb = mma_blocks(n);
        RT_HIP(hipStreamCreate(&M->s));
        const size_t nv = 15, fl = (size_t)n * (nv + 2 * (m + 1) + m + 1);
        const size_t pslots = (size_t)M->nb * std::max(mma_newton_slots(), mma_trial_slots());
        const size_t bytes = sizeof(float) * fl + 256 + sizeof(double) * pslots;
        RT_HIP(hipMalloc(&M->mem, bytes));
        RT_HIP(hipMemset(M->mem, 0, bytes));
        float* f = (float*)M->mem;
        float** vs[15] = {&M->V.x, &M->V.xmin, &M->V.xmax, &M->V.al, &M->V.au, &M->V.alpha, &M->V.beta,
                          &M->V.lastdx, &M->V.xi, &M->V.eta, &M->V.curx, &M->V.dx, &M->V.nx, &M->V.nxi,
                          &M->V.neta};
        for (int i = 0; i < 15; i++) { *vs[i] = f; f += n; }
        M->p = f; f += (size_t)n * (m + 1);
        M->q = f; f += (size_t)n * (m + 1);
        M->dg_dev = f; f += (size_t)n * m;
        M->df_dev = f; f += n;
        M->part = (double*)(((uintptr_t)f + 255) & ~(uintptr_t)255);
        M->hpart.resize(pslots);
        M->y.assign(m, 1.f); M->lambda.assign(m, 1.f); M->mu.assign(m, 0.f); M->s_.assign(m, 1.f);
        M->a.assign(m, 0.f); M->c.assign(m, 10.f); M->d.assign(m, 1.f); M->b.assign(m, 0.f);
        for (int i = 0; i < m; i++) M->mu[i] = std::max(M->c[i] / 2.f, 1.f);
        M->lam_pq = M->lambda;
        // bounds may be host or device arrays: stage them in dx / nx
        RT_HIP(hipMemcpyAsync(M->V.dx, lower, sizeof(float) * n, hipMemcpyDefault, M->s));
        RT_HIP(hipMemcpyAsync(M->V.nx, upper, sizeof(float) * n, hipMemcpyDefault, M->s));
        launch_mma_init(M->V, n, M->V.dx, M->V.nx, M->s);
        RT_HIP(hipStreamSynchronize(M->s));
        *out = M.release();
    })
}

int rt_mma_destroy(rt_mma* h) {
    RT_MMA_TRY({ delete h; })
}

int rt_mma_update(rt_mma* h, const float* df, const float* dg, const float* g, int* converged) {
    RT_MMA_TRY({
        rt_mma& M = MM_(h);
        if (!df || !dg || !g) throw std::invalid_argument("null pointer argument");
        const bool c = mma_update(M, df, dg, g);
        if (converged) *converged = c ? 1 : 0;
    })
}

int rt_mma_update_rows(rt_mma* h, const float* df, const float* const* dg_rows, const float* g, int* converged) {
    RT_MMA_TRY({
        rt_mma& M = MM_(h);
        if (!df || !dg_rows || !g) throw std::invalid_argument("null pointer argument");
        for (int i = 0; i < M.m; i++)
            if (!dg_rows[i]) throw std::invalid_argument("null constraint-gradient row");
        const bool c = mma_update(M, df, nullptr, g, dg_rows);
        if (converged) *converged = c ? 1 : 0;
    })
}

int rt_mma_get_x(const rt_mma* h, float* x) {
    RT_MMA_TRY({
        rt_mma& M = MM_(const_cast<rt_mma*>(h));
        if (!x) throw std::invalid_argument("null pointer argument");
        RT_HIP(hipMemcpyAsync(x, M.V.x, sizeof(float) * M.n, hipMemcpyDefault, M.s));
        RT_HIP(hipStreamSynchronize(M.s));
    })
}

int rt_mma_stats(const rt_mma* h, float* lambda, float* z, int* sub_iters, int* overflowed) {
    RT_MMA_TRY({
        const rt_mma& M = MM_(const_cast<rt_mma*>(h));
        if (lambda) std::memcpy(lambda, M.lambda.data(), sizeof(float) * M.m);
        if (z) *z = M.z;
        if (sub_iters) *sub_iters = M.sub_iters;
        if (overflowed) *overflowed = M.overflowed ? 1 : 0;
    })
}

}  // extern "C"
