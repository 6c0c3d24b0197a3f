// MMA (method of moving asymptotes) kernels: every length-n pass of one
// update, each fused with the reductions that follow it. Per-block results go
// to `part[block][slot]` (FP64) and are combined on the host in block order,
// so the update is deterministic. Vectors are FP32; m <= RT_MMA_MAX_M.
#include "rt_mma.h"

namespace rt {

namespace {

constexpr float kF1 = -1.f / 0.99f, kF2 = 1.f / 0.99f;   // asym_clamp_factor (mma_t.h:21)

// block reduction of K values per thread into part[blockIdx.x * K + k]; MAX
// selects max instead of sum for each slot
template <int K>
__device__ void block_reduce(double (&v)[K], const bool (&is_max)[K], double* part) {
    __shared__ double red[K][4];
#pragma unroll
    for (int k = 0; k < K; k++) {
        double a = v[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double b = __shfl_down(a, o, 64);
            a = is_max[k] ? fmax(a, b) : a + b;
        }
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        double a = red[k][0];
        for (int w = 1; w < 4; w++) a = is_max[k] ? fmax(a, red[k][w]) : a + red[k][w];
        part[blockIdx.x * K + k] = a;
    }
}

struct PQ {
    const float* p;   // [(m + 1) * n]
    const float* q;
};

__device__ __forceinline__ float plam(const PQ& P, const MmaLam& L, int m, int n, int j) {
    float v = P.p[j];
    for (int i = 0; i < m; i++) v += L.v[i] * P.p[(size_t)(i + 1) * n + j];
    return v;
}
__device__ __forceinline__ float qlam(const PQ& P, const MmaLam& L, int m, int n, int j) {
    float v = P.q[j];
    for (int i = 0; i < m; i++) v += L.v[i] * P.q[(size_t)(i + 1) * n + j];
    return v;
}

}  // namespace

// p, q of the convex approximation at the current design (mma_t.cu:69-77) and
// the constant of each constraint proxy: sums of p_{i+1} / (u - x) and
// q_{i+1} / (x - l) (mma_t.cu:80-87); x is saved as the update's start point
__global__ void __launch_bounds__(256) k_mma_pq(MmaVecs V, int n, int m, const float* __restrict__ df,
                                                const float* __restrict__ dg, float* __restrict__ p,
                                                float* __restrict__ q, double* __restrict__ part) {
    double acc[2 * RT_MMA_MAX_M];
    bool mx[2 * RT_MMA_MAX_M];
#pragma unroll
    for (int k = 0; k < 2 * RT_MMA_MAX_M; k++) { acc[k] = 0; mx[k] = false; }
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const float x = V.x[j];
        V.curx[j] = x;
        const float ux = V.au[j] - x, xl = x - V.al[j];
        const float r = 1e-5f / (V.xmax[j] - V.xmin[j]);
        for (int i = 0; i <= m; i++) {
            const float D = i == 0 ? df[j] : dg[(size_t)(i - 1) * n + j];
            const float dp = fmaxf(D, 0.f), dn = fmaxf(-D, 0.f);
            const float pv = (ux * ux) * (1.001f * dp + 0.001f * dn + r);
            const float qv = (xl * xl) * (0.001f * dp + 1.001f * dn + r);
            p[(size_t)i * n + j] = pv;
            q[(size_t)i * n + j] = qv;
            if (i > 0) {
                acc[i - 1] += (double)(pv * (1.f / ux));
                acc[RT_MMA_MAX_M + i - 1] += (double)(qv * (1.f / xl));
            }
        }
    }
    block_reduce<2 * RT_MMA_MAX_M>(acc, mx, part);
}

// reset of the subproblem's primal-dual point (mma_t.cu:775-786), n part
__global__ void __launch_bounds__(256) k_mma_reset(MmaVecs V, int n) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const float x = (V.alpha[j] + V.beta[j]) / 2.f;
        V.x[j] = x;
        V.xi[j] = fmaxf(1.f / (x - V.alpha[j]), 1.f);
        V.eta[j] = fmaxf(1.f / (V.beta[j] - x), 1.f);
    }
}

// element terms of the Newton system (mma_t.cu:371-395 and :232-328):
// Dx = Psi + xi/(x-alpha) + eta/(beta-x), hat_dx, G_i; sums of the proxy
// terms, of G_i G_k / Dx and of G_i hat_dx / Dx
struct NewtonTerms {
    float Dx, hdx, G[RT_MMA_MAX_M];
};
__device__ __forceinline__ NewtonTerms newton_terms(const MmaVecs& V, const PQ& P, const MmaLam& L, int m, int n,
                                                    int j, float eps) {
    NewtonTerms t;
    const float x = V.x[j];
    const float ux = V.au[j] - x, xl = x - V.al[j];
    const float pl = plam(P, L, m, n, j), ql = qlam(P, L, m, n, j);
    const float psi = 2.f * pl / (ux * ux * ux) + 2.f * ql / (xl * xl * xl);
    t.Dx = psi + V.xi[j] / (x - V.alpha[j]) + V.eta[j] / (V.beta[j] - x);
    const float varphi = pl / (ux * ux) - ql / (xl * xl);
    t.hdx = -(eps / (x - V.alpha[j]) - eps / (V.beta[j] - x) - varphi);
    for (int i = 0; i < m; i++)
        t.G[i] = P.p[(size_t)(i + 1) * n + j] / (ux * ux) - P.q[(size_t)(i + 1) * n + j] / (xl * xl);
    return t;
}

// slots: [0, m) sum p/(u-x); [M, M+m) sum q/(x-l); [2M, 2M+m*m) G_i (G_k/Dx);
// [2M+M*M, 2M+M*M+m) G_i (hat_dx/Dx)
constexpr int kNewtonSlots = 2 * RT_MMA_MAX_M + RT_MMA_MAX_M * RT_MMA_MAX_M + RT_MMA_MAX_M;
__global__ void __launch_bounds__(256) k_mma_newton(MmaVecs V, const float* __restrict__ p,
                                                    const float* __restrict__ q, MmaLam L, int n, int m, float eps,
                                                    double* __restrict__ part) {
    double acc[kNewtonSlots];
    bool mx[kNewtonSlots];
#pragma unroll
    for (int k = 0; k < kNewtonSlots; k++) { acc[k] = 0; mx[k] = false; }
    const PQ P{p, q};
    constexpr int M = RT_MMA_MAX_M;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const NewtonTerms t = newton_terms(V, P, L, m, n, j, eps);
        const float x = V.x[j];
        const float ux = V.au[j] - x, xl = x - V.al[j];
        for (int i = 0; i < m; i++) {
            acc[i] += (double)(p[(size_t)(i + 1) * n + j] * (1.f / ux));
            acc[M + i] += (double)(q[(size_t)(i + 1) * n + j] * (1.f / xl));
            for (int k = 0; k < m; k++) acc[2 * M + i * M + k] += (double)(t.G[i] * (t.G[k] / t.Dx));
            acc[2 * M + M * M + i] += (double)(t.G[i] * (t.hdx / t.Dx));
        }
    }
    block_reduce<kNewtonSlots>(acc, mx, part);
}

// Newton direction of x and the largest admissible step factors
// (mma_t.cu:293-297, :412-416): slot 0 max f1 dx/(x-alpha), 1 max f2
// dx/(beta-x), 2 max of f1 dxi/xi and f1 deta/eta
__global__ void __launch_bounds__(256) k_mma_dx(MmaVecs V, const float* __restrict__ p, const float* __restrict__ q,
                                                MmaLam L, MmaLam dlam, int n, int m, float eps,
                                                double* __restrict__ part) {
    double acc[3] = {-INFINITY, -INFINITY, -INFINITY};
    const bool mx[3] = {true, true, true};
    const PQ P{p, q};
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const NewtonTerms t = newton_terms(V, P, L, m, n, j, eps);
        float d = -t.hdx / t.Dx;
        for (int i = 0; i < m; i++) d -= dlam.v[i] * t.G[i] / t.Dx;
        V.dx[j] = d;
        const float x = V.x[j], xa = x - V.alpha[j], bx = V.beta[j] - x;
        const float dxi = eps / xa - d / xa * V.xi[j] - V.xi[j];
        const float deta = eps / bx + V.eta[j] / bx * d - V.eta[j];
        acc[0] = fmax(acc[0], (double)(kF1 * (d / xa)));
        acc[1] = fmax(acc[1], (double)(kF2 * (d / bx)));
        acc[2] = fmax(acc[2], fmax((double)(kF1 * dxi / V.xi[j]), (double)(kF1 * deta / V.eta[j])));
    }
    block_reduce<3>(acc, mx, part);
}

// one trial step of the line search (mma_t.cu:562-575) and the n part of the
// KKT residual at the trial point (mma_t.cu:605-687): slots 0 |rex|^2, 1
// max|rex|, 2 |rexsi|^2, 3 max|rexsi|, 4 |reeta|^2, 5 max|reeta|, 6.. g_i
constexpr int kTrialSlots = 6 + RT_MMA_MAX_M;
__global__ void __launch_bounds__(256) k_mma_trial(MmaVecs V, const float* __restrict__ p,
                                                   const float* __restrict__ q, MmaLam Lnew, int n, int m, float eps,
                                                   float step, double* __restrict__ part) {
    double acc[kTrialSlots];
    bool mx[kTrialSlots];
#pragma unroll
    for (int k = 0; k < kTrialSlots; k++) { acc[k] = 0; mx[k] = (k == 1 || k == 3 || k == 5); }
    const PQ P{p, q};
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const float x = V.x[j], d = V.dx[j], xa = x - V.alpha[j], bx = V.beta[j] - x;
        const float dxi = eps / xa - d / xa * V.xi[j] - V.xi[j];
        const float deta = eps / bx + V.eta[j] / bx * d - V.eta[j];
        const float tx = x + d * step, txi = V.xi[j] + dxi * step, teta = V.eta[j] + deta * step;
        V.nx[j] = tx; V.nxi[j] = txi; V.neta[j] = teta;
        const float ux = V.au[j] - tx, xl = tx - V.al[j];
        const float rex = plam(P, Lnew, m, n, j) / (ux * ux) - qlam(P, Lnew, m, n, j) / (xl * xl) - txi + teta;
        const float rxsi = txi * (tx - V.alpha[j]) - eps;
        const float reeta = teta * (V.beta[j] - tx) - eps;
        acc[0] += (double)rex * rex; acc[1] = fmax(acc[1], (double)fabsf(rex));
        acc[2] += (double)rxsi * rxsi; acc[3] = fmax(acc[3], (double)fabsf(rxsi));
        acc[4] += (double)reeta * reeta; acc[5] = fmax(acc[5], (double)fabsf(reeta));
        for (int i = 0; i < m; i++)
            acc[6 + i] += (double)(p[(size_t)(i + 1) * n + j] / ux + q[(size_t)(i + 1) * n + j] / xl);
    }
    block_reduce<kTrialSlots>(acc, mx, part);
}

// moving-asymptote and move-limit update after the subproblem
// (adjust_asym, mma_t.cu:161-226); slot 0 = |x - x_start|^2
__global__ void __launch_bounds__(256) k_mma_adjust(MmaVecs V, int n, double* __restrict__ part) {
    double acc[1] = {0};
    const bool mx[1] = {false};
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const float x = V.x[j], d1 = x - V.curx[j], d2 = V.lastdx[j];
        float gamma = 1.f;
        if (d1 * d2 > 0) gamma = 1.2f;
        if (d1 * d2 < 0) gamma = 0.7f;
        const float st = V.xmax[j] - V.xmin[j];
        float nal = x - gamma * (x - d1 - V.al[j]);
        float nau = x + gamma * (V.au[j] - x + d1);
        nal = fminf(fmaxf(nal, x - 10.f * st), x - 0.01f * st);
        nau = fminf(fmaxf(nau, x + 0.01f * st), x + 10.f * st);
        V.al[j] = nal; V.au[j] = nau;
        float na = V.xmin[j];
        if (na < nal + 0.1f * (x - nal)) na = nal + 0.1f * (x - nal);
        if (na < x - 0.5f * st) na = x - 0.5f * st;
        float nb = V.xmax[j];
        if (nb > nau - 0.1f * (nau - x)) nb = nau - 0.1f * (nau - x);
        if (nb > x + 0.5f * st) nb = x + 0.5f * st;
        V.alpha[j] = na; V.beta[j] = nb;
        V.lastdx[j] = d1;
        acc[0] += (double)d1 * d1;
    }
    block_reduce<1>(acc, mx, part);
}

// bounds and first point (set_bound + clamp_asymptotes + init_subproblem_variable,
// mma_t.cu:803-866)
__global__ void __launch_bounds__(256) k_mma_init(MmaVecs V, int n, const float* __restrict__ lo,
                                                  const float* __restrict__ hi) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const float l = lo[j], h = hi[j];
        V.xmin[j] = l; V.xmax[j] = h;
        float x = (l + h) / 2.f;
        const float st = h - l;
        const float al = fminf(fmaxf(l, x - 10.f * st), x - 0.01f * st);
        const float au = fminf(fmaxf(h, x + 0.01f * st), x + 10.f * st);
        V.al[j] = al; V.au[j] = au;
        const float a = fmaxf(l, fmaxf(al + 0.1f * (x - al), x - 0.5f * st));
        const float b = fminf(h, fminf(au - 0.1f * (au - x), x + 0.5f * st));
        V.alpha[j] = a; V.beta[j] = b;
        x = (a + b) / 2.f;
        V.x[j] = x;
        V.xi[j] = fmaxf(1.f / (x - a), 1.f);
        V.eta[j] = fmaxf(1.f / (b - x), 1.f);
        V.lastdx[j] = 0.f;
    }
}

int mma_blocks(int n) { return std::max(1, std::min(1024, (n + 255) / 256)); }
int mma_newton_slots() { return kNewtonSlots; }
int mma_trial_slots() { return kTrialSlots; }

#define RT_LAUNCH(k, ...)                                                          \
    do {                                                                           \
        hipLaunchKernelGGL(k, dim3(mma_blocks(n)), dim3(256), 0, s, __VA_ARGS__); \
        RT_HIP(hipGetLastError());                                                 \
    } while (0)

void launch_mma_init(const MmaVecs& V, int n, const float* lo, const float* hi, hipStream_t s) {
    RT_LAUNCH(k_mma_init, V, n, lo, hi);
}
void launch_mma_pq(const MmaVecs& V, int n, int m, const float* df, const float* dg, float* p, float* q,
                   double* part, hipStream_t s) {
    RT_LAUNCH(k_mma_pq, V, n, m, df, dg, p, q, part);
}
void launch_mma_reset(const MmaVecs& V, int n, hipStream_t s) { RT_LAUNCH(k_mma_reset, V, n); }
void launch_mma_newton(const MmaVecs& V, const float* p, const float* q, const MmaLam& L, int n, int m, float eps,
                       double* part, hipStream_t s) {
    RT_LAUNCH(k_mma_newton, V, p, q, L, n, m, eps, part);
}
void launch_mma_dx(const MmaVecs& V, const float* p, const float* q, const MmaLam& L, const MmaLam& dlam, int n,
                   int m, float eps, double* part, hipStream_t s) {
    RT_LAUNCH(k_mma_dx, V, p, q, L, dlam, n, m, eps, part);
}
void launch_mma_trial(const MmaVecs& V, const float* p, const float* q, const MmaLam& Lnew, int n, int m, float eps,
                      float step, double* part, hipStream_t s) {
    RT_LAUNCH(k_mma_trial, V, p, q, Lnew, n, m, eps, step, part);
}
void launch_mma_adjust(const MmaVecs& V, int n, double* part, hipStream_t s) { RT_LAUNCH(k_mma_adjust, V, n, part); }

#undef RT_LAUNCH

}  // namespace rt
