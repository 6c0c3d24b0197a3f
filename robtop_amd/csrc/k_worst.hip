// Worst-case load search (modified power method) on the finest level: random
// start force, projection of a nodal field onto the admissible loads, and the
// unit normalisation / force-change reductions of each step.
//
// Load nodes are a compact list of slots; the projected force lives in a
// compact SoA array fs[3][nl] and is scattered into the full force vector,
// which is zero away from the load nodes.
#include "rt_internal.h"

namespace rt {

__device__ __forceinline__ double blk_sum_w(double v) {
    __shared__ double red[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; i++) t += red[i];
    }
    return t;
}

// splitmix64 finaliser: a counter-based generator, so the force does not
// depend on the launch shape
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// randForce (Grid.cu:2242-2245): components uniform in [-1, 1) on existing nodes
__global__ void __launch_bounds__(256) k_rand_force(int N, const uint8_t* __restrict__ vflag, uint64_t seed,
                                                    Vec3<double> F) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const bool act = vflag[i] & NF_ACTIVE;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const uint64_t r = mix64(seed ^ mix64((uint64_t)i * 3 + c));
            const double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);   // [0, 1)
            F.c[c][i] = act ? 2.0 * u - 1.0 : 0.0;
        }
    }
}
void launch_rand_force(const Geom& g, const uint8_t* vflag, uint64_t seed, Vec3<double> F, hipStream_t s) {
    hipLaunchKernelGGL(k_rand_force, dim3(1024), dim3(256), 0, s, g.N, vflag, seed, F);
    RT_HIP(hipGetLastError());
}

// fs = P x at the load nodes (projection.cu:165-231): the nodal value, or
// its normal component (f.n) n in the constrained-direction mode
// (projection.cpp:338-340); per-block partial sums of |fs|^2
__global__ void __launch_bounds__(256) k_pm_gather(int nl, const int* __restrict__ slots, Vec3<const double> X,
                                                   const double* __restrict__ nrm, double* __restrict__ fs,
                                                   double* __restrict__ partials) {
    double s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nl; i += gridDim.x * blockDim.x) {
        const int v = slots[i];
        double f0 = X.c[0][v], f1 = X.c[1][v], f2 = X.c[2][v];
        if (nrm) {
            const double n0 = nrm[i], n1 = nrm[nl + i], n2 = nrm[2 * nl + i];
            const double d = f0 * n0 + f1 * n1 + f2 * n2;
            f0 = d * n0; f1 = d * n1; f2 = d * n2;
        }
        fs[i] = f0; fs[nl + i] = f1; fs[2 * nl + i] = f2;
        s += f0 * f0 + f1 * f1 + f2 * f2;
    }
    s = blk_sum_w(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
void launch_pm_gather(int nl, const int* slots, Vec3<const double> X, const double* nrm, double* fs,
                      double* partials, int nparts, hipStream_t s) {
    hipLaunchKernelGGL(k_pm_gather, dim3(nparts), dim3(256), 0, s, nl, slots, X, nrm, fs, partials);
    RT_HIP(hipGetLastError());
}

// F[slot] = scale * fs; partials[b] = sum |new - old|^2, partials[nparts + b] =
// sum |old|^2 (supportForceCh / supportForceNorm, Grid.cu:2255-2286); old <- new
__global__ void __launch_bounds__(256) k_pm_commit(int nl, const int* __restrict__ slots,
                                                   const double* __restrict__ fs, double scale, Vec3<double> F,
                                                   double* __restrict__ fold, double* __restrict__ partials) {
    double d = 0, o = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nl; i += gridDim.x * blockDim.x) {
        const int v = slots[i];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double nv = fs[c * nl + i] * scale, ov = fold[c * nl + i];
            F.c[c][v] = nv;
            d += (nv - ov) * (nv - ov);
            o += ov * ov;
            fold[c * nl + i] = nv;
        }
    }
    d = blk_sum_w(d);
    o = blk_sum_w(o);
    if (threadIdx.x == 0) { partials[blockIdx.x] = d; partials[gridDim.x + blockIdx.x] = o; }
}
void launch_pm_commit(int nl, const int* slots, const double* fs, double scale, Vec3<double> F, double* fold,
                      double* partials, int nparts, hipStream_t s) {
    hipLaunchKernelGGL(k_pm_commit, dim3(nparts), dim3(256), 0, s, nl, slots, fs, scale, F, fold, partials);
    RT_HIP(hipGetLastError());
}

}  // namespace rt
