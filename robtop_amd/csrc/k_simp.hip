// SIMP-loop kernels on the finest level: rho^p, sensitivity, sensitivity
// filter, OC trial update and the reductions they need.
#include "rt_internal.h"

namespace rt {

__constant__ double c_KE64_s[576];

void set_ke_simp(const double* ke576, hipStream_t s) {
    RT_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_KE64_s), ke576, sizeof(double) * 576, 0, hipMemcpyHostToDevice, s));
}

__device__ __forceinline__ double blk_sum_s(double v) {
    __shared__ double red[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; i++) t += red[i];
    }
    return t;
}

// rho^p through FP32 powf as the reference does (Grid.cu:834, :963); 0 for absent elements
__global__ void k_rhop(int N, const float* __restrict__ rho, const uint8_t* __restrict__ eflag, float p,
                       float* __restrict__ rhop) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    rhop[i] = (eflag[i] & EF_ACTIVE) ? powf(rho[i], p) : 0.f;
}
void launch_rhop(const Geom& g, const float* rho, const uint8_t* eflag, float p, float* rhop, hipStream_t s) {
    hipLaunchKernelGGL(k_rhop, dim3((g.N + 255) / 256), dim3(256), 0, s, g.N, rho, eflag, p, rhop);
    RT_HIP(hipGetLastError());
}

// Sensitivity g_e = -p rho^(p-1) u_e^T KE u_e (optimization.cu:43-107). The
// reference scatters per-vertex partials with FP32 atomics in arbitrary order;
// here each element gathers its 24 displacements and sums in FP64.
template <typename T>
__global__ void __launch_bounds__(256) k_sens(Geom g, Vec3<const T> U, const float* __restrict__ rho,
                                              const uint8_t* __restrict__ eflag, float p, float* __restrict__ sens) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ex = g.n[0] - 1, ey = g.n[1] - 1, ez = g.n[2] - 1;
    if (t >= ex * ey * ez) return;
    const int x = t % ex, y = (t / ex) % ey, z = t / (ex * ey);
    const int es = oct_p(g, x, y, z);
    if (!(eflag[es] & EF_ACTIVE)) return;
    double ue[24];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int s = oct_p(g, x + (k & 1), y + ((k >> 1) & 1), z + (k >> 2));
        ue[3 * k + 0] = (double)U.c[0][s];
        ue[3 * k + 1] = (double)U.c[1][s];
        ue[3 * k + 2] = (double)U.c[2][s];
    }
    double en = 0;
#pragma unroll
    for (int a = 0; a < 24; a++) {
        double ku = 0;
#pragma unroll
        for (int b = 0; b < 24; b++) ku += c_KE64_s[a * 24 + b] * ue[b];
        en += ue[a] * ku;
    }
    const float r = rho[es];
    const double penal = (double)(p * powf(r, p - 1.f));
    sens[es] = (float)(-penal * en);
}
template <typename T>
void launch_sensitivity(const Geom& g, Vec3<const T> U, const float* rho, const uint8_t* eflag, float p, float* sens,
                        hipStream_t s) {
    int ne = (g.n[0] - 1) * (g.n[1] - 1) * (g.n[2] - 1);
    hipLaunchKernelGGL((k_sens<T>), dim3((ne + 255) / 256), dim3(256), 0, s, g, U, rho, eflag, p, sens);
    RT_HIP(hipGetLastError());
}
template void launch_sensitivity<double>(const Geom&, Vec3<const double>, const float*, const uint8_t*, float, float*,
                                         hipStream_t);
template void launch_sensitivity<float>(const Geom&, Vec3<const float>, const float*, const uint8_t*, float, float*,
                                        hipStream_t);

// Sensitivity filter (Grid.cu:2323-2409): weighted mean over solid elements with
// |d|^2 <= R^2 in the cube of half width (int)(R + 0.5), weight
// w(r) = 1 - 6r^2 + 8r^3 - 3r^4 with r = sqrt(|d|^2 / R^2) in FP32 (Grid.cu:2420-2423);
// numerator in FP64, weight sum in FP32, x outer / z inner as the reference loops.
__global__ void __launch_bounds__(256) k_filter(Geom g, const float* __restrict__ gin,
                                                const uint8_t* __restrict__ eflag, float Rf, float* __restrict__ gout) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ex = g.n[0] - 1, ey = g.n[1] - 1, ez = g.n[2] - 1;
    if (t >= ex * ey * ez) return;
    const int x = t % ex, y = (t / ex) % ey, z = t / (ex * ey);
    const int es = oct_p(g, x, y, z);
    if (!(eflag[es] & EF_ACTIVE)) return;
    const float R2 = Rf * Rf;
    const int Ri = (int)(Rf + 0.5f);
    float wsum = 0.f;
    double gsum = 0.0;
    for (int dx = -Ri; dx <= Ri; dx++) {
        const int nx = x + dx;
        if (nx < 0 || nx >= ex) continue;
        for (int dy = -Ri; dy <= Ri; dy++) {
            const int ny = y + dy;
            if (ny < 0 || ny >= ey) continue;
            for (int dz = -Ri; dz <= Ri; dz++) {
                const int nz = z + dz;
                if (nz < 0 || nz >= ez) continue;
                const float r2 = (float)(dx * dx + dy * dy + dz * dz);
                if (r2 > R2) continue;
                const int ns = oct_p(g, nx, ny, nz);
                if (!(eflag[ns] & EF_ACTIVE)) continue;
                const float r = sqrtf(r2 / R2);
                const float rr = r * r;
                const float w = 1.f - 6.f * rr + 8.f * rr * r - 3.f * rr * rr;
                gsum += (double)(w * gin[ns]);
                wsum += w;
            }
        }
    }
    gout[es] = (float)(gsum / (double)wsum);
}
void launch_filter(const Geom& g, const float* gin, const uint8_t* eflag, float R, float* gout, hipStream_t s) {
    int ne = (g.n[0] - 1) * (g.n[1] - 1) * (g.n[2] - 1);
    hipLaunchKernelGGL(k_filter, dim3((ne + 255) / 256), dim3(256), 0, s, g, gin, eflag, R, gout);
    RT_HIP(hipGetLastError());
}

// OC trial (trySensMultiplier_kernel, optimization.cu:136-162) with per-block
// partial sums of the new densities (for the volume, optimization.cu:199-208).
__global__ void __launch_bounds__(256) k_oc(int N, const float* __restrict__ rho, const float* __restrict__ sens,
                                            const uint8_t* __restrict__ eflag, float thr, float step, float damp,
                                            float rhomin, float* __restrict__ rho_out, double* __restrict__ partials) {
    double s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const uint8_t f = eflag[i];
        if (!(f & EF_ACTIVE)) continue;
        float gg = sens[i];
        if (gg > 0) gg = 0;
        gg = fabsf(gg);
        const float ro = rho[i];
        float rn = ro * powf(gg / thr, damp);
        rn = fminf(fmaxf(rn, ro - step), ro + step);
        rn = fminf(fmaxf(rn, rhomin), 1.f);
        if (f & EF_SHELL) rn = 1.f;
        if (rho_out) rho_out[i] = rn;
        s += (double)rn;
    }
    s = blk_sum_s(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
void launch_oc_trial(const Geom& g, const float* rho, const float* sens, const uint8_t* eflag, float thr, float step,
                     float damp, float rhomin, float* rho_out, double* partials, int nparts, hipStream_t s) {
    hipLaunchKernelGGL(k_oc, dim3(nparts), dim3(256), 0, s, g.N, rho, sens, eflag, thr, step, damp, rhomin, rho_out,
                       partials);
    RT_HIP(hipGetLastError());
}

__global__ void __launch_bounds__(256) k_maxabs(int N, const float* __restrict__ a, const uint8_t* __restrict__ eflag,
                                                double* __restrict__ partials) {
    float m = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        if (eflag[i] & EF_ACTIVE) m = fmaxf(m, fabsf(a[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_down(m, o, 64));
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = (double)fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}
void launch_maxabs(const Geom& g, const float* a, const uint8_t* eflag, double* partials, int nparts, hipStream_t s) {
    hipLaunchKernelGGL(k_maxabs, dim3(nparts), dim3(256), 0, s, g.N, a, eflag, partials);
    RT_HIP(hipGetLastError());
}

// sum of rho over solid elements (volume), per-block partials
__global__ void __launch_bounds__(256) k_sum_active(int N, const float* __restrict__ a,
                                                    const uint8_t* __restrict__ eflag, double* __restrict__ partials) {
    double s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        if (eflag[i] & EF_ACTIVE) s += (double)a[i];
    s = blk_sum_s(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
void launch_sum_active(const Geom& g, const float* a, const uint8_t* eflag, double* partials, int nparts,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_sum_active, dim3(nparts), dim3(256), 0, s, g.N, a, eflag, partials);
    RT_HIP(hipGetLastError());
}

}  // namespace rt
