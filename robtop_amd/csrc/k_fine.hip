// Finest-level matrix-free operators on the octant layout (see rt_internal.h).
//
// For a node v the reference forms, element by element (Grid.cu:922-1066 for
// the smoother, :1501-1557 for the residual),
//     KU_v = sum_{e ni v} rho_e^p * KE[v_loc(e), :] * u_e
// where the element incident at offset w = (w%2-1, w/2%2-1, w/4-1) holds v as
// local vertex 7-w. Support ("fixed") neighbours contribute nothing; a fixed
// node gets S = n_e I and KU = 0 in the smoother and r = f in the residual.
//
// Here each thread owns one node. The colour is a template parameter, so every
// neighbour and element slot is base + (compile-time constant combination of
// the runtime strides); threads of a wavefront own consecutive nodes of one
// colour block and every load is a contiguous run. KE lives in constant memory
// and is addressed with compile-time indices (scalar loads).
#include "rt_internal.h"

namespace rt {

__constant__ float c_KE32[576];
__constant__ double c_KE64[576];

template <typename T> __device__ __forceinline__ T ke_at(int i);
template <> __device__ __forceinline__ float ke_at<float>(int i) { return c_KE32[i]; }
template <> __device__ __forceinline__ double ke_at<double>(int i) { return c_KE64[i]; }

void set_ke_constants(const double* ke576, hipStream_t s) {
    float k32[576];
    for (int i = 0; i < 576; i++) k32[i] = (float)ke576[i];
    RT_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_KE64), ke576, sizeof(double) * 576, 0, hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_KE32), k32, sizeof(float) * 576, 0, hipMemcpyHostToDevice, s));
    RT_HIP(hipStreamSynchronize(s));
}

struct ColourGrid {
    int nh[3];   // number of interior positions of this colour per axis
    int total;
};

template <int C>
__device__ __forceinline__ bool colour_thread(const Geom& g, int t, int& base) {
    constexpr int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    const int nhx = (g.n[0] - cx + 1) >> 1, nhy = (g.n[1] - cy + 1) >> 1, nhz = (g.n[2] - cz + 1) >> 1;
    if (t >= nhx * nhy * nhz) return false;
    const int i = t % nhx;
    const int r = t / nhx;
    const int j = r % nhy;
    const int k = r / nhy;
    base = ((k + 1) * g.h[1] + (j + 1)) * g.h[0] + (i + 1);
    return true;
}

// slot of the neighbour at (dx,dy,dz) of a colour-C node with half-index base
template <int C, int DX, int DY, int DZ>
__device__ __forceinline__ int nb_slot(const Geom& g, int base) {
    constexpr int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    constexpr int col = C ^ ((DX & 1) | ((DY & 1) << 1) | ((DZ & 1) << 2));
    constexpr int ox = (cx + DX) >> 1, oy = (cy + DY) >> 1, oz = (cz + DZ) >> 1;
    return col * g.cs + base + (oz * g.h[1] + oy) * g.h[0] + ox;
}

template <int C>
__device__ __forceinline__ int nb_slot_rt(const Geom& g, int base, int nb) {
    constexpr int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    const int dx = nb % 3 - 1, dy = nb / 3 % 3 - 1, dz = nb / 9 - 1;
    const int col = C ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
    const int ox = (cx + dx) >> 1, oy = (cy + dy) >> 1, oz = (cz + dz) >> 1;
    return col * g.cs + base + (oz * g.h[1] + oy) * g.h[0] + ox;
}

// Element incident at w (offset (w%2-1, w/2%2-1, w/4-1)) of a colour-C node.
template <int C>
__device__ __forceinline__ int elem_slot(const Geom& g, int base, int w) {
    constexpr int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    const int ox = w % 2 - 1, oy = w / 2 % 2 - 1, oz = w / 4 - 1;
    const int col = C ^ ((ox & 1) | ((oy & 1) << 1) | ((oz & 1) << 2));
    const int hx = (cx + ox) >> 1, hy = (cy + oy) >> 1, hz = (cz + oz) >> 1;
    return col * g.cs + base + (hz * g.h[1] + hy) * g.h[0] + hx;
}

// KU (and, for the smoother, the self block S) of one colour-C node.
// WITH_SELF: include the node's own displacement (residual/apply);
// CHECK_FIX: zero the displacement of fixed neighbours (Grid.cu:981-984, :1533-1535).
template <typename T, int C, bool WITH_SELF, bool CHECK_FIX, bool WANT_S>
__device__ __forceinline__ void node_KU(const Geom& g, int base, const T* __restrict__ U0, const T* __restrict__ U1,
                                        const T* __restrict__ U2, const float* __restrict__ rhop,
                                        const uint8_t* __restrict__ vflag, T KU[3], T S[9]) {
    T u[27][3];
#pragma unroll
    for (int nb = 0; nb < 27; nb++) {
        if (!WITH_SELF && nb == 13) { u[nb][0] = u[nb][1] = u[nb][2] = T(0); continue; }
        const int s = nb_slot_rt<C>(g, base, nb);
        T a = U0[s], b = U1[s], c = U2[s];
        if (CHECK_FIX) {
            if (vflag[s] & NF_FIXED) { a = T(0); b = T(0); c = T(0); }
        }
        u[nb][0] = a; u[nb][1] = b; u[nb][2] = c;
    }
    KU[0] = KU[1] = KU[2] = T(0);
    if (WANT_S) {
#pragma unroll
        for (int i = 0; i < 9; i++) S[i] = T(0);
    }
#pragma unroll
    for (int w = 0; w < 8; w++) {
        const T pen = (T)rhop[elem_slot<C>(g, base, w)];
        const int vi = 7 - w;
        T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int vj = 0; vj < 8; vj++) {
            if (!WITH_SELF && vj == vi) continue;
            const int nbj = (vj % 2 + w % 2) + 3 * (vj / 2 % 2 + w / 2 % 2) + 9 * (vj / 4 + w / 4);
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int q = 0; q < 3; q++) acc[r] += ke_at<T>((3 * vi + r) * 24 + 3 * vj + q) * u[nbj][q];
        }
#pragma unroll
        for (int r = 0; r < 3; r++) KU[r] += pen * acc[r];
        if (WANT_S) {
#pragma unroll
            for (int i = 0; i < 9; i++) S[i] += pen * ke_at<T>((3 * vi + i / 3) * 24 + 3 * vi + i % 3);
        }
    }
}

template <typename T, int C>
__device__ __forceinline__ void fine_gs_body(const Geom& g, Vec3<T> U, Vec3<const T> F, const float* __restrict__ rhop,
                                             const uint8_t* __restrict__ vflag, bool reverse) {
    int base;
    if (!colour_thread<C>(g, blockIdx.x * blockDim.x + threadIdx.x, base)) return;
    const int v = C * g.cs + base;
    const uint8_t fl = vflag[v];
    if (!(fl & NF_ACTIVE)) return;
    T f[3] = {F.c[0][v], F.c[1][v], F.c[2][v]};
    if (fl & NF_FIXED) {
        // S = n_e I, KeU = 0 (Grid.cu:1007-1014)
        const T ne = (T)(fl >> 4);
        U.c[0][v] = f[0] / ne; U.c[1][v] = f[1] / ne; U.c[2][v] = f[2] / ne;
        return;
    }
    T KU[3], S[9];
    if (fl & NF_NEARFIX)
        node_KU<T, C, false, true, true>(g, base, U.c[0], U.c[1], U.c[2], rhop, vflag, KU, S);
    else
        node_KU<T, C, false, false, true>(g, base, U.c[0], U.c[1], U.c[2], rhop, vflag, KU, S);
    T u[3] = {U.c[0][v], U.c[1][v], U.c[2][v]};
    // component-wise Gauss-Seidel with the 3x3 self block (Grid.cu:1056-1062)
    if (!reverse) {
        u[0] = (f[0] - S[1] * u[1] - S[2] * u[2] - KU[0]) / S[0];
        u[1] = (f[1] - S[3] * u[0] - S[5] * u[2] - KU[1]) / S[4];
        u[2] = (f[2] - S[6] * u[0] - S[7] * u[1] - KU[2]) / S[8];
    } else {
        u[2] = (f[2] - S[6] * u[0] - S[7] * u[1] - KU[2]) / S[8];
        u[1] = (f[1] - S[3] * u[0] - S[5] * u[2] - KU[1]) / S[4];
        u[0] = (f[0] - S[1] * u[1] - S[2] * u[2] - KU[0]) / S[0];
    }
    U.c[0][v] = u[0]; U.c[1][v] = u[1]; U.c[2][v] = u[2];
}

template <typename T, int C>
__global__ void __launch_bounds__(256) k_fine_gs(Geom g, Vec3<T> U, Vec3<const T> F, const float* __restrict__ rhop,
                                                 const uint8_t* __restrict__ vflag, int reverse) {
    fine_gs_body<T, C>(g, U, F, rhop, vflag, reverse != 0);
}

template <typename T>
void launch_fine_gs(const Geom& g, Vec3<T> U, Vec3<const T> F, const float* rhop, const uint8_t* vflag, int colour,
                    bool reverse, hipStream_t s) {
    const int cx = colour & 1, cy = (colour >> 1) & 1, cz = colour >> 2;
    const long long tot = (long long)((g.n[0] - cx + 1) >> 1) * ((g.n[1] - cy + 1) >> 1) * ((g.n[2] - cz + 1) >> 1);
    if (tot <= 0) return;
    dim3 grid((unsigned)((tot + 255) / 256));
    switch (colour) {
#define RT_CASE(c) case c: hipLaunchKernelGGL((k_fine_gs<T, c>), grid, dim3(256), 0, s, g, U, F, rhop, vflag, (int)reverse); break;
        RT_CASE(0) RT_CASE(1) RT_CASE(2) RT_CASE(3) RT_CASE(4) RT_CASE(5) RT_CASE(6) RT_CASE(7)
#undef RT_CASE
    }
    RT_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// residual r = f - K u and apply y = K x (WS rules). One launch covers all
// colours: blockIdx.y is the colour, so the compile-time offsets still apply.
// ---------------------------------------------------------------------------
template <typename T, int C, bool APPLY>
__device__ __forceinline__ T fine_res_body(const Geom& g, Vec3<const T> U, Vec3<const T> F, Vec3<T> R,
                                          const float* __restrict__ rhop, const uint8_t* __restrict__ vflag) {
    int base;
    if (!colour_thread<C>(g, blockIdx.x * blockDim.x + threadIdx.x, base)) return T(0);
    const int v = C * g.cs + base;
    const uint8_t fl = vflag[v];
    if (!(fl & NF_ACTIVE)) return T(0);
    if (fl & NF_FIXED) {
        // KU = 0 at fixed nodes (Grid.cu:1547-1549)
        if (APPLY) { R.c[0][v] = T(0); R.c[1][v] = T(0); R.c[2][v] = T(0); return T(0); }
        R.c[0][v] = F.c[0][v]; R.c[1][v] = F.c[1][v]; R.c[2][v] = F.c[2][v];
        return T(0);
    }
    T KU[3], S[9];
    if (fl & NF_NEARFIX)
        node_KU<T, C, true, true, false>(g, base, U.c[0], U.c[1], U.c[2], rhop, vflag, KU, S);
    else
        node_KU<T, C, true, false, false>(g, base, U.c[0], U.c[1], U.c[2], rhop, vflag, KU, S);
    if (APPLY) {
        R.c[0][v] = KU[0]; R.c[1][v] = KU[1]; R.c[2][v] = KU[2];
        return KU[0] * U.c[0][v] + KU[1] * U.c[1][v] + KU[2] * U.c[2][v];
    }
    R.c[0][v] = F.c[0][v] - KU[0]; R.c[1][v] = F.c[1][v] - KU[1]; R.c[2][v] = F.c[2][v] - KU[2];
    return T(0);
}

__device__ __forceinline__ double block_sum_256(double v) {
    __shared__ double red[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0) t = (red[0] + red[1]) + (red[2] + red[3]);
    return t;
}

template <typename T, bool APPLY>
__global__ void __launch_bounds__(256) k_fine_res(Geom g, Vec3<const T> U, Vec3<const T> F, Vec3<T> R,
                                                  const float* __restrict__ rhop, const uint8_t* __restrict__ vflag,
                                                  double* __restrict__ partials) {
    T d = T(0);
    switch (blockIdx.y) {
#define RT_CASE(c) case c: d = fine_res_body<T, c, APPLY>(g, U, F, R, rhop, vflag); break;
        RT_CASE(0) RT_CASE(1) RT_CASE(2) RT_CASE(3) RT_CASE(4) RT_CASE(5) RT_CASE(6) RT_CASE(7)
#undef RT_CASE
    }
    if (APPLY && partials) {
        double t = block_sum_256((double)d);
        if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = t;
    }
}

static inline unsigned max_colour_blocks(const Geom& g) {
    long long tot = (long long)((g.n[0] + 1) >> 1) * ((g.n[1] + 1) >> 1) * ((g.n[2] + 1) >> 1);
    return (unsigned)((tot + 255) / 256);
}

template <typename T>
void launch_fine_residual(const Geom& g, Vec3<const T> U, Vec3<const T> F, Vec3<T> R, const float* rhop,
                          const uint8_t* vflag, hipStream_t s) {
    dim3 grid(max_colour_blocks(g), 8);
    hipLaunchKernelGGL((k_fine_res<T, false>), grid, dim3(256), 0, s, g, U, F, R, rhop, vflag, (double*)nullptr);
    RT_HIP(hipGetLastError());
}

template <typename T>
void launch_fine_apply(const Geom& g, Vec3<const T> X, Vec3<T> Y, const float* rhop, const uint8_t* vflag,
                       double* partials, int* nparts, hipStream_t s) {
    dim3 grid(max_colour_blocks(g), 8);
    hipLaunchKernelGGL((k_fine_res<T, true>), grid, dim3(256), 0, s, g, X, X, Y, rhop, vflag, partials);
    RT_HIP(hipGetLastError());
    if (nparts) *nparts = (int)(grid.x * grid.y);
}

int fine_apply_partials(const Geom& g) { return (int)(max_colour_blocks(g) * 8); }

#define RT_INST(T)                                                                                               \
    template void launch_fine_gs<T>(const Geom&, Vec3<T>, Vec3<const T>, const float*, const uint8_t*, int, bool, \
                                    hipStream_t);                                                                \
    template void launch_fine_residual<T>(const Geom&, Vec3<const T>, Vec3<const T>, Vec3<T>, const float*,       \
                                          const uint8_t*, hipStream_t);                                          \
    template void launch_fine_apply<T>(const Geom&, Vec3<const T>, Vec3<T>, const float*, const uint8_t*, double*, \
                                       int*, hipStream_t);
RT_INST(float)
RT_INST(double)
#undef RT_INST

}  // namespace rt
