// Coarse-level stencil smoother/residual, grid transfers, coarsest solve,
// deterministic FP64 reductions and the PCG vector updates.
#include "rt_internal.h"

#include <type_traits>
#include <utility>

namespace rt {

// ---------------------------------------------------------------------------
// helpers shared with k_fine.hip (duplicated as inline templates)
// ---------------------------------------------------------------------------
template <int C>
__device__ __forceinline__ bool c_colour_thread(const Geom& g, int t, int& base) {
    constexpr int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    const int nhx = (g.n[0] - cx + 1) >> 1, nhy = (g.n[1] - cy + 1) >> 1, nhz = (g.n[2] - cz + 1) >> 1;
    if (t >= nhx * nhy * nhz) return false;
    const int i = t % nhx;
    const int r = t / nhx;
    base = ((r / nhy + 1) * g.h[1] + (r % nhy + 1)) * g.h[0] + (i + 1);
    return true;
}

template <int C>
__device__ __forceinline__ int c_nb_slot(const Geom& g, int base, int nb) {
    constexpr int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    const int dx = nb % 3 - 1, dy = nb / 3 % 3 - 1, dz = nb / 9 - 1;
    const int col = C ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
    const int ox = (cx + dx) >> 1, oy = (cy + dy) >> 1, oz = (cz + dz) >> 1;
    return col * g.cs + base + (oz * g.h[1] + oy) * g.h[0] + ox;
}

constexpr int kProlongRowMax = 1026;        // longest fine row of the row-per-workgroup transfer kernels

static inline unsigned colour_blocks(const Geom& g, int colour, int bs) {
    const int cx = colour & 1, cy = (colour >> 1) & 1, cz = colour >> 2;
    long long tot = (long long)((g.n[0] - cx + 1) >> 1) * ((g.n[1] - cy + 1) >> 1) * ((g.n[2] - cz + 1) >> 1);
    return (unsigned)((tot + bs - 1) / bs);
}

template <typename TS> __device__ __forceinline__ double ld_st(const TS* p) { return (double)*p; }

// ---------------------------------------------------------------------------
// stored-stencil operator of the coarse levels
// ---------------------------------------------------------------------------
// Au over the stencil blocks of node v (colour C, half-index base), with raw
// buffer loads at wave-uniform scalar offsets. The 27 neighbours are taken in
// three planes of nine (dz = -1, 0, 1): a plane's 9 x (3 + 9) loads are issued
// together into registers, a scheduling barrier keeps them ahead of the
// plane's FMAs, so each plane costs one memory latency rather than one per
// load. CENTRE selects whether the centre block (nb = 13) is included
// (residual) or left to the caller (smoother).
template <typename T, typename TS, bool CENTRE, int DZ>
__device__ __forceinline__ void coarse_apply_plane(const Geom& g, int C, int vbu, int vbs, int sh,
                                                   const __amdgpu_buffer_rsrc_t (&ru)[3], const TS* __restrict__ st,
                                                   T (&Au)[3]) {
    const int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    const int N = g.N;
    T u[9][3];
    TS a[9][9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int nb = 9 * (DZ + 1) + k;
        if (!CENTRE && nb == 13) continue;
        const int dx = k % 3 - 1, dy = k / 3 - 1;
        const int col = C ^ ((dx & 1) | ((dy & 1) << 1) | ((DZ & 1) << 2));
        const int so = (col * g.cs + (((cz + DZ) >> 1) * g.h[1] + ((cy + dy) >> 1)) * g.h[0] + ((cx + dx) >> 1) + sh) *
                       (int)sizeof(T);
#pragma unroll
        for (int c = 0; c < 3; c++) u[k][c] = bload<T>(ru[c], vbu, so);
        const __amdgpu_buffer_rsrc_t rs = rsrc(st + (size_t)nb * 9 * N, 9 * N * (int)sizeof(TS));
#pragma unroll
        for (int q = 0; q < 9; q++) a[k][q] = bload<TS>(rs, vbs, q * N * (int)sizeof(TS));
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int nb = 9 * (DZ + 1) + k;
        if (!CENTRE && nb == 13) continue;
#pragma unroll
        for (int r = 0; r < 3; r++)
            Au[r] += (T)a[k][3 * r + 0] * u[k][0] + (T)a[k][3 * r + 1] * u[k][1] + (T)a[k][3 * r + 2] * u[k][2];
    }
}

template <typename T, typename TS, bool CENTRE = true>
__device__ __forceinline__ void coarse_apply(const Geom& g, int C, int base, int v, Vec3<const T> U,
                                             const TS* __restrict__ st, T (&Au)[3]) {
    const int N = g.N;
    const int sh = g.h[0] * g.h[1] + g.h[0] + 1;                 // keeps the neighbour offsets >= 0
    const int vbu = (base - sh) * (int)sizeof(T), vbs = v * (int)sizeof(TS);
    const __amdgpu_buffer_rsrc_t ru[3] = {rsrc(U.c[0], N * (int)sizeof(T)), rsrc(U.c[1], N * (int)sizeof(T)),
                                          rsrc(U.c[2], N * (int)sizeof(T))};
    coarse_apply_plane<T, TS, CENTRE, -1>(g, C, vbu, vbs, sh, ru, st, Au);
    coarse_apply_plane<T, TS, CENTRE, 0>(g, C, vbu, vbs, sh, ru, st, Au);
    coarse_apply_plane<T, TS, CENTRE, 1>(g, C, vbu, vbs, sh, ru, st, Au);
}

// coarse smoother: Grid.cu:678-790. Au over the 26 off-centre stencil blocks,
// then component-wise GS with the centre block (stencil index (nb*9 + 3r + q)).
template <typename T, typename TS, int C>
__device__ __forceinline__ void coarse_gs_body(const Geom& g, Vec3<T> U, Vec3<const T> F, const TS* __restrict__ st,
                                               const uint8_t* __restrict__ vflag, bool reverse) {
    int base;
    if (!c_colour_thread<C>(g, blockIdx.x * blockDim.x + threadIdx.x, base)) return;
    const int v = C * g.cs + base;
    if (!(vflag[v] & NF_ACTIVE)) return;
    const size_t N = (size_t)g.N;
    T Au[3] = {T(0), T(0), T(0)};
    coarse_apply<T, TS, false>(g, C, base, v, Vec3<const T>{{U.c[0], U.c[1], U.c[2]}}, st, Au);
    const TS* sc = st + (size_t)13 * 9 * N + v;
    T S[9];
#pragma unroll
    for (int i = 0; i < 9; i++) S[i] = (T)sc[i * N];
    T u[3] = {U.c[0][v], U.c[1][v], U.c[2][v]};
    const T f[3] = {F.c[0][v], F.c[1][v], F.c[2][v]};
    if (!reverse) {
        u[0] = (f[0] - Au[0] - (S[1] * u[1] + S[2] * u[2])) / S[0];
        u[1] = (f[1] - Au[1] - (S[3] * u[0] + S[5] * u[2])) / S[4];
        u[2] = (f[2] - Au[2] - (S[6] * u[0] + S[7] * u[1])) / S[8];
    } else {
        u[2] = (f[2] - Au[2] - (S[6] * u[0] + S[7] * u[1])) / S[8];
        u[1] = (f[1] - Au[1] - (S[3] * u[0] + S[5] * u[2])) / S[4];
        u[0] = (f[0] - Au[0] - (S[1] * u[1] + S[2] * u[2])) / S[0];
    }
    U.c[0][v] = u[0]; U.c[1][v] = u[1]; U.c[2][v] = u[2];
}

template <typename T, typename TS, int C>
__global__ void __launch_bounds__(256) k_coarse_gs(Geom g, Vec3<T> U, Vec3<const T> F, const TS* __restrict__ st,
                                                   const uint8_t* __restrict__ vflag, int reverse) {
    coarse_gs_body<T, TS, C>(g, U, F, st, vflag, reverse != 0);
}

// Sliced form of the same smoother, for levels whose colour passes are too
// short to hide memory latency with one thread per node: a workgroup of nine
// waves owns 64 nodes, wave w reads the three neighbours of neighbour row
// (dy, dz) = (w % 3 - 1, w / 3 - 1) (12 loads per lane, one round trip) and
// the partial products meet in LDS; the wave holding the centre block does the
// component-wise update. Lanes of a wave take consecutive nodes, so every load
// is a unit-stride run, and the row's offsets are wave-uniform scalars.
template <typename T, typename TS, int C>
__global__ void __launch_bounds__(576) k_coarse_gs9(Geom g, Vec3<T> U, Vec3<const T> F, const TS* __restrict__ st,
                                                    const uint8_t* __restrict__ vflag, int reverse) {
    __shared__ T part[9][3][64];
    constexpr int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int dy = w % 3 - 1, dz = w / 3 - 1;
    int base = 0;
    bool act = c_colour_thread<C>(g, blockIdx.x * 64 + lane, base);
    const int v = C * g.cs + base;
    if (act) act = (vflag[v] & NF_ACTIVE) != 0;
    const int N = g.N;
    const int sh = g.h[0] * g.h[1] + g.h[0] + 1;
    const int vbu = act ? (base - sh) * (int)sizeof(T) : 0, vbs = act ? v * (int)sizeof(TS) : 0;
    // f of the node, needed after the barrier by the wave that updates it, is
    // loaded with the neighbour values so it does not add a second load round
    T fo[3] = {T(0), T(0), T(0)};
    if (act && w == 4) { fo[0] = F.c[0][v]; fo[1] = F.c[1][v]; fo[2] = F.c[2][v]; }
    T Au[3] = {T(0), T(0), T(0)};
    T S[9], uc[3] = {T(0), T(0), T(0)};
    if (act) {
        T u[3][3];
        TS a[3][9];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int dx = k - 1;
            const int nb = 9 * (dz + 1) + 3 * (dy + 1) + k;
            const int col = C ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
            const int so = (col * g.cs + (((cz + dz) >> 1) * g.h[1] + ((cy + dy) >> 1)) * g.h[0] + ((cx + dx) >> 1) +
                            sh) * (int)sizeof(T);
            const __amdgpu_buffer_rsrc_t rs = rsrc(st + (size_t)nb * 9 * N, 9 * N * (int)sizeof(TS));
#pragma unroll
            for (int c = 0; c < 3; c++) u[k][c] = bload<T>(rsrc(U.c[c], N * (int)sizeof(T)), vbu, so);
#pragma unroll
            for (int q = 0; q < 9; q++) a[k][q] = bload<TS>(rs, vbs, q * N * (int)sizeof(TS));
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (w == 4 && k == 1) {          // the node itself: centre block and own value
#pragma unroll
                for (int q = 0; q < 9; q++) S[q] = (T)a[k][q];
#pragma unroll
                for (int c = 0; c < 3; c++) uc[c] = u[k][c];
                continue;
            }
#pragma unroll
            for (int r = 0; r < 3; r++)
                Au[r] += (T)a[k][3 * r + 0] * u[k][0] + (T)a[k][3 * r + 1] * u[k][1] + (T)a[k][3 * r + 2] * u[k][2];
        }
    }
#pragma unroll
    for (int r = 0; r < 3; r++) part[w][r][lane] = Au[r];
    __syncthreads();
    if (w != 4 || !act) return;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        T acc = T(0);
#pragma unroll
        for (int q = 0; q < 9; q++) acc += part[q][r][lane];
        Au[r] = acc;
    }
    T u[3] = {uc[0], uc[1], uc[2]};
    const T f[3] = {fo[0], fo[1], fo[2]};
    if (!reverse) {
        u[0] = (f[0] - Au[0] - (S[1] * u[1] + S[2] * u[2])) / S[0];
        u[1] = (f[1] - Au[1] - (S[3] * u[0] + S[5] * u[2])) / S[4];
        u[2] = (f[2] - Au[2] - (S[6] * u[0] + S[7] * u[1])) / S[8];
    } else {
        u[2] = (f[2] - Au[2] - (S[6] * u[0] + S[7] * u[1])) / S[8];
        u[1] = (f[1] - Au[1] - (S[3] * u[0] + S[5] * u[2])) / S[4];
        u[0] = (f[0] - Au[0] - (S[1] * u[1] + S[2] * u[2])) / S[0];
    }
    U.c[0][v] = u[0]; U.c[1][v] = u[1]; U.c[2][v] = u[2];
}

// Colour-pair form of k_coarse_gs9 for levels whose rows hold at most 64
// nodes of one colour: colours C and C^1 (same y, z parity) couple only
// through the x-neighbours of their own row, so a workgroup owns one row
// (y, z) of the pair's class and relaxes its nodes of colour CA, then, after a
// barrier that makes those values visible to the workgroup, its nodes of
// colour CA^1. Each node's arithmetic is that of k_coarse_gs9, and CA, CA^1
// are consecutive in the sweep's colour order, so a sweep gives the same bits
// in half the launches.
template <typename T, typename TS, int C>
__device__ __forceinline__ void coarse_gs9_row(const Geom& g, int j, int k, Vec3<T> U, Vec3<const T> F,
                                               const TS* __restrict__ st, const uint8_t* __restrict__ vflag,
                                               bool reverse, T (&part)[9][3][64]) {
    constexpr int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int dy = w % 3 - 1, dz = w / 3 - 1;
    const int nhx = (g.n[0] - cx + 1) >> 1;
    const int base = ((k + 1) * g.h[1] + (j + 1)) * g.h[0] + (lane + 1);
    bool act = lane < nhx;
    const int v = C * g.cs + base;
    if (act) act = (vflag[v] & NF_ACTIVE) != 0;
    const int N = g.N;
    const int sh = g.h[0] * g.h[1] + g.h[0] + 1;
    const int vbu = act ? (base - sh) * (int)sizeof(T) : 0, vbs = act ? v * (int)sizeof(TS) : 0;
    T fo[3] = {T(0), T(0), T(0)};
    if (act && w == 4) { fo[0] = F.c[0][v]; fo[1] = F.c[1][v]; fo[2] = F.c[2][v]; }
    T Au[3] = {T(0), T(0), T(0)};
    T S[9], uc[3] = {T(0), T(0), T(0)};
    if (act) {
        T u[3][3];
        TS a[3][9];
#pragma unroll
        for (int kk = 0; kk < 3; kk++) {
            const int dx = kk - 1;
            const int nb = 9 * (dz + 1) + 3 * (dy + 1) + kk;
            const int col = C ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
            const int so = (col * g.cs + (((cz + dz) >> 1) * g.h[1] + ((cy + dy) >> 1)) * g.h[0] + ((cx + dx) >> 1) +
                            sh) * (int)sizeof(T);
            const __amdgpu_buffer_rsrc_t rs = rsrc(st + (size_t)nb * 9 * N, 9 * N * (int)sizeof(TS));
#pragma unroll
            for (int c = 0; c < 3; c++) u[kk][c] = bload<T>(rsrc(U.c[c], N * (int)sizeof(T)), vbu, so);
#pragma unroll
            for (int q = 0; q < 9; q++) a[kk][q] = bload<TS>(rs, vbs, q * N * (int)sizeof(TS));
        }
#pragma unroll
        for (int kk = 0; kk < 3; kk++) {
            if (w == 4 && kk == 1) {
#pragma unroll
                for (int q = 0; q < 9; q++) S[q] = (T)a[kk][q];
#pragma unroll
                for (int c = 0; c < 3; c++) uc[c] = u[kk][c];
                continue;
            }
#pragma unroll
            for (int r = 0; r < 3; r++)
                Au[r] += (T)a[kk][3 * r + 0] * u[kk][0] + (T)a[kk][3 * r + 1] * u[kk][1] + (T)a[kk][3 * r + 2] * u[kk][2];
        }
    }
#pragma unroll
    for (int r = 0; r < 3; r++) part[w][r][lane] = Au[r];
    __syncthreads();
    if (w == 4 && act) {
#pragma unroll
        for (int r = 0; r < 3; r++) {
            T acc = T(0);
#pragma unroll
            for (int q = 0; q < 9; q++) acc += part[q][r][lane];
            Au[r] = acc;
        }
        T u[3] = {uc[0], uc[1], uc[2]};
        const T f[3] = {fo[0], fo[1], fo[2]};
        if (!reverse) {
            u[0] = (f[0] - Au[0] - (S[1] * u[1] + S[2] * u[2])) / S[0];
            u[1] = (f[1] - Au[1] - (S[3] * u[0] + S[5] * u[2])) / S[4];
            u[2] = (f[2] - Au[2] - (S[6] * u[0] + S[7] * u[1])) / S[8];
        } else {
            u[2] = (f[2] - Au[2] - (S[6] * u[0] + S[7] * u[1])) / S[8];
            u[1] = (f[1] - Au[1] - (S[3] * u[0] + S[5] * u[2])) / S[4];
            u[0] = (f[0] - Au[0] - (S[1] * u[1] + S[2] * u[2])) / S[0];
        }
        U.c[0][v] = u[0]; U.c[1][v] = u[1]; U.c[2][v] = u[2];
    }
}

template <typename T, typename TS, int CA>
__global__ void __launch_bounds__(576) k_coarse_gs9_pair(Geom g, Vec3<T> U, Vec3<const T> F,
                                                         const TS* __restrict__ st,
                                                         const uint8_t* __restrict__ vflag, int reverse) {
    __shared__ T part[9][3][64];
    constexpr int cy = (CA >> 1) & 1;
    const int nhy = (g.n[1] - cy + 1) >> 1;
    const int j = blockIdx.x % nhy, k = blockIdx.x / nhy;
    coarse_gs9_row<T, TS, CA>(g, j, k, U, F, st, vflag, reverse != 0, part);
    __syncthreads();   // the first colour's new values, and part free again
    coarse_gs9_row<T, TS, CA ^ 1>(g, j, k, U, F, st, vflag, reverse != 0, part);
}

// whether a level's sweeps can use the colour-pair smoother: at most 64 nodes
// of a colour per row, and the sliced smoother selected
static inline bool coarse_pairs_fit(const Geom& g) { return g.coarse_gs == 1 && ((g.n[0] + 1) >> 1) <= 64; }

template <typename T, typename TS>
void launch_coarse_gs_pair(const Geom& g, Vec3<T> U, Vec3<const T> F, const TS* st, const uint8_t* vflag, int pair,
                           bool reverse, hipStream_t s) {
    if (!coarse_pairs_fit(g)) throw std::logic_error("launch_coarse_gs_pair: rows too long or smoother not sliced");
    const int cy = pair & 1, cz = pair >> 1;
    const long long rows = (long long)((g.n[1] - cy + 1) >> 1) * ((g.n[2] - cz + 1) >> 1);
    if (rows <= 0) return;
    // forward: colour 2 pair first (x even), reverse: 2 pair + 1 first
    const int ca = 2 * pair + (reverse ? 1 : 0);
    switch (ca) {
#define RT_CASE(c) case c: hipLaunchKernelGGL((k_coarse_gs9_pair<T, TS, c>), dim3((unsigned)rows), dim3(576), 0, s, g, U, F, st, vflag, (int)reverse); break;
        RT_CASE(0) RT_CASE(1) RT_CASE(2) RT_CASE(3) RT_CASE(4) RT_CASE(5) RT_CASE(6) RT_CASE(7)
#undef RT_CASE
    }
    RT_HIP(hipGetLastError());
}

bool coarse_gs_pairs(const Geom& g) { return coarse_pairs_fit(g); }

template <typename T, typename TS>
void launch_coarse_gs(const Geom& g, Vec3<T> U, Vec3<const T> F, const TS* st, const uint8_t* vflag, int colour,
                      bool reverse, hipStream_t s) {
    if (g.coarse_gs) {
        unsigned nb = colour_blocks(g, colour, 64);
        if (!nb) return;
        switch (colour) {
#define RT_CASE(c) case c: hipLaunchKernelGGL((k_coarse_gs9<T, TS, c>), dim3(nb), dim3(576), 0, s, g, U, F, st, vflag, (int)reverse); break;
            RT_CASE(0) RT_CASE(1) RT_CASE(2) RT_CASE(3) RT_CASE(4) RT_CASE(5) RT_CASE(6) RT_CASE(7)
#undef RT_CASE
        }
        RT_HIP(hipGetLastError());
        return;
    }
    unsigned nb = colour_blocks(g, colour, 256);
    if (!nb) return;
    switch (colour) {
#define RT_CASE(c) case c: hipLaunchKernelGGL((k_coarse_gs<T, TS, c>), dim3(nb), dim3(256), 0, s, g, U, F, st, vflag, (int)reverse); break;
        RT_CASE(0) RT_CASE(1) RT_CASE(2) RT_CASE(3) RT_CASE(4) RT_CASE(5) RT_CASE(6) RT_CASE(7)
#undef RT_CASE
    }
    RT_HIP(hipGetLastError());
}

// coarse residual: Grid.cu:1349-1448. All colours in one launch (blockIdx.y =
// colour). Loads are raw buffer loads whose neighbour / stencil-component
// offsets are wave-uniform scalars, so all 27 x 12 loads of a node are issued
// back to back.
template <typename T, typename TS>
__global__ void __launch_bounds__(256) k_coarse_res(Geom g, Vec3<const T> U, Vec3<const T> F, Vec3<T> R,
                                                    const TS* __restrict__ st, const uint8_t* __restrict__ vflag) {
    const int C = blockIdx.y;
    const int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    const int nhx = (g.n[0] - cx + 1) >> 1, nhy = (g.n[1] - cy + 1) >> 1, nhz = (g.n[2] - cz + 1) >> 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nhx * nhy * nhz) return;
    const int rr = t / nhx;
    const int base = ((rr / nhy + 1) * g.h[1] + (rr % nhy + 1)) * g.h[0] + (t % nhx + 1);
    const int v = C * g.cs + base;
    if (!(vflag[v] & NF_ACTIVE)) return;
    T Au[3] = {T(0), T(0), T(0)};
    coarse_apply<T, TS>(g, C, base, v, U, st, Au);
    R.c[0][v] = F.c[0][v] - Au[0];
    R.c[1][v] = F.c[1][v] - Au[1];
    R.c[2][v] = F.c[2][v] - Au[2];
}

// sliced residual (same partition as k_coarse_gs9): colour = blockIdx.y
template <typename T, typename TS>
__global__ void __launch_bounds__(576) k_coarse_res9(Geom g, Vec3<const T> U, Vec3<const T> F, Vec3<T> R,
                                                     const TS* __restrict__ st, const uint8_t* __restrict__ vflag) {
    __shared__ T part[9][3][64];
    const int C = blockIdx.y;
    const int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    const int nhx = (g.n[0] - cx + 1) >> 1, nhy = (g.n[1] - cy + 1) >> 1, nhz = (g.n[2] - cz + 1) >> 1;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int dy = w % 3 - 1, dz = w / 3 - 1;
    const int t = blockIdx.x * 64 + lane;
    bool act = t < nhx * nhy * nhz;
    int base = 0, v = 0;
    if (act) {
        const int rr = t / nhx;
        base = ((rr / nhy + 1) * g.h[1] + (rr % nhy + 1)) * g.h[0] + (t % nhx + 1);
        v = C * g.cs + base;
        act = (vflag[v] & NF_ACTIVE) != 0;
    }
    const int N = g.N;
    const int sh = g.h[0] * g.h[1] + g.h[0] + 1;
    T fo[3] = {T(0), T(0), T(0)};   // f for wave 0's output, loaded before the barrier
    if (act && w == 0) { fo[0] = F.c[0][v]; fo[1] = F.c[1][v]; fo[2] = F.c[2][v]; }
    T Au[3] = {T(0), T(0), T(0)};
    if (act) {
        const int vbu = (base - sh) * (int)sizeof(T), vbs = v * (int)sizeof(TS);
        T u[3][3];
        TS a[3][9];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int dx = k - 1;
            const int nb = 9 * (dz + 1) + 3 * (dy + 1) + k;
            const int col = C ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
            const int so = (col * g.cs + (((cz + dz) >> 1) * g.h[1] + ((cy + dy) >> 1)) * g.h[0] + ((cx + dx) >> 1) +
                            sh) * (int)sizeof(T);
            const __amdgpu_buffer_rsrc_t rs = rsrc(st + (size_t)nb * 9 * N, 9 * N * (int)sizeof(TS));
#pragma unroll
            for (int c = 0; c < 3; c++) u[k][c] = bload<T>(rsrc(U.c[c], N * (int)sizeof(T)), vbu, so);
#pragma unroll
            for (int q = 0; q < 9; q++) a[k][q] = bload<TS>(rs, vbs, q * N * (int)sizeof(TS));
        }
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int r = 0; r < 3; r++)
                Au[r] += (T)a[k][3 * r + 0] * u[k][0] + (T)a[k][3 * r + 1] * u[k][1] + (T)a[k][3 * r + 2] * u[k][2];
    }
#pragma unroll
    for (int r = 0; r < 3; r++) part[w][r][lane] = Au[r];
    __syncthreads();
    if (w != 0 || !act) return;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        T acc = T(0);
#pragma unroll
        for (int q = 0; q < 9; q++) acc += part[q][r][lane];
        R.c[r][v] = fo[r] - acc;
    }
}

template <typename T, typename TS>
void launch_coarse_residual(const Geom& g, Vec3<const T> U, Vec3<const T> F, Vec3<T> R, const TS* st,
                            const uint8_t* vflag, hipStream_t s) {
    // sliced on the small levels only: on a level with many nodes per colour
    // (L2 at 511 x 256 x 256: 68K) one thread per node streams the stencil faster
    // (115 us against 145 us there)
    const bool sliced = g.coarse_res < 0 ? colour_blocks(g, 0, 1) < 32768 : g.coarse_res != 0;
    if (sliced) {
        dim3 grid(colour_blocks(g, 0, 64), 8);
        if (!grid.x) return;
        hipLaunchKernelGGL((k_coarse_res9<T, TS>), grid, dim3(576), 0, s, g, U, F, R, st, vflag);
    } else {
        dim3 grid(colour_blocks(g, 0, 256), 8);
        if (!grid.x) return;
        hipLaunchKernelGGL((k_coarse_res<T, TS>), grid, dim3(256), 0, s, g, U, F, R, st, vflag);
    }
    RT_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// restriction F_c = P^T r_f over the (2 ratio - 1)^3 window with weights
// prod (ratio - |d|) / ratio (x4: Grid.cu:1733-1794; x2: Grid.cu:1680-1731).
// One thread per coarse node; rows of fine nodes are read with unit stride in x.
// ---------------------------------------------------------------------------
template <typename T, int RATIO>
__global__ void __launch_bounds__(256) k_restrict(Geom gf, int ofx, int ofy, int ofz, Vec3<const T> Rf, Geom gc,
                                                  int ocx, int ocy, int ocz, Vec3<T> Fc,
                                                  const uint8_t* __restrict__ vflag_c) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncx = gc.n[0], ncy = gc.n[1], ncz = gc.n[2];
    if (t >= ncx * ncy * ncz) return;
    const int x = t % ncx, y = (t / ncx) % ncy, z = t / (ncx * ncy);
    const int vc = oct_p(gc, x, y, z);
    if (!(vflag_c[vc] & NF_ACTIVE)) return;
    // coarse node position in fine-box coords
    const int fx = (ocx + x) * RATIO - ofx, fy = (ocy + y) * RATIO - ofy, fz = (ocz + z) * RATIO - ofz;
    constexpr int W = RATIO - 1;
    constexpr T inv = T(1) / T(RATIO * RATIO * RATIO);
    T s0 = T(0), s1 = T(0), s2 = T(0);
    for (int dz = -W; dz <= W; dz++) {
        const int pz = fz + dz;
        if (pz < 0 || pz >= gf.n[2]) continue;
        const int wz = RATIO - (dz < 0 ? -dz : dz);
        for (int dy = -W; dy <= W; dy++) {
            const int py = fy + dy;
            if (py < 0 || py >= gf.n[1]) continue;
            const int wyz = wz * (RATIO - (dy < 0 ? -dy : dy));
#pragma unroll
            for (int dx = -W; dx <= W; dx++) {
                const int px = fx + dx;
                if (px < 0 || px >= gf.n[0]) continue;
                const T w = (T)(wyz * (RATIO - (dx < 0 ? -dx : dx))) * inv;
                const int s = oct_p(gf, px, py, pz);
                s0 += w * Rf.c[0][s];
                s1 += w * Rf.c[1][s];
                s2 += w * Rf.c[2][s];
            }
        }
    }
    Fc.c[0][vc] = s0; Fc.c[1][vc] = s1; Fc.c[2][vc] = s2;
}

// x4 restriction in two separable stages (the weights factor per axis):
//   stage 1: W(xc, yc, z) = sum_{dx,dy} (4 - |dx|)(4 - |dy|) r_f(fx + dx, fy + dy, z)
//            for every fine plane z, into a plain [z][yc][xc] scratch array;
//   stage 2: F_c = sum_dz (4 - |dz|) W(xc, yc, fz + dz) / 64.
// Stage 1 reads each fine value from HBM once (its 7 x 7 window is shared by
// neighbouring lanes through L1/L2); stage 2 reads the 1/16-size scratch.
// Stage 1 as one workgroup per (coarse y, fine z): the 7 fine rows of the y
// window are staged in LDS one after another with unit-stride loads of each
// parity block, and each thread accumulates the (x, y) window of its coarse x
// from LDS. Workgroups are mapped so that each XCD takes a contiguous run of
// coarse y (consecutive windows share 3 fine rows in its L2).
template <typename T, int RATIO>
__global__ void __launch_bounds__(256) k_restrict_xy(Geom gf, int ofx, int ofy, Vec3<const T> Rf, int ncx, int ncy,
                                                     int ocx, int ocy, T* __restrict__ W) {
    __shared__ T row[3][kProlongRowMax];
    const int nx = gf.n[0], ny = gf.n[1], nz = gf.n[2];
    const int b = xcd_row(blockIdx.x, gridDim.x);
    const int yc = b % ncy, z = b / ncy;
    const int fy = (ocy + yc) * RATIO - ofy;
    constexpr int Wd = RATIO - 1;
    constexpr int MAXC = (kProlongRowMax + 255) / 256;
    T s[MAXC][3];
#pragma unroll
    for (int m = 0; m < MAXC; m++) s[m][0] = s[m][1] = s[m][2] = T(0);
    const int nh = (nx + 1) >> 1;
    // rows of the window are double-buffered through registers: the loads of
    // the next row are issued once the current one is in LDS, so they overlap
    // its accumulation
    constexpr int MAXK = (kProlongRowMax / 2 + 255) / 256;
    T nxt[MAXK][2][3];
    auto load_row = [&](int py) {
#pragma unroll
        for (int m = 0; m < MAXK; m++) {
            const int k = threadIdx.x + 256 * m;
#pragma unroll
            for (int par = 0; par < 2; par++) {
                const int x = 2 * k + par;
                if (k < nh && x < nx) {
                    const int sl = oct_p(gf, x, py, z);
                    nxt[m][par][0] = Rf.c[0][sl]; nxt[m][par][1] = Rf.c[1][sl]; nxt[m][par][2] = Rf.c[2][sl];
                }
            }
        }
    };
    const int dy_lo = max(-Wd, -fy), dy_hi = min(Wd, ny - 1 - fy);   // rows inside the box
    if (dy_lo <= dy_hi) load_row(fy + dy_lo);
    for (int dy = dy_lo; dy <= dy_hi; dy++) {
        __syncthreads();                             // the previous row has been consumed
#pragma unroll
        for (int m = 0; m < MAXK; m++) {
            const int k = threadIdx.x + 256 * m;
#pragma unroll
            for (int par = 0; par < 2; par++) {
                const int x = 2 * k + par;
                if (k < nh && x < nx) {
                    row[0][x] = nxt[m][par][0]; row[1][x] = nxt[m][par][1]; row[2][x] = nxt[m][par][2];
                }
            }
        }
        if (dy < dy_hi) load_row(fy + dy + 1);
        __syncthreads();
        const int wy = RATIO - (dy < 0 ? -dy : dy);
#pragma unroll
        for (int m = 0; m < MAXC; m++) {
            const int xc = threadIdx.x + 256 * m;
            if (xc >= ncx) continue;
            const int fx = (ocx + xc) * RATIO - ofx;
#pragma unroll
            for (int dx = -Wd; dx <= Wd; dx++) {
                const int px = fx + dx;
                if (px < 0 || px >= nx) continue;
                const T w = (T)(wy * (RATIO - (dx < 0 ? -dx : dx)));
                s[m][0] += w * row[0][px];
                s[m][1] += w * row[1][px];
                s[m][2] += w * row[2][px];
            }
        }
    }
    const size_t tot = (size_t)ncx * ncy * nz;
#pragma unroll
    for (int m = 0; m < MAXC; m++) {
        const int xc = threadIdx.x + 256 * m;
        if (xc >= ncx) continue;
        const size_t t = ((size_t)z * ncy + yc) * ncx + xc;
        W[t] = s[m][0];
        W[tot + t] = s[m][1];
        W[2 * tot + t] = s[m][2];
    }
}

template <typename T, int RATIO>
__global__ void __launch_bounds__(256) k_restrict_z(int nfz, int ofz, const T* __restrict__ W, Geom gc, int ocz,
                                                    Vec3<T> Fc, const uint8_t* __restrict__ vflag_c) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncx = gc.n[0], ncy = gc.n[1];
    if (t >= ncx * ncy * gc.n[2]) return;
    const int xy = t % (ncx * ncy), zc = t / (ncx * ncy);
    const int vc = oct_p(gc, xy % ncx, xy / ncx, zc);
    if (!(vflag_c[vc] & NF_ACTIVE)) return;
    const int fz = (ocz + zc) * RATIO - ofz;
    const size_t tot = (size_t)ncx * ncy * nfz;
    constexpr int Wd = RATIO - 1;
    constexpr T inv = T(1) / T(RATIO * RATIO * RATIO);
    T s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
    for (int dz = -Wd; dz <= Wd; dz++) {
        const int pz = fz + dz;
        if (pz < 0 || pz >= nfz) continue;
        const T w = (T)(RATIO - (dz < 0 ? -dz : dz));
        const size_t i = (size_t)pz * ncx * ncy + xy;
        s0 += w * W[i];
        s1 += w * W[tot + i];
        s2 += w * W[2 * tot + i];
    }
    Fc.c[0][vc] = s0 * inv; Fc.c[1][vc] = s1 * inv; Fc.c[2][vc] = s2 * inv;
}

size_t restrict_scratch_elems(const Geom& gf, const Geom& gc) {
    return 3 * (size_t)gc.n[0] * gc.n[1] * gf.n[2];
}

template <typename T>
void launch_restrict(const Geom& gf, const int of[3], Vec3<const T> Rf, const Geom& gc, const int oc[3], Vec3<T> Fc,
                     const uint8_t* vflag_c, int ratio, T* scratch, hipStream_t s) {
    long long tot = (long long)gc.n[0] * gc.n[1] * gc.n[2];
    dim3 grid((unsigned)((tot + 255) / 256));
    if (ratio == 4 && scratch && gf.n[0] <= kProlongRowMax) {
        const long long wgs = (long long)gc.n[1] * gf.n[2];
        hipLaunchKernelGGL((k_restrict_xy<T, 4>), dim3((unsigned)wgs), dim3(256), 0, s, gf, of[0], of[1], Rf,
                           gc.n[0], gc.n[1], oc[0], oc[1], scratch);
        RT_HIP(hipGetLastError());
        hipLaunchKernelGGL((k_restrict_z<T, 4>), grid, dim3(256), 0, s, gf.n[2], of[2], (const T*)scratch, gc, oc[2],
                           Fc, vflag_c);
    } else if (ratio == 4)
        hipLaunchKernelGGL((k_restrict<T, 4>), grid, dim3(256), 0, s, gf, of[0], of[1], of[2], Rf, gc, oc[0], oc[1],
                           oc[2], Fc, vflag_c);
    else
        hipLaunchKernelGGL((k_restrict<T, 2>), grid, dim3(256), 0, s, gf, of[0], of[1], of[2], Rf, gc, oc[0], oc[1],
                           oc[2], Fc, vflag_c);
    RT_HIP(hipGetLastError());
}

// prolongation u_f += P u_c (x4: Grid.cu:1859-1895; x2: Grid.cu:1818-1856)
template <typename T, int RATIO>
__global__ void __launch_bounds__(256) k_prolong(Geom gf, int ofx, int ofy, int ofz, Vec3<T> Uf,
                                                 const uint8_t* __restrict__ vflag_f, Geom gc, int ocx, int ocy,
                                                 int ocz, Vec3<const T> Uc, int skip_fixed) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = gf.n[0], ny = gf.n[1], nz = gf.n[2];
    if (t >= nx * ny * nz) return;
    const int x = t % nx, y = (t / nx) % ny, z = t / (nx * ny);
    const int v = oct_p(gf, x, y, z);
    const uint8_t fl = vflag_f[v];
    if (!(fl & NF_ACTIVE) || (skip_fixed && (fl & NF_FIXED))) return;
    const int ax = ofx + x, ay = ofy + y, az = ofz + z;
    const int px = ax % RATIO, py = ay % RATIO, pz = az % RATIO;
    const int bx = ax / RATIO - ocx, by = ay / RATIO - ocy, bz = az / RATIO - ocz;
    constexpr T inv = T(1) / T(RATIO * RATIO * RATIO);
    T c0 = T(0), c1 = T(0), c2 = T(0);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int wx = (i & 1) ? px : RATIO - px;
        const int wy = (i & 2) ? py : RATIO - py;
        const int wz = (i & 4) ? pz : RATIO - pz;
        const int w = wx * wy * wz;
        if (w == 0) continue;
        const int s = oct_p(gc, bx + (i & 1), by + ((i >> 1) & 1), bz + (i >> 2));
        const T ww = (T)w * inv;
        c0 += ww * Uc.c[0][s];
        c1 += ww * Uc.c[1][s];
        c2 += ww * Uc.c[2][s];
    }
    Uf.c[0][v] += c0; Uf.c[1][v] += c1; Uf.c[2][v] += c2;
}

// x4 prolongation, one workgroup per fine row (y, z). Each thread takes coarse
// x-intervals of the row: it interpolates the coarse values in y and z at the
// interval's two ends (8 coarse nodes) and then in x for the RATIO fine nodes
// of the interval, so a fine node costs 6 coarse loads instead of 24. The
// increments meet in LDS by (parity, x / 2), and the row is then updated with
// unit-stride loads and stores of each parity block. Interval gxl covers
// absolute fine x in [RATIO (gx0 + gxl), RATIO (gx0 + gxl + 1)).
template <typename T, int RATIO>
__global__ void __launch_bounds__(256) k_prolong_row(Geom gf, int ofx, int ofy, int ofz, Vec3<T> Uf,
                                                     const uint8_t* __restrict__ vflag_f, Geom gc, int ocx, int ocy,
                                                     int ocz, Vec3<const T> Uc, int skip_fixed, int gx0, int ngx) {
    constexpr int H = kProlongRowMax / 2;
    __shared__ T inc[3][2][H];
    const int nx = gf.n[0], ny = gf.n[1];
    const int y = blockIdx.x % ny, z = blockIdx.x / ny;
    const int ay = ofy + y, az = ofz + z;
    const int py = ay % RATIO, pz = az % RATIO;
    const int by = ay / RATIO - ocy, bz = az / RATIO - ocz;
    constexpr T inv = T(1) / T(RATIO * RATIO * RATIO);
    // the fine row's flags and values do not depend on the interpolation: the
    // thread's first pair of fine nodes is loaded before it, so the two load
    // rounds overlap (each fine node is read and written by one thread only)
    const int nh = (nx + 1) >> 1;
    const int k0 = threadIdx.x;
    uint8_t pf[2] = {0, 0};
    T pu[2][3] = {{T(0), T(0), T(0)}, {T(0), T(0), T(0)}};
#pragma unroll
    for (int par = 0; par < 2; par++) {
        const int x = 2 * k0 + par;
        if (k0 < nh && x < nx) {
            const int v = oct_p(gf, x, y, z);
            pf[par] = vflag_f[v];
            pu[par][0] = Uf.c[0][v]; pu[par][1] = Uf.c[1][v]; pu[par][2] = Uf.c[2][v];
        }
    }
    for (int gxl = threadIdx.x; gxl < ngx; gxl += blockDim.x) {
        const int bxa = gx0 + gxl;
        const int xs = bxa * RATIO - ofx;             // fine x of the interval start
        const int ilo = xs < 0 ? -xs : 0, ihi = nx - xs < RATIO ? nx - xs : RATIO;
        const int bx = bxa - ocx;
        const bool right = ihi > 1;                   // some node of the interval has a weight on bx + 1
        T a0 = T(0), a1 = T(0), a2 = T(0), b0 = T(0), b1 = T(0), b2 = T(0);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int wy = (j & 1) ? py : RATIO - py;
            const int wz = (j & 2) ? pz : RATIO - pz;
            if (wy * wz == 0) continue;
            const T w = (T)(wy * wz);
            const int sa = oct_p(gc, bx, by + (j & 1), bz + (j >> 1));
            a0 += w * Uc.c[0][sa]; a1 += w * Uc.c[1][sa]; a2 += w * Uc.c[2][sa];
            if (right) {
                const int sb = oct_p(gc, bx + 1, by + (j & 1), bz + (j >> 1));
                b0 += w * Uc.c[0][sb]; b1 += w * Uc.c[1][sb]; b2 += w * Uc.c[2][sb];
            }
        }
#pragma unroll
        for (int i = 0; i < RATIO; i++) {
            if (i < ilo || i >= ihi) continue;
            const int x = xs + i;
            const T wl = (T)(RATIO - i), wr = (T)i;
            inc[0][x & 1][x >> 1] = (wl * a0 + wr * b0) * inv;
            inc[1][x & 1][x >> 1] = (wl * a1 + wr * b1) * inv;
            inc[2][x & 1][x >> 1] = (wl * a2 + wr * b2) * inv;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nh; k += blockDim.x) {
#pragma unroll
        for (int par = 0; par < 2; par++) {
            const int x = 2 * k + par;
            if (x >= nx) continue;
            const int v = oct_p(gf, x, y, z);
            uint8_t fl;
            T u0, u1, u2;
            if (k == k0) {
                fl = pf[par]; u0 = pu[par][0]; u1 = pu[par][1]; u2 = pu[par][2];
            } else {
                fl = vflag_f[v];                          // loaded together with the values
                u0 = Uf.c[0][v]; u1 = Uf.c[1][v]; u2 = Uf.c[2][v];
            }
            if (!(fl & NF_ACTIVE) || (skip_fixed && (fl & NF_FIXED))) continue;
            Uf.c[0][v] = u0 + inc[0][par][k];
            Uf.c[1][v] = u1 + inc[1][par][k];
            Uf.c[2][v] = u2 + inc[2][par][k];
        }
    }
}

template <typename T>
void launch_prolong(const Geom& gf, const int of[3], Vec3<T> Uf, const uint8_t* vflag_f, const Geom& gc,
                    const int oc[3], Vec3<const T> Uc, int ratio, bool skip_fixed, hipStream_t s) {
    if (ratio == 4 && gf.n[0] <= kProlongRowMax) {
        const int gx0 = of[0] / 4, ngx = (of[0] + gf.n[0] - 1) / 4 - gx0 + 1;
        const long long rows = (long long)gf.n[1] * gf.n[2];
        hipLaunchKernelGGL((k_prolong_row<T, 4>), dim3((unsigned)rows), dim3(256), 0, s, gf, of[0], of[1], of[2], Uf,
                           vflag_f, gc, oc[0], oc[1], oc[2], Uc, (int)skip_fixed, gx0, ngx);
    } else if (ratio == 4) {
        long long tot = (long long)gf.n[0] * gf.n[1] * gf.n[2];
        hipLaunchKernelGGL((k_prolong<T, 4>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, gf, of[0], of[1],
                           of[2], Uf, vflag_f, gc, oc[0], oc[1], oc[2], Uc, (int)skip_fixed);
    } else {
        long long tot = (long long)gf.n[0] * gf.n[1] * gf.n[2];
        hipLaunchKernelGGL((k_prolong<T, 2>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, gf, of[0], of[1],
                           of[2], Uf, vflag_f, gc, oc[0], oc[1], oc[2], Uc, (int)skip_fixed);
    }
    RT_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// coarsest: u = Ainv f on the active nodes (rowslot[r] = slot of node r)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_gather_rows(const int* __restrict__ rowslot, int nrows, Vec3<const T> F, double* __restrict__ fb) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int s = rowslot[r];
    fb[3 * r + 0] = (double)F.c[0][s];
    fb[3 * r + 1] = (double)F.c[1][s];
    fb[3 * r + 2] = (double)F.c[2][s];
}

template <typename T>
__global__ void __launch_bounds__(64) k_dense_gemv(const double* __restrict__ A, int n, const double* __restrict__ f,
                                                   const int* __restrict__ rowslot, Vec3<T> U) {
    const int row = blockIdx.x;
    double s = 0;
    const double* a = A + (size_t)row * n;
    for (int j = threadIdx.x; j < n; j += 64) s += a[j] * f[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (threadIdx.x == 0) U.c[row % 3][rowslot[row / 3]] = (T)s;
}

template <typename T>
void launch_coarsest(const double* Ainv, int n3, const int* rowslot, int, Vec3<const T> F, Vec3<T> U, double* fbuf,
                     hipStream_t s) {
    const int nrows = n3 / 3;
    if (nrows <= 0) return;
    hipLaunchKernelGGL((k_gather_rows<T>), dim3((nrows + 255) / 256), dim3(256), 0, s, rowslot, nrows, F, fbuf);
    hipLaunchKernelGGL((k_dense_gemv<T>), dim3(n3), dim3(64), 0, s, Ainv, n3, (const double*)fbuf, rowslot, U);
    RT_HIP(hipGetLastError());
}

__global__ void k_to_bf16(const double* __restrict__ a, bf16* __restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t u = __builtin_bit_cast(uint32_t, (float)a[i]);
        u += 0x7FFFu + ((u >> 16) & 1u);
        b[i] = bf16{(uint16_t)(u >> 16)};
    }
}
void launch_to_bf16(const double* a, bf16* b, size_t n, hipStream_t s) {
    unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_to_bf16, dim3(nb), dim3(256), 0, s, a, b, n);
    RT_HIP(hipGetLastError());
}

// A[r][c] = A[c][r] for c < r (row-major n x n): completes a symmetric matrix
// whose upper triangle holds the values
__global__ void k_mirror_upper(double* __restrict__ A, int n) {
    const int r = blockIdx.y;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < r; c += gridDim.x * blockDim.x)
        A[(size_t)r * n + c] = A[(size_t)c * n + r];
}
void launch_mirror_upper(double* A, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_mirror_upper, dim3((n + 255) / 256, n), dim3(256), 0, s, A, n);
    RT_HIP(hipGetLastError());
}

// dense n x n (n = 3 nr) matrix of a level from its FP64 stencils: row block r
// (slot slots[r]) takes neighbour nb's 3 x 3 block at column block
// nbcol[27 r + nb] (-1: outside the box or inactive). A must be zero on entry.
__global__ void k_dense_from_stencil(const double* __restrict__ st, size_t N, const int* __restrict__ slots,
                                     const int* __restrict__ nbcol, int nr, double* __restrict__ A) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)nr * 243) return;
    const int r = (int)(i / 243), k = (int)(i % 243);
    const int nb = k / 9, a = k % 9 / 3, b = k % 3;
    const int c = nbcol[27 * r + nb];
    if (c < 0) return;
    const size_t n = 3 * (size_t)nr;
    A[(3 * (size_t)r + a) * n + 3 * (size_t)c + b] = st[(size_t)k * N + slots[r]];
}
void launch_dense_from_stencil(const double* st, size_t N, const int* slots, const int* nbcol, int nr, double* A,
                               hipStream_t s) {
    const size_t n = 3 * (size_t)nr;
    RT_HIP(hipMemsetAsync(A, 0, sizeof(double) * n * n, s));
    const size_t work = (size_t)nr * 243;
    if (work == 0) return;
    hipLaunchKernelGGL(k_dense_from_stencil, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, st, N, slots,
                       nbcol, nr, A);
    RT_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// vector helpers and deterministic two-stage FP64 reductions
// ---------------------------------------------------------------------------
template <typename TA, typename TB>
__global__ void k_convert(const TA* __restrict__ a, TB* __restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (TB)a[i];
}
template <typename TA, typename TB>
void launch_convert(const TA* a, TB* b, size_t n, hipStream_t s) {
    unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    if (!nb) return;
    hipLaunchKernelGGL((k_convert<TA, TB>), dim3(nb), dim3(256), 0, s, a, b, n);
    RT_HIP(hipGetLastError());
}

template <typename T>
__global__ void k_fill(T* __restrict__ a, T v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}
template <typename T>
void launch_fill(T* a, T v, size_t n, hipStream_t s) {
    unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    if (!nb) return;
    hipLaunchKernelGGL((k_fill<T>), dim3(nb), dim3(256), 0, s, a, v, n);
    RT_HIP(hipGetLastError());
}

__device__ __forceinline__ double blk_sum(double v) {
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

// partials of a.b, four FP32 values per lane
// chunk (optional): one byte per 1024 float4 (4096 floats) of the vectors,
// 0 where every value is known to be 0 (no active node): those are skipped,
// which leaves every sum and vector unchanged
__global__ void __launch_bounds__(256) k_dot4_partials(const float* __restrict__ a, const float* __restrict__ b,
                                                       size_t n4, double* __restrict__ partials,
                                                       const uint8_t* __restrict__ chunk) {
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* b4 = reinterpret_cast<const float4*>(b);
    double s = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
        if (chunk && !chunk[i >> 10]) continue;
        const float4 x = a4[i], y = b4[i];
        s += ((double)x.x * y.x + (double)x.y * y.y) + ((double)x.z * y.z + (double)x.w * y.w);
    }
    s = blk_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename TA, typename TB>
__global__ void __launch_bounds__(256) k_dot_partials(const TA* __restrict__ a, const TB* __restrict__ b, size_t n,
                                                      double* __restrict__ partials) {
    double s = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += (double)a[i] * (double)b[i];
    s = blk_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename TA, typename TB>
void launch_dot_partials_t(const TA* a, const TB* b, size_t n, double* partials, int nparts, hipStream_t s,
                           const uint8_t* chunk) {
    if constexpr (std::is_same<TA, float>::value && std::is_same<TB, float>::value) {
        if (n % 4 == 0) {
            hipLaunchKernelGGL(k_dot4_partials, dim3(nparts), dim3(256), 0, s, a, b, n / 4, partials, chunk);
            RT_HIP(hipGetLastError());
            return;
        }
    }
    hipLaunchKernelGGL((k_dot_partials<TA, TB>), dim3(nparts), dim3(256), 0, s, a, b, n, partials);
    RT_HIP(hipGetLastError());
}
void launch_dot_partials(const double* a, const double* b, size_t n, double* partials, int nparts, hipStream_t s) {
    launch_dot_partials_t<double, double>(a, b, n, partials, nparts, s, nullptr);
}

// chunk mask of a finest-level vector (3 components of N slots): byte c = 1 if
// floats [4096 c, 4096 (c + 1)) hold an active node's value
__global__ void k_vec_chunks(const uint8_t* __restrict__ vflag, int N, uint8_t* __restrict__ chunk) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < N; s += gridDim.x * blockDim.x)
        if (vflag[s] & NF_ACTIVE)
            for (int k = 0; k < 3; k++) chunk[((size_t)k * N + s) >> 12] = 1;
}
void launch_vec_chunks(const uint8_t* vflag, int N, uint8_t* chunk, hipStream_t s) {
    hipLaunchKernelGGL(k_vec_chunks, dim3((unsigned)std::min<long long>(((long long)N + 255) / 256, 16384)), dim3(256),
                       0, s, vflag, N, chunk);
    RT_HIP(hipGetLastError());
}

// one workgroup of 1024 threads; each thread keeps four running sums over
// interleaved runs of the partials (independent loads in flight, fixed order)
__global__ void __launch_bounds__(1024) k_sum_partials(const double* __restrict__ p, int n, double* __restrict__ out) {
    double a[4] = {0, 0, 0, 0};
    const int step = 4 * (int)blockDim.x;
    for (int i = threadIdx.x; i < n; i += step) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int j = i + k * (int)blockDim.x;
            if (j < n) a[k] += p[j];
        }
    }
    double s = (a[0] + a[1]) + (a[2] + a[3]);
    s = blk_sum(s);
    if (threadIdx.x == 0) out[0] = s;
}
void launch_sum_partials(const double* partials, int nparts, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, partials, nparts, out);
    RT_HIP(hipGetLastError());
}

// x += a p; r -= a q; partials of r.r; rz = (TZ) r
template <typename T, typename TZ>
__global__ void __launch_bounds__(256) k_pcg_xr(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
                                                const T* __restrict__ q, double alpha, size_t n,
                                                double* __restrict__ partials, TZ* __restrict__ rz) {
    const T a = (T)alpha;
    double s = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        x[i] += a * p[i];
        const T ri = r[i] - a * q[i];
        r[i] = ri;
        s += (double)ri * (double)ri;
        if (rz) rz[i] = (TZ)ri;
    }
    s = blk_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
template <typename T, typename TZ>
void launch_pcg_update_x_r(T* x, T* r, const T* p, const T* q, double alpha, size_t n, double* partials, int nparts,
                           TZ* rz_out, hipStream_t s) {
    hipLaunchKernelGGL((k_pcg_xr<T, TZ>), dim3(nparts), dim3(256), 0, s, x, r, p, q, alpha, n, partials, rz_out);
    RT_HIP(hipGetLastError());
}

// reliable-update PCG step: r (FP32) -= a q with partials of r.r, and, when
// DX, the FP32 solution increment dx += a p in the same pass.
// The FP32 streaming kernels of the reliable-update PCG take four values per
// lane per step (16 B loads) when n is a multiple of 4 (it is: n = 3 * 8 cs),
// over kStreamBlocks workgroups, so enough bytes are in flight to approach
// the HBM rate.
template <bool DX>
__global__ void __launch_bounds__(256) k_pcg_ru_dxr(float* __restrict__ r, const float* __restrict__ q,
                                                    float* __restrict__ dx, const float* __restrict__ p, float a,
                                                    size_t n4, double* __restrict__ partials,
                                                    const uint8_t* __restrict__ chunk) {
    float4* r4 = reinterpret_cast<float4*>(r);
    float4* d4 = reinterpret_cast<float4*>(dx);
    const float4* q4 = reinterpret_cast<const float4*>(q);
    const float4* p4 = reinterpret_cast<const float4*>(p);
    double s = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
        if (chunk && !chunk[i >> 10]) continue;
        float4 ri = r4[i];
        const float4 qi = q4[i];
        ri.x -= a * qi.x; ri.y -= a * qi.y; ri.z -= a * qi.z; ri.w -= a * qi.w;
        r4[i] = ri;
        if constexpr (DX) {
            const float4 pi = p4[i];
            float4 di = d4[i];
            di.x += a * pi.x; di.y += a * pi.y; di.z += a * pi.z; di.w += a * pi.w;
            d4[i] = di;
        }
        s += ((double)ri.x * ri.x + (double)ri.y * ri.y) + ((double)ri.z * ri.z + (double)ri.w * ri.w);
    }
    s = blk_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
void launch_pcg_ru_dxr(float* r, const float* q, float* dx, const float* p, double alpha, size_t n, double* partials,
                       int nparts, hipStream_t s, const uint8_t* chunk) {
    if (n % 4) throw std::logic_error("launch_pcg_ru_dxr: n must be a multiple of 4");
    if (dx) {
        if (!p) throw std::logic_error("launch_pcg_ru_dxr: dx without p");
        hipLaunchKernelGGL(k_pcg_ru_dxr<true>, dim3(nparts), dim3(256), 0, s, r, q, dx, p, (float)alpha, n / 4,
                           partials, chunk);
    } else {
        hipLaunchKernelGGL(k_pcg_ru_dxr<false>, dim3(nparts), dim3(256), 0, s, r, q, dx, p, (float)alpha, n / 4,
                           partials, chunk);
    }
    RT_HIP(hipGetLastError());
}

// x += dx (FP64), dx = 0. With p, the increment a p still pending on dx is
// added first, in FP32 as the streaming kernels add it: x += (dx + a p).
__global__ void k_pcg_flush_x(double* __restrict__ x, float* __restrict__ dx, const float* __restrict__ p, float a,
                              size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float d = dx[i];
        if (p) d += a * p[i];
        x[i] += (double)d;
        dx[i] = 0.f;
    }
}
void launch_pcg_flush_x(double* x, float* dx, size_t n, hipStream_t s, const float* p, double alpha) {
    unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_pcg_flush_x, dim3(nb), dim3(256), 0, s, x, dx, p, (float)alpha, n);
    RT_HIP(hipGetLastError());
}

// dx += a p, then p = z + b p: the direction update with the previous step's
// solution increment, deferred from k_pcg_ru_dxr so that p is read once for
// both (same FP32 operations, so the same bits as updating dx with r)
__global__ void __launch_bounds__(256) k_pcg_p4_dx(float* __restrict__ p, const float* __restrict__ z, float b,
                                                   float* __restrict__ dx, float a, size_t n4,
                                                   const uint8_t* __restrict__ chunk) {
    float4* p4 = reinterpret_cast<float4*>(p);
    float4* d4 = reinterpret_cast<float4*>(dx);
    const float4* z4 = reinterpret_cast<const float4*>(z);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
        if (chunk && !chunk[i >> 10]) continue;
        const float4 zi = z4[i];
        float4 pi = p4[i];
        float4 di = d4[i];
        di.x += a * pi.x; di.y += a * pi.y; di.z += a * pi.z; di.w += a * pi.w;
        pi.x = zi.x + b * pi.x; pi.y = zi.y + b * pi.y; pi.z = zi.z + b * pi.z; pi.w = zi.w + b * pi.w;
        d4[i] = di;
        p4[i] = pi;
    }
}
void launch_pcg_p_dx(float* p, const float* z, double beta, float* dx, double alpha, size_t n, hipStream_t s,
                     const uint8_t* chunk) {
    if (n % 4) throw std::logic_error("launch_pcg_p_dx: n must be a multiple of 4");
    hipLaunchKernelGGL(k_pcg_p4_dx, dim3(kStreamBlocks), dim3(256), 0, s, p, z, (float)beta, dx, (float)alpha, n / 4,
                       chunk);
    RT_HIP(hipGetLastError());
}

// p = z + b p, four FP32 values per lane
__global__ void __launch_bounds__(256) k_pcg_p4(float* __restrict__ p, const float* __restrict__ z, float b,
                                                size_t n4, const uint8_t* __restrict__ chunk) {
    float4* p4 = reinterpret_cast<float4*>(p);
    const float4* z4 = reinterpret_cast<const float4*>(z);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
        if (chunk && !chunk[i >> 10]) continue;
        const float4 zi = z4[i];
        float4 pi = p4[i];
        pi.x = zi.x + b * pi.x; pi.y = zi.y + b * pi.y; pi.z = zi.z + b * pi.z; pi.w = zi.w + b * pi.w;
        p4[i] = pi;
    }
}


// p = z + b p
template <typename T, typename TZ>
__global__ void k_pcg_p(T* __restrict__ p, const TZ* __restrict__ z, double beta, size_t n) {
    const T b = (T)beta;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = (T)z[i] + b * p[i];
}
template <typename T, typename TZ>
void launch_pcg_update_p(T* p, const TZ* z, double beta, size_t n, hipStream_t s, const uint8_t* chunk) {
    if constexpr (std::is_same<T, float>::value && std::is_same<TZ, float>::value) {
        if (n % 4 == 0) {
            hipLaunchKernelGGL(k_pcg_p4, dim3(kStreamBlocks), dim3(256), 0, s, p, z, (float)beta, n / 4, chunk);
            RT_HIP(hipGetLastError());
            return;
        }
    }
    unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL((k_pcg_p<T, TZ>), dim3(nb), dim3(256), 0, s, p, z, beta, n);
    RT_HIP(hipGetLastError());
}

// r = b - q; partials of r.r; rz = (TZ) r
template <typename T, typename TZ>
__global__ void __launch_bounds__(256) k_sub_conv(const T* __restrict__ b, const T* __restrict__ q, T* __restrict__ r,
                                                  TZ* __restrict__ rz, size_t n, double* __restrict__ partials) {
    double s = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const T ri = b[i] - q[i];
        r[i] = ri;
        if (rz) rz[i] = (TZ)ri;
        s += (double)ri * (double)ri;
    }
    s = blk_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
template <typename T, typename TZ>
void launch_sub_conv(const T* b, const T* q, T* r, TZ* rz, size_t n, double* partials, int nparts, hipStream_t s) {
    hipLaunchKernelGGL((k_sub_conv<T, TZ>), dim3(nparts), dim3(256), 0, s, b, q, r, rz, n, partials);
    RT_HIP(hipGetLastError());
}

// x *= a
template <typename T>
__global__ void k_scale(T* __restrict__ x, double a, size_t n) {
    const T s = (T)a;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] *= s;
}
template <typename T>
void launch_scale(T* x, double a, size_t n, hipStream_t s) {
    unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL((k_scale<T>), dim3(nb), dim3(256), 0, s, x, a, n);
    RT_HIP(hipGetLastError());
}
template void launch_scale<float>(float*, double, size_t, hipStream_t);
template void launch_scale<double>(double*, double, size_t, hipStream_t);

// x = sum_i c[i] h[i] over k <= 8 vectors (FP64)
struct LinComb {
    const double* h[8];
    double c[8];
    int k;
};
__global__ void k_lincomb(double* __restrict__ x, LinComb lc, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double s = 0;
        for (int j = 0; j < lc.k; j++) s += lc.c[j] * lc.h[j][i];
        x[i] = s;
    }
}
void launch_lincomb(double* x, const double* const* h, const double* c, int k, size_t n, hipStream_t s) {
    if (k < 1 || k > 8) throw std::invalid_argument("launch_lincomb: 1..8 vectors");
    LinComb lc{};
    for (int j = 0; j < k; j++) { lc.h[j] = h[j]; lc.c[j] = c[j]; }
    lc.k = k;
    unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_lincomb, dim3(nb), dim3(256), 0, s, x, lc, n);
    RT_HIP(hipGetLastError());
}

// x += e (TE -> T)
template <typename T, typename TE>
__global__ void k_add_conv(T* __restrict__ x, const TE* __restrict__ e, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] += (T)e[i];
}
template <typename T, typename TE>
void launch_add_conv(T* x, const TE* e, size_t n, hipStream_t s) {
    unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL((k_add_conv<T, TE>), dim3(nb), dim3(256), 0, s, x, e, n);
    RT_HIP(hipGetLastError());
}

// scalar helpers: out = a / b (device scalars), optionally negated
__global__ void k_ratio(const double* a, const double* b, double* out) { out[0] = a[0] / b[0]; }
void launch_ratio(const double* a, const double* b, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_ratio, dim3(1), dim3(1), 0, s, a, b, out);
    RT_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// box (natural, x fastest) <-> octant slot layout
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_box_oct(Geom g, int bx, int by, int bz, const T* __restrict__ src, T* __restrict__ dst, int to_oct) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)bx * by * bz) return;
    const int x = (int)(t % bx), y = (int)((t / bx) % by), z = (int)(t / ((size_t)bx * by));
    const int s = oct_p(g, x, y, z);
    if (to_oct) dst[s] = src[t]; else dst[t] = src[s];
}
template <typename T>
void launch_box_to_oct(const Geom& g, const T* box, T* oct, int node_or_elem, hipStream_t s) {
    int d = node_or_elem ? 1 : 0;
    size_t tot = (size_t)(g.n[0] - d) * (g.n[1] - d) * (g.n[2] - d);
    if (!tot) return;
    hipLaunchKernelGGL((k_box_oct<T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, g.n[0] - d,
                       g.n[1] - d, g.n[2] - d, box, oct, 1);
    RT_HIP(hipGetLastError());
}
template <typename T>
void launch_oct_to_box(const Geom& g, const T* oct, T* box, int node_or_elem, hipStream_t s) {
    int d = node_or_elem ? 1 : 0;
    size_t tot = (size_t)(g.n[0] - d) * (g.n[1] - d) * (g.n[2] - d);
    if (!tot) return;
    hipLaunchKernelGGL((k_box_oct<T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, g.n[0] - d,
                       g.n[1] - d, g.n[2] - d, oct, box, 0);
    RT_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// instantiations
// ---------------------------------------------------------------------------
#define RT_INST_CS(T, TS)                                                                                           \
    template void launch_coarse_gs_pair<T, TS>(const Geom&, Vec3<T>, Vec3<const T>, const TS*, const uint8_t*, int, \
                                               bool, hipStream_t);                                               \
    template void launch_coarse_gs<T, TS>(const Geom&, Vec3<T>, Vec3<const T>, const TS*, const uint8_t*, int, bool, \
                                          hipStream_t);                                                             \
    template void launch_coarse_residual<T, TS>(const Geom&, Vec3<const T>, Vec3<const T>, Vec3<T>, const TS*,      \
                                                const uint8_t*, hipStream_t);
RT_INST_CS(double, double)
RT_INST_CS(float, float)
RT_INST_CS(float, bf16)
#undef RT_INST_CS

#define RT_INST_T(T)                                                                                                  \
    template void launch_restrict<T>(const Geom&, const int*, Vec3<const T>, const Geom&, const int*, Vec3<T>,        \
                                     const uint8_t*, int, T*, hipStream_t);                                           \
    template void launch_prolong<T>(const Geom&, const int*, Vec3<T>, const uint8_t*, const Geom&, const int*,        \
                                    Vec3<const T>, int, bool, hipStream_t);                                           \
    template void launch_coarsest<T>(const double*, int, const int*, int, Vec3<const T>, Vec3<T>, double*,            \
                                     hipStream_t);                                                                    \
    template void launch_fill<T>(T*, T, size_t, hipStream_t);                                                         \
    template void launch_box_to_oct<T>(const Geom&, const T*, T*, int, hipStream_t);                                  \
    template void launch_oct_to_box<T>(const Geom&, const T*, T*, int, hipStream_t);
RT_INST_T(double)
RT_INST_T(float)
#undef RT_INST_T
template void launch_fill<uint8_t>(uint8_t*, uint8_t, size_t, hipStream_t);
template void launch_box_to_oct<uint8_t>(const Geom&, const uint8_t*, uint8_t*, int, hipStream_t);
template void launch_oct_to_box<uint8_t>(const Geom&, const uint8_t*, uint8_t*, int, hipStream_t);
template void launch_convert<double, float>(const double*, float*, size_t, hipStream_t);
template void launch_convert<float, double>(const float*, double*, size_t, hipStream_t);
template void launch_convert<float, float>(const float*, float*, size_t, hipStream_t);
template void launch_convert<double, double>(const double*, double*, size_t, hipStream_t);
template void launch_dot_partials_t<double, double>(const double*, const double*, size_t, double*, int, hipStream_t,
                                                  const uint8_t*);
template void launch_dot_partials_t<double, float>(const double*, const float*, size_t, double*, int, hipStream_t,
                                                  const uint8_t*);
template void launch_dot_partials_t<float, float>(const float*, const float*, size_t, double*, int, hipStream_t,
                                                  const uint8_t*);
#define RT_INST_PCG(T, TZ)                                                                                        \
    template void launch_pcg_update_x_r<T, TZ>(T*, T*, const T*, const T*, double, size_t, double*, int, TZ*,        \
                                               hipStream_t);                                                       \
    template void launch_pcg_update_p<T, TZ>(T*, const TZ*, double, size_t, hipStream_t, const uint8_t*);                          \
    template void launch_sub_conv<T, TZ>(const T*, const T*, T*, TZ*, size_t, double*, int, hipStream_t);          \
    template void launch_add_conv<T, TZ>(T*, const TZ*, size_t, hipStream_t);
RT_INST_PCG(double, float)
RT_INST_PCG(double, double)
RT_INST_PCG(float, float)
RT_INST_PCG(float, double)
#undef RT_INST_PCG

// ---------------------------------------------------------------------------
// Batched forms for nb <= kBatchMax right-hand sides at once (the worst-case
// search pushes a Lanczos block of load vectors through one V-cycle). Vector b
// of a batch starts at base + b * bs (bs = 3N), its component c at + c * N.
// A workgroup loads each stencil coefficient, and the coarsest inverse each
// matrix entry, once and applies it to all nb vectors; per vector the
// arithmetic and its order are those of the single-vector kernels, so a
// batched cycle gives bit for bit the single cycles' results.
// ---------------------------------------------------------------------------
template <typename T, typename TS, int C>
__global__ void __launch_bounds__(576) k_coarse_gs9b(Geom g, T* __restrict__ U, const T* __restrict__ F, int bs,
                                                     int nb, const TS* __restrict__ st,
                                                     const uint8_t* __restrict__ vflag, int reverse) {
    __shared__ T part[kBatchMax][9][3][64];
    constexpr int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int dy = w % 3 - 1, dz = w / 3 - 1;
    int base = 0;
    bool act = c_colour_thread<C>(g, blockIdx.x * 64 + lane, base);
    const int v = C * g.cs + base;
    if (act) act = (vflag[v] & NF_ACTIVE) != 0;
    const int N = g.N;
    const int sh = g.h[0] * g.h[1] + g.h[0] + 1;
    const int vbu = act ? (base - sh) * (int)sizeof(T) : 0, vbs = act ? v * (int)sizeof(TS) : 0;
    T fo[kBatchMax][3], uc[kBatchMax][3];
    T S[9];
    TS a[3][9];
    int so[3];
    if (act) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int dx = k - 1;
            const int nbk = 9 * (dz + 1) + 3 * (dy + 1) + k;
            const int col = C ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
            so[k] = (col * g.cs + (((cz + dz) >> 1) * g.h[1] + ((cy + dy) >> 1)) * g.h[0] + ((cx + dx) >> 1) + sh) *
                    (int)sizeof(T);
            const __amdgpu_buffer_rsrc_t rs = rsrc(st + (size_t)nbk * 9 * N, 9 * N * (int)sizeof(TS));
#pragma unroll
            for (int q = 0; q < 9; q++) a[k][q] = bload<TS>(rs, vbs, q * N * (int)sizeof(TS));
        }
    }
#pragma unroll
    for (int b = 0; b < kBatchMax; b++) {
        if (b >= nb) break;
        const T* Ub = U + (size_t)b * bs;
        T Au[3] = {T(0), T(0), T(0)};
        if (act && w == 4) {
            const T* Fb = F + (size_t)b * bs;
            fo[b][0] = Fb[v]; fo[b][1] = Fb[N + v]; fo[b][2] = Fb[2 * N + v];
        }
        if (act) {
            T u[3][3];
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int c = 0; c < 3; c++) u[k][c] = bload<T>(rsrc(Ub + (size_t)c * N, N * (int)sizeof(T)), vbu, so[k]);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                if (w == 4 && k == 1) {
#pragma unroll
                    for (int c = 0; c < 3; c++) uc[b][c] = u[k][c];
                    continue;
                }
#pragma unroll
                for (int r = 0; r < 3; r++)
                    Au[r] += (T)a[k][3 * r + 0] * u[k][0] + (T)a[k][3 * r + 1] * u[k][1] + (T)a[k][3 * r + 2] * u[k][2];
            }
        }
#pragma unroll
        for (int r = 0; r < 3; r++) part[b][w][r][lane] = Au[r];
    }
    if (act && w == 4)
#pragma unroll
        for (int q = 0; q < 9; q++) S[q] = (T)a[1][q];
    __syncthreads();
    if (w != 4 || !act) return;
#pragma unroll
    for (int b = 0; b < kBatchMax; b++) {
        if (b >= nb) break;
        T Au[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            T acc = T(0);
#pragma unroll
            for (int q = 0; q < 9; q++) acc += part[b][q][r][lane];
            Au[r] = acc;
        }
        T u[3] = {uc[b][0], uc[b][1], uc[b][2]};
        const T f[3] = {fo[b][0], fo[b][1], fo[b][2]};
        if (!reverse) {
            u[0] = (f[0] - Au[0] - (S[1] * u[1] + S[2] * u[2])) / S[0];
            u[1] = (f[1] - Au[1] - (S[3] * u[0] + S[5] * u[2])) / S[4];
            u[2] = (f[2] - Au[2] - (S[6] * u[0] + S[7] * u[1])) / S[8];
        } else {
            u[2] = (f[2] - Au[2] - (S[6] * u[0] + S[7] * u[1])) / S[8];
            u[1] = (f[1] - Au[1] - (S[3] * u[0] + S[5] * u[2])) / S[4];
            u[0] = (f[0] - Au[0] - (S[1] * u[1] + S[2] * u[2])) / S[0];
        }
        T* Ub = U + (size_t)b * bs;
        Ub[v] = u[0]; Ub[N + v] = u[1]; Ub[2 * N + v] = u[2];
    }
}

template <typename T, typename TS>
__global__ void __launch_bounds__(576) k_coarse_res9b(Geom g, const T* __restrict__ U, const T* __restrict__ F,
                                                      T* __restrict__ R, int bs, int nb, const TS* __restrict__ st,
                                                      const uint8_t* __restrict__ vflag) {
    __shared__ T part[kBatchMax][9][3][64];
    const int C = blockIdx.y;
    const int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    const int nhx = (g.n[0] - cx + 1) >> 1, nhy = (g.n[1] - cy + 1) >> 1, nhz = (g.n[2] - cz + 1) >> 1;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int dy = w % 3 - 1, dz = w / 3 - 1;
    const int t = blockIdx.x * 64 + lane;
    bool act = t < nhx * nhy * nhz;
    int base = 0, v = 0;
    if (act) {
        const int rr = t / nhx;
        base = ((rr / nhy + 1) * g.h[1] + (rr % nhy + 1)) * g.h[0] + (t % nhx + 1);
        v = C * g.cs + base;
        act = (vflag[v] & NF_ACTIVE) != 0;
    }
    const int N = g.N;
    const int sh = g.h[0] * g.h[1] + g.h[0] + 1;
    const int vbu = act ? (base - sh) * (int)sizeof(T) : 0, vbs = act ? v * (int)sizeof(TS) : 0;
    TS a[3][9];
    int so[3];
    if (act) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int dx = k - 1;
            const int nbk = 9 * (dz + 1) + 3 * (dy + 1) + k;
            const int col = C ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
            so[k] = (col * g.cs + (((cz + dz) >> 1) * g.h[1] + ((cy + dy) >> 1)) * g.h[0] + ((cx + dx) >> 1) + sh) *
                    (int)sizeof(T);
            const __amdgpu_buffer_rsrc_t rs = rsrc(st + (size_t)nbk * 9 * N, 9 * N * (int)sizeof(TS));
#pragma unroll
            for (int q = 0; q < 9; q++) a[k][q] = bload<TS>(rs, vbs, q * N * (int)sizeof(TS));
        }
    }
    T fo[kBatchMax][3];
#pragma unroll
    for (int b = 0; b < kBatchMax; b++) {
        if (b >= nb) break;
        T Au[3] = {T(0), T(0), T(0)};
        if (act && w == 0) {
            const T* Fb = F + (size_t)b * bs;
            fo[b][0] = Fb[v]; fo[b][1] = Fb[N + v]; fo[b][2] = Fb[2 * N + v];
        }
        if (act) {
            const T* Ub = U + (size_t)b * bs;
            T u[3][3];
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int c = 0; c < 3; c++) u[k][c] = bload<T>(rsrc(Ub + (size_t)c * N, N * (int)sizeof(T)), vbu, so[k]);
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int r = 0; r < 3; r++)
                    Au[r] += (T)a[k][3 * r + 0] * u[k][0] + (T)a[k][3 * r + 1] * u[k][1] + (T)a[k][3 * r + 2] * u[k][2];
        }
#pragma unroll
        for (int r = 0; r < 3; r++) part[b][w][r][lane] = Au[r];
    }
    __syncthreads();
    if (w != 0 || !act) return;
#pragma unroll
    for (int b = 0; b < kBatchMax; b++) {
        if (b >= nb) break;
        T* Rb = R + (size_t)b * bs;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            T acc = T(0);
#pragma unroll
            for (int q = 0; q < 9; q++) acc += part[b][q][r][lane];
            Rb[(size_t)r * N + v] = fo[b][r] - acc;
        }
    }
}

template <typename T, typename TS>
void launch_coarse_gs_batch(const Geom& g, int nb, T* U, const T* F, const TS* st, const uint8_t* vflag, int colour,
                            bool reverse, hipStream_t s) {
    if (nb < 1 || nb > kBatchMax) throw std::invalid_argument("batch size out of range");
    const unsigned blocks = colour_blocks(g, colour, 64);
    if (!blocks) return;
    const int bs = 3 * g.N;
    switch (colour) {
#define RT_CASE(c) case c: hipLaunchKernelGGL((k_coarse_gs9b<T, TS, c>), dim3(blocks), dim3(576), 0, s, g, U, F, bs, nb, st, vflag, (int)reverse); break;
        RT_CASE(0) RT_CASE(1) RT_CASE(2) RT_CASE(3) RT_CASE(4) RT_CASE(5) RT_CASE(6) RT_CASE(7)
#undef RT_CASE
    }
    RT_HIP(hipGetLastError());
}

template <typename T, typename TS>
void launch_coarse_residual_batch(const Geom& g, int nb, const T* U, const T* F, T* R, const TS* st,
                                  const uint8_t* vflag, hipStream_t s) {
    if (nb < 1 || nb > kBatchMax) throw std::invalid_argument("batch size out of range");
    dim3 grid(colour_blocks(g, 0, 64), 8);
    if (!grid.x) return;
    hipLaunchKernelGGL((k_coarse_res9b<T, TS>), grid, dim3(576), 0, s, g, U, F, R, 3 * g.N, nb, st, vflag);
    RT_HIP(hipGetLastError());
}

// coarsest for a batch: fb[b][3r + c] = F_b at node r; U_b = Ainv fb[b]
template <typename T>
__global__ void k_gather_rows_b(const int* __restrict__ rowslot, int nrows, const T* __restrict__ F, int N, int bs,
                                int nb, double* __restrict__ fb) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int s = rowslot[r];
    for (int b = 0; b < nb; b++) {
        const T* Fb = F + (size_t)b * bs;
        double* f = fb + (size_t)b * 3 * nrows;
        f[3 * r + 0] = (double)Fb[s];
        f[3 * r + 1] = (double)Fb[N + s];
        f[3 * r + 2] = (double)Fb[2 * N + s];
    }
}

template <typename T>
__global__ void __launch_bounds__(64) k_dense_gemv_b(const double* __restrict__ A, int n, const double* __restrict__ f,
                                                     const int* __restrict__ rowslot, T* __restrict__ U, int N, int bs,
                                                     int nb) {
    const int row = blockIdx.x;
    double s[kBatchMax];
#pragma unroll
    for (int b = 0; b < kBatchMax; b++) s[b] = 0;
    const double* a = A + (size_t)row * n;
    for (int j = threadIdx.x; j < n; j += 64) {
        const double aj = a[j];
#pragma unroll
        for (int b = 0; b < kBatchMax; b++)
            if (b < nb) s[b] += aj * f[(size_t)b * n + j];
    }
#pragma unroll
    for (int b = 0; b < kBatchMax; b++) {
        if (b >= nb) break;
        double x = s[b];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (threadIdx.x == 0) U[(size_t)b * bs + (size_t)(row % 3) * N + rowslot[row / 3]] = (T)x;
    }
}

template <typename T>
void launch_coarsest_batch(const double* Ainv, int n3, const int* rowslot, int nb, int N, const T* F, T* U,
                           double* fbuf, hipStream_t s) {
    if (nb < 1 || nb > kBatchMax) throw std::invalid_argument("batch size out of range");
    const int nrows = n3 / 3;
    if (nrows <= 0) return;
    hipLaunchKernelGGL((k_gather_rows_b<T>), dim3((nrows + 255) / 256), dim3(256), 0, s, rowslot, nrows, F, N, 3 * N,
                       nb, fbuf);
    hipLaunchKernelGGL((k_dense_gemv_b<T>), dim3(n3), dim3(64), 0, s, Ainv, n3, (const double*)fbuf, rowslot, U, N,
                       3 * N, nb);
    RT_HIP(hipGetLastError());
}

template void launch_coarse_gs_batch<float, float>(const Geom&, int, float*, const float*, const float*,
                                                   const uint8_t*, int, bool, hipStream_t);
template void launch_coarse_residual_batch<float, float>(const Geom&, int, const float*, const float*, float*,
                                                         const float*, const uint8_t*, hipStream_t);
template void launch_coarsest_batch<float>(const double*, int, const int*, int, int, const float*, float*, double*,
                                           hipStream_t);

}  // namespace rt
