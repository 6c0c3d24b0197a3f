// Finest-level nodal Gauss-Seidel on the octant layout (see rt_internal.h).
//
// For a node v the reference forms, element by element (Grid.cu:922-1066),
//     KeU_v = sum_{e ni v} rho_e^p * sum_{j in e, j != v} KE[v_loc(e)][j_loc(e)] u_j
//     S_v   = sum_{e ni v} rho_e^p * KE[v_loc(e)][v_loc(e)]
// where the element incident at w = (w%2-1, w/2%2-1, w/4-1) holds v as local
// vertex 7-w; support ("fixed") neighbours contribute nothing, a fixed node
// gets S = n_e I and KeU = 0; then x, y, z are relaxed in turn.
//
// One thread per node of the colour being relaxed; the colour is a template
// parameter, so each of the 27 neighbour slots and 8 element slots is the
// thread's base index plus a wave-uniform offset. Loads are raw buffer loads
// with that offset in a scalar register (no per-load address arithmetic). The
// loop runs over neighbours and accumulates into 8 per-element partials, so a
// neighbour's displacement is live only while it is used. KE enters as 10
// signed scalar constants (ke_pattern.h): every FMA takes its coefficient from
// an SGPR with a sign modifier.
#include "rt_internal.h"
#include "ke_pattern.h"

#include <algorithm>
#include <cmath>
#include <utility>

namespace rt {

// the smoother addresses u, rho^p and the flags through buffer resources whose
// byte ranges are 32-bit: a level whose vectors pass 2 GiB is refused instead of
// being read as zeros past the limit
template <typename T>
static void check_buffer_range(const Geom& g) {
    if ((long long)g.N * (long long)sizeof(T) >= (1ll << 31))
        throw std::invalid_argument(std::string("finest level of ") + std::to_string(g.N) +
                                    " slots is too large for 32-bit buffer offsets in this precision's smoother" +
                                    (sizeof(T) == 8 ? " (FP32 cycles reach twice the size)" : ""));
}


template <typename T> struct KeC { T c[RT_KE_NVAL]; };

static double g_kec[RT_KE_NVAL];

// the 10 distinct magnitudes of KE, read off the table (validated entry by entry)
void set_ke_constants(const double* ke576, hipStream_t) {
    bool seen[RT_KE_NVAL] = {false};
    for (int i = 0; i < 576; i++) {
        const int c = rt_ke_code[i];
        if (c && !seen[std::abs(c) - 1]) { g_kec[std::abs(c) - 1] = c > 0 ? ke576[i] : -ke576[i]; seen[std::abs(c) - 1] = true; }
    }
    double scale = 0;
    for (int i = 0; i < 576; i++) scale = std::max(scale, std::fabs(ke576[i]));
    for (int i = 0; i < 576; i++) {
        const int c = rt_ke_code[i];
        const double v = c == 0 ? 0.0 : (c > 0 ? g_kec[c - 1] : -g_kec[-c - 1]);
        if (std::fabs(v - ke576[i]) > 1e-12 * scale)
            throw std::runtime_error("element matrix does not follow the isotropic hexahedron pattern");
    }
}

template <typename T>
static KeC<T> kec() {
    KeC<T> k;
    for (int i = 0; i < RT_KE_NVAL; i++) k.c[i] = (T)g_kec[i];
    return k;
}

// offset (in slots, shifted to be >= 0) of the point at (DX,DY,DZ) in {-1,0,1}^3
// from a colour-C node whose half-index base is b: slot = (b - hxy - hx - 1) + off
template <int C, int DX, int DY, int DZ>
__device__ __forceinline__ int rel_off(const Geom& g) {
    constexpr int cx = C & 1, cy = (C >> 1) & 1, cz = C >> 2;
    constexpr int col = C ^ ((DX & 1) | ((DY & 1) << 1) | ((DZ & 1) << 2));
    constexpr int ox = (cx + DX) >> 1, oy = (cy + DY) >> 1, oz = (cz + DZ) >> 1;
    return col * g.cs + ((oz + 1) * g.h[1] + (oy + 1)) * g.h[0] + (ox + 1);
}

constexpr int ke_code_at(int i) { return rt_ke_code[i]; }

template <int CODE, typename T>
__device__ __forceinline__ T kc(const KeC<T>& K) {
    static_assert(CODE != 0, "isotropic KE has no zero entry");
    if constexpr (CODE > 0) return K.c[CODE - 1];
    else return -K.c[-CODE - 1];
}

// acc[W] += KE[vi rows][vj cols] u_nb for neighbour NB (position NB%3, NB/3%3,
// NB/9 in the 3x3x3 block around the node) if it is vertex vj != vi of element W
template <typename T, int NB, int W>
__device__ __forceinline__ void nb_elem(T (&acc)[8][3], const KeC<T>& K, T u0, T u1, T u2) {
    constexpr int px = NB % 3, py = NB / 3 % 3, pz = NB / 9;
    constexpr int jx = px - (W & 1), jy = py - ((W >> 1) & 1), jz = pz - (W >> 2);
    if constexpr (jx >= 0 && jx <= 1 && jy >= 0 && jy <= 1 && jz >= 0 && jz <= 1) {
        constexpr int vj = jx + 2 * jy + 4 * jz, vi = 7 - W;
        if constexpr (vj != vi) {
            acc[W][0] += kc<ke_code_at((3 * vi + 0) * 24 + 3 * vj + 0)>(K) * u0;
            acc[W][0] += kc<ke_code_at((3 * vi + 0) * 24 + 3 * vj + 1)>(K) * u1;
            acc[W][0] += kc<ke_code_at((3 * vi + 0) * 24 + 3 * vj + 2)>(K) * u2;
            acc[W][1] += kc<ke_code_at((3 * vi + 1) * 24 + 3 * vj + 0)>(K) * u0;
            acc[W][1] += kc<ke_code_at((3 * vi + 1) * 24 + 3 * vj + 1)>(K) * u1;
            acc[W][1] += kc<ke_code_at((3 * vi + 1) * 24 + 3 * vj + 2)>(K) * u2;
            acc[W][2] += kc<ke_code_at((3 * vi + 2) * 24 + 3 * vj + 0)>(K) * u0;
            acc[W][2] += kc<ke_code_at((3 * vi + 2) * 24 + 3 * vj + 1)>(K) * u1;
            acc[W][2] += kc<ke_code_at((3 * vi + 2) * 24 + 3 * vj + 2)>(K) * u2;
        }
    }
}

template <typename T, int NB, int... Ws>
__device__ __forceinline__ void nb_elems(T (&acc)[8][3], const KeC<T>& K, T u0, T u1, T u2,
                                         std::integer_sequence<int, Ws...>) {
    (nb_elem<T, NB, Ws>(acc, K, u0, u1, u2), ...);
}

struct GsBufs {
    __amdgpu_buffer_rsrc_t u0, u1, u2, rho, flag;
};

// colour of the neighbour at (DX, DY, DZ) of a colour-C node
template <int C, int DX, int DY, int DZ>
constexpr int nb_colour() { return C ^ ((DX & 1) | ((DY & 1) << 1) | ((DZ & 1) << 2)); }

// Z0 ("zero start"): the sweep is the first of a forward cycle from u = 0, so
// a neighbour whose colour comes later in the order 0..7 still holds 0 and is
// neither loaded nor multiplied (adding KE * 0 leaves the sums unchanged).
template <typename T, int C, bool CHECK_FIX, int NB, bool Z0 = false>
__device__ __forceinline__ void nb_step(const Geom& g, int vb, const KeC<T>& K, const GsBufs& B, T (&acc)[8][3]) {
    if constexpr (NB != 13 && !(Z0 && nb_colour<C, NB % 3 - 1, NB / 3 % 3 - 1, NB / 9 - 1>() > C)) {
        const int so = rel_off<C, NB % 3 - 1, NB / 3 % 3 - 1, NB / 9 - 1>(g);
        const int vo = vb * (int)sizeof(T), sob = so * (int)sizeof(T);
        T u0 = bload<T>(B.u0, vo, sob);
        T u1 = bload<T>(B.u1, vo, sob);
        T u2 = bload<T>(B.u2, vo, sob);
        if constexpr (CHECK_FIX) {
            if (bload_u8(B.flag, vb, so) & NF_FIXED) { u0 = T(0); u1 = T(0); u2 = T(0); }
        }
        nb_elems<T, NB>(acc, K, u0, u1, u2, std::make_integer_sequence<int, 8>{});
    }
}

// two consecutive values of one stream with one load (dword-aligned 8- or 16-byte
// buffer load)
template <typename T> struct Two { T lo, hi; };
template <typename T> __device__ __forceinline__ Two<T> bload2(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ Two<float> bload2<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(Two<float>, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ Two<double> bload2<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(Two<double>, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

// The x-neighbours (dx = -1, +1) of a row (dy, dz) have the opposite x parity,
// so both sit in one colour block at consecutive slots (offset of dx = -1, then
// + 1): one two-wide load fetches both. The dx = 0 neighbour is a single load.
template <typename T, int C, bool CHECK_FIX, int ROW, bool Z0 = false>
__device__ __forceinline__ void nb_row(const Geom& g, int vb, const KeC<T>& K, const GsBufs& B, T (&acc)[8][3]) {
    constexpr int DY = ROW % 3 - 1, DZ = ROW / 3 - 1;
    constexpr int NBM = 9 * (DZ + 1) + 3 * (DY + 1);   // dx = -1; dx = 0, +1 follow
    const int vo = vb * (int)sizeof(T);
    if constexpr (!(Z0 && nb_colour<C, -1, DY, DZ>() > C)) {
        const int so = rel_off<C, -1, DY, DZ>(g), sob = so * (int)sizeof(T);
        const Two<T> a = bload2<T>(B.u0, vo, sob), b = bload2<T>(B.u1, vo, sob), c = bload2<T>(B.u2, vo, sob);
        T l0 = a.lo, l1 = b.lo, l2 = c.lo, r0 = a.hi, r1 = b.hi, r2 = c.hi;
        if constexpr (CHECK_FIX) {
            if (bload_u8(B.flag, vb, so) & NF_FIXED) { l0 = T(0); l1 = T(0); l2 = T(0); }
            if (bload_u8(B.flag, vb, so + 1) & NF_FIXED) { r0 = T(0); r1 = T(0); r2 = T(0); }
        }
        nb_elems<T, NBM>(acc, K, l0, l1, l2, std::make_integer_sequence<int, 8>{});
        nb_elems<T, NBM + 2>(acc, K, r0, r1, r2, std::make_integer_sequence<int, 8>{});
    }
    nb_step<T, C, CHECK_FIX, NBM + 1, Z0>(g, vb, K, B, acc);   // skips the node itself
}

template <typename T, int C, bool CHECK_FIX, bool Z0, int... ROWs>
__device__ __forceinline__ void nb_all(const Geom& g, int vb, const KeC<T>& K, const GsBufs& B, T (&acc)[8][3],
                                       std::integer_sequence<int, ROWs...>) {
    (nb_row<T, C, CHECK_FIX, ROWs, Z0>(g, vb, K, B, acc), ...);
}

template <typename T, int C, int W>
__device__ __forceinline__ void elem_step(const Geom& g, int vb, const KeC<T>& K, const GsBufs& B,
                                          const T (&acc)[8][3], T KU[3], T S[9]) {
    constexpr int vi = 7 - W;
    const int so = rel_off<C, (W & 1) - 1, ((W >> 1) & 1) - 1, (W >> 2) - 1>(g);
    const T pen = (T)bload<float>(B.rho, vb * 4, so * 4);
    KU[0] += pen * acc[W][0];
    KU[1] += pen * acc[W][1];
    KU[2] += pen * acc[W][2];
    S[0] += pen * kc<ke_code_at((3 * vi + 0) * 24 + 3 * vi + 0)>(K);
    S[1] += pen * kc<ke_code_at((3 * vi + 0) * 24 + 3 * vi + 1)>(K);
    S[2] += pen * kc<ke_code_at((3 * vi + 0) * 24 + 3 * vi + 2)>(K);
    S[3] += pen * kc<ke_code_at((3 * vi + 1) * 24 + 3 * vi + 0)>(K);
    S[4] += pen * kc<ke_code_at((3 * vi + 1) * 24 + 3 * vi + 1)>(K);
    S[5] += pen * kc<ke_code_at((3 * vi + 1) * 24 + 3 * vi + 2)>(K);
    S[6] += pen * kc<ke_code_at((3 * vi + 2) * 24 + 3 * vi + 0)>(K);
    S[7] += pen * kc<ke_code_at((3 * vi + 2) * 24 + 3 * vi + 1)>(K);
    S[8] += pen * kc<ke_code_at((3 * vi + 2) * 24 + 3 * vi + 2)>(K);
}

template <typename T, int C, int... Ws>
__device__ __forceinline__ void elem_all(const Geom& g, int vb, const KeC<T>& K, const GsBufs& B, const T (&acc)[8][3],
                                         T KU[3], T S[9], std::integer_sequence<int, Ws...>) {
    (elem_step<T, C, Ws>(g, vb, K, B, acc, KU, S), ...);
}

template <typename T, int C, bool CHECK_FIX, bool Z0 = false>
__device__ __forceinline__ void gs_node(const Geom& g, int base, const KeC<T>& K, const GsBufs& B, T KU[3], T S[9]) {
    const int vb = base - g.h[0] * g.h[1] - g.h[0] - 1;
    T acc[8][3];
#pragma unroll
    for (int w = 0; w < 8; w++) acc[w][0] = acc[w][1] = acc[w][2] = T(0);
    nb_all<T, C, CHECK_FIX, Z0>(g, vb, K, B, acc, std::make_integer_sequence<int, 9>{});
    KU[0] = KU[1] = KU[2] = T(0);
#pragma unroll
    for (int i = 0; i < 9; i++) S[i] = T(0);
    elem_all<T, C>(g, vb, K, B, acc, KU, S, std::make_integer_sequence<int, 8>{});
}

// the three divisions of a node update; FP32 optionally by the hardware
// reciprocal (1 ulp) times the numerator instead of the IEEE division sequence
#ifndef RT_GS_FAST_DIV
#define RT_GS_FAST_DIV 0
#endif
template <typename T> __device__ __forceinline__ T gs_div(T a, T b) { return a / b; }
template <> __device__ __forceinline__ float gs_div<float>(float a, float b) {
    return RT_GS_FAST_DIV ? a * __builtin_amdgcn_rcpf(b) : a / b;
}

// GS update of the colour-C node with half-index base (Grid.cu:1007-1062)
// ZS ("zero supports"): the caller guarantees u = 0 at every support node for
// the whole sweep (f = 0 there, so the relaxation keeps it 0, and the
// prolongation skips supports), so a neighbour never needs its flag checked.
// Z0: the node's own old value is 0 as well, and an inactive node, which the
// sweep otherwise leaves alone, gets the 0 that a cleared vector would hold.
// fu: f . u_new of the node is added to it (the last sweep of a cycle, whose
// values are final: the preconditioned residual's r . z without another pass).
template <typename T, int C, bool ZS, bool Z0 = false>
__device__ __forceinline__ void gs_update(const Geom& g, int base, Vec3<T> U, Vec3<const T> F,
                                          const uint8_t* __restrict__ vflag, const KeC<T>& K, const GsBufs& B,
                                          bool reverse, double* fu = nullptr) {
    const int v = C * g.cs + base;
    const uint8_t fl = vflag[v];
    if (!(fl & NF_ACTIVE)) {
        if constexpr (Z0) { U.c[0][v] = T(0); U.c[1][v] = T(0); U.c[2][v] = T(0); }
        return;
    }
    const T f[3] = {F.c[0][v], F.c[1][v], F.c[2][v]};
    if (fl & NF_FIXED) {
        // S = n_e I, KeU = 0 (Grid.cu:1007-1014)
        const T ne = (T)(fl >> 4);
        const T u0 = f[0] / ne, u1 = f[1] / ne, u2 = f[2] / ne;
        U.c[0][v] = u0; U.c[1][v] = u1; U.c[2][v] = u2;
        if (fu) *fu += ((double)f[0] * (double)u0 + (double)f[1] * (double)u1) + (double)f[2] * (double)u2;
        return;
    }
    T KU[3], S[9];
    if (!ZS && (fl & NF_NEARFIX)) gs_node<T, C, true, Z0>(g, base, K, B, KU, S);
    else gs_node<T, C, false, Z0>(g, base, K, B, KU, S);
    T u[3];
    if constexpr (Z0) {
        u[0] = u[1] = u[2] = T(0);
    } else {
        u[0] = U.c[0][v]; u[1] = U.c[1][v]; u[2] = U.c[2][v];
    }
    // component-wise Gauss-Seidel with the 3x3 self block (Grid.cu:1056-1062)
    if (!reverse) {
        u[0] = gs_div<T>(f[0] - S[1] * u[1] - S[2] * u[2] - KU[0], S[0]);
        u[1] = gs_div<T>(f[1] - S[3] * u[0] - S[5] * u[2] - KU[1], S[4]);
        u[2] = gs_div<T>(f[2] - S[6] * u[0] - S[7] * u[1] - KU[2], S[8]);
    } else {
        u[2] = gs_div<T>(f[2] - S[6] * u[0] - S[7] * u[1] - KU[2], S[8]);
        u[1] = gs_div<T>(f[1] - S[3] * u[0] - S[5] * u[2] - KU[1], S[4]);
        u[0] = gs_div<T>(f[0] - S[1] * u[1] - S[2] * u[2] - KU[0], S[0]);
    }
    U.c[0][v] = u[0]; U.c[1][v] = u[1]; U.c[2][v] = u[2];
    if (fu) *fu += ((double)f[0] * (double)u[0] + (double)f[1] * (double)u[1]) + (double)f[2] * (double)u[2];
}

// Colours C and C^1 (same y, z parity, opposite x parity) couple only through
// the x-neighbours in the same row, so one workgroup owns a whole row
// (y, z) and relaxes its even-x nodes, then (after a barrier, which makes the
// new values visible to the block) its odd-x nodes; REV swaps the two.
// Every other neighbour belongs to a row of another colour pair and is
// read-only during the launch. Four launches make a sweep, in the reference's
// colour order 0..7 (or 7..0).
template <typename T, int C, bool ZS, bool Z0 = false>
__device__ __forceinline__ void gs_row_phase(const Geom& g, int j, int k, Vec3<T> U, Vec3<const T> F,
                                             const uint8_t* __restrict__ vflag, const KeC<T>& K, const GsBufs& B,
                                             bool reverse, double* fu = nullptr) {
    const int nhx = (g.n[0] - (C & 1) + 1) >> 1;
    const int rowbase = ((k + 1) * g.h[1] + (j + 1)) * g.h[0] + 1;
    for (int i = threadIdx.x; i < nhx; i += blockDim.x)
        gs_update<T, C, ZS, Z0>(g, rowbase + i, U, F, vflag, K, B, reverse, fu);
}

// sum of a double over the workgroup, valid in thread 0
__device__ __forceinline__ double gs_blk_sum(double v) {
    __shared__ double red[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)((blockDim.x + 63) >> 6); i++) t += red[i];
    return t;
}

// rows relaxed one after another by each workgroup of the smoother (rows of one
// launch never neighbour each other). Measured at 511 x 256 x 256 elements: one
// row per workgroup 1.06 ms per sweep, two 1.13 ms, four 1.24 ms. Taking the
// rows of a launch in compact (y, z) tiles (4x32 to 64x4 rows) instead of
// y-fastest order changed the sweep by less than 1 %, and loading both
// phases' node flags before the first phase (so no phase waits on its flag)
// changed a launch by 0.1 %.
#ifndef RT_GS_ROWS
#define RT_GS_ROWS 1
#endif

// optional occupancy target (waves per SIMD) for the smoother, a build-time knob
#if defined(RT_GS_WAVES_PER_EU) && RT_GS_WAVES_PER_EU > 0
#define RT_GS_ATTR __attribute__((amdgpu_waves_per_eu(RT_GS_WAVES_PER_EU)))
#else
#define RT_GS_ATTR
#endif

// DOT: partial sums of f . u_new over each workgroup's rows, one per workgroup
// in fu_part[blockIdx.x]
template <typename T, int CY, int CZ, bool REV, bool ZS, bool Z0 = false, bool DOT = false>
__global__ void __launch_bounds__(256) RT_GS_ATTR k_fine_gs_pair(Geom g, Vec3<T> U, Vec3<const T> F,
                                                      const float* __restrict__ rhop,
                                                      const uint8_t* __restrict__ vflag, KeC<T> K, int nrows,
                                                      double* __restrict__ fu_part = nullptr) {
    static_assert(!(Z0 && REV), "a zero-start sweep runs forward");
    double fu = 0;
    double* fup = DOT ? &fu : nullptr;
    const int nhy = (g.n[1] - CY + 1) >> 1;
    const int first = xcd_row(blockIdx.x, gridDim.x) * RT_GS_ROWS;
    const int nbytes = g.N * (int)sizeof(T);
    const GsBufs B = {rsrc(U.c[0], nbytes), rsrc(U.c[1], nbytes), rsrc(U.c[2], nbytes), rsrc(rhop, g.N * 4),
                      rsrc(vflag, g.N)};
    constexpr int CA = (REV ? 1 : 0) | (CY << 1) | (CZ << 2);
    for (int r = 0; r < RT_GS_ROWS; r++) {
        const int row = first + r;   // uniform across the workgroup
        if (row >= nrows) break;
        const int j = row % nhy, k = row / nhy;
        gs_row_phase<T, CA, ZS, Z0>(g, j, k, U, F, vflag, K, B, REV, fup);
        __syncthreads();
        gs_row_phase<T, CA ^ 1, ZS, Z0>(g, j, k, U, F, vflag, K, B, REV, fup);
    }
    if constexpr (DOT) {
        const double t = gs_blk_sum(fu);
        if (threadIdx.x == 0) fu_part[blockIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// Tiled sweep: two launches per sweep instead of four, reading the old values
// from Uin and writing the new ones to Uout (a second buffer).
//
// In the reference's colour order all colours of the z-parity class relaxed
// first (z even forward, z odd in reverse) come before any colour of the other
// class, and planes of one class never neighbour each other. So a launch
// relaxes the four colours of one class, plane by plane, and a workgroup owns a
// tile of TR whole rows (all x) of one plane. Inside the tile the four colours
// follow the sweep order with a barrier between them; a neighbour in the same
// plane is new (read from Uout, written earlier by this workgroup) if its
// colour comes earlier, old (read from Uin) otherwise; a neighbour in the other
// z class is old in the first launch and new (Uout, written by the first
// launch) in the second. The second pair of colours of the plane (the "lag"
// rows, y parity relaxed last) needs the new values of the first pair on the
// rows on either side, so the first pair is also relaxed on the one row of the
// next tile it needs (below the tile in reverse order, above it forward). That
// row is computed from the same values and with the same arithmetic as its
// owner computes it, so both write identical bits. Uin is never written, so
// the result does not depend on the order in which workgroups run, and each
// node gets exactly the value of the colour-pair sweep above.
// ---------------------------------------------------------------------------
struct GsBufs2 {
    __amdgpu_buffer_rsrc_t in0, in1, in2, out0, out1, out2, rho, flag;
};

// whether the neighbour at (DX, DY, DZ) of a colour-C node already holds its
// new value when the node is relaxed (FIRST: this launch relaxes the first z
// class of the sweep; REV: reverse colour order)
template <int C, int DX, int DY, int DZ, bool FIRST, bool REV>
constexpr bool nb_is_new() {
    if constexpr (DZ != 0) {
        return !FIRST;
    } else {
        constexpr int cn = C ^ ((DX & 1) | ((DY & 1) << 1));
        constexpr int o = ((C >> 1) & 1) * 2 + (C & 1), on = ((cn >> 1) & 1) * 2 + (cn & 1);
        return REV ? (on > o) : (on < o);
    }
}

template <typename T, int C, bool CHECK_FIX, int NB, bool FIRST, bool REV>
__device__ __forceinline__ void nb_step2(const Geom& g, int vb, const KeC<T>& K, const GsBufs2& B, T (&acc)[8][3]) {
    if constexpr (NB != 13) {
        constexpr int DX = NB % 3 - 1, DY = NB / 3 % 3 - 1, DZ = NB / 9 - 1;
        constexpr bool NW = nb_is_new<C, DX, DY, DZ, FIRST, REV>();
        const int so = rel_off<C, DX, DY, DZ>(g);
        const int vo = vb * (int)sizeof(T), sob = so * (int)sizeof(T);
        T u0 = bload<T>(NW ? B.out0 : B.in0, vo, sob);
        T u1 = bload<T>(NW ? B.out1 : B.in1, vo, sob);
        T u2 = bload<T>(NW ? B.out2 : B.in2, vo, sob);
        if constexpr (CHECK_FIX) {
            if (bload_u8(B.flag, vb, so) & NF_FIXED) { u0 = T(0); u1 = T(0); u2 = T(0); }
        }
        nb_elems<T, NB>(acc, K, u0, u1, u2, std::make_integer_sequence<int, 8>{});
    }
}

template <typename T, int C, bool CHECK_FIX, int ROW, bool FIRST, bool REV>
__device__ __forceinline__ void nb_row2(const Geom& g, int vb, const KeC<T>& K, const GsBufs2& B, T (&acc)[8][3]) {
    constexpr int DY = ROW % 3 - 1, DZ = ROW / 3 - 1;
    constexpr int NBM = 9 * (DZ + 1) + 3 * (DY + 1);
    constexpr bool NW = nb_is_new<C, -1, DY, DZ, FIRST, REV>();   // dx = -1 and +1 share a colour
    const int vo = vb * (int)sizeof(T);
    {
        const int so = rel_off<C, -1, DY, DZ>(g), sob = so * (int)sizeof(T);
        const Two<T> a = bload2<T>(NW ? B.out0 : B.in0, vo, sob), b = bload2<T>(NW ? B.out1 : B.in1, vo, sob),
                     c = bload2<T>(NW ? B.out2 : B.in2, vo, sob);
        T l0 = a.lo, l1 = b.lo, l2 = c.lo, r0 = a.hi, r1 = b.hi, r2 = c.hi;
        if constexpr (CHECK_FIX) {
            if (bload_u8(B.flag, vb, so) & NF_FIXED) { l0 = T(0); l1 = T(0); l2 = T(0); }
            if (bload_u8(B.flag, vb, so + 1) & NF_FIXED) { r0 = T(0); r1 = T(0); r2 = T(0); }
        }
        nb_elems<T, NBM>(acc, K, l0, l1, l2, std::make_integer_sequence<int, 8>{});
        nb_elems<T, NBM + 2>(acc, K, r0, r1, r2, std::make_integer_sequence<int, 8>{});
    }
    nb_step2<T, C, CHECK_FIX, NBM + 1, FIRST, REV>(g, vb, K, B, acc);
}

template <typename T, int C, bool CHECK_FIX, bool FIRST, bool REV, int... ROWs>
__device__ __forceinline__ void nb_all2(const Geom& g, int vb, const KeC<T>& K, const GsBufs2& B, T (&acc)[8][3],
                                        std::integer_sequence<int, ROWs...>) {
    (nb_row2<T, C, CHECK_FIX, ROWs, FIRST, REV>(g, vb, K, B, acc), ...);
}

template <typename T, int C, bool CHECK_FIX, bool FIRST, bool REV>
__device__ __forceinline__ void gs_node2(const Geom& g, int base, const KeC<T>& K, const GsBufs2& B, T KU[3],
                                         T S[9]) {
    const int vb = base - g.h[0] * g.h[1] - g.h[0] - 1;
    T acc[8][3];
#pragma unroll
    for (int w = 0; w < 8; w++) acc[w][0] = acc[w][1] = acc[w][2] = T(0);
    nb_all2<T, C, CHECK_FIX, FIRST, REV>(g, vb, K, B, acc, std::make_integer_sequence<int, 9>{});
    KU[0] = KU[1] = KU[2] = T(0);
#pragma unroll
    for (int i = 0; i < 9; i++) S[i] = T(0);
    const GsBufs Be = {B.in0, B.in1, B.in2, B.rho, B.flag};
    elem_all<T, C>(g, vb, K, Be, acc, KU, S, std::make_integer_sequence<int, 8>{});
}

// the GS update of gs_update, old values from Uin, the result to Uout
template <typename T, int C, bool ZS, bool FIRST, bool REV>
__device__ __forceinline__ void gs_update2(const Geom& g, int base, Vec3<const T> Uin, Vec3<T> Uout, Vec3<const T> F,
                                           const uint8_t* __restrict__ vflag, const KeC<T>& K, const GsBufs2& B) {
    const int v = C * g.cs + base;
    const uint8_t fl = vflag[v];
    if (!(fl & NF_ACTIVE)) return;
    const T f[3] = {F.c[0][v], F.c[1][v], F.c[2][v]};
    if (fl & NF_FIXED) {
        const T ne = (T)(fl >> 4);
        Uout.c[0][v] = f[0] / ne; Uout.c[1][v] = f[1] / ne; Uout.c[2][v] = f[2] / ne;
        return;
    }
    T KU[3], S[9];
    if (!ZS && (fl & NF_NEARFIX)) gs_node2<T, C, true, FIRST, REV>(g, base, K, B, KU, S);
    else gs_node2<T, C, false, FIRST, REV>(g, base, K, B, KU, S);
    T u[3] = {Uin.c[0][v], Uin.c[1][v], Uin.c[2][v]};
    if (!REV) {
        u[0] = gs_div<T>(f[0] - S[1] * u[1] - S[2] * u[2] - KU[0], S[0]);
        u[1] = gs_div<T>(f[1] - S[3] * u[0] - S[5] * u[2] - KU[1], S[4]);
        u[2] = gs_div<T>(f[2] - S[6] * u[0] - S[7] * u[1] - KU[2], S[8]);
    } else {
        u[2] = gs_div<T>(f[2] - S[6] * u[0] - S[7] * u[1] - KU[2], S[8]);
        u[1] = gs_div<T>(f[1] - S[3] * u[0] - S[5] * u[2] - KU[1], S[4]);
        u[0] = gs_div<T>(f[0] - S[1] * u[1] - S[2] * u[2] - KU[0], S[0]);
    }
    Uout.c[0][v] = u[0]; Uout.c[1][v] = u[1]; Uout.c[2][v] = u[2];
}

// colour-C nodes of the rows y in [ya, yb) (clipped to the box) of plane
// half-index k, spread over the workgroup
template <typename T, int C, bool ZS, bool FIRST, bool REV>
__device__ __forceinline__ void tile_phase(const Geom& g, int k, int ya, int yb, Vec3<const T> Uin, Vec3<T> Uout,
                                           Vec3<const T> F, const uint8_t* __restrict__ vflag, const KeC<T>& K,
                                           const GsBufs2& B) {
    constexpr int cx = C & 1, cy = (C >> 1) & 1;
    const int nhx = (g.n[0] - cx + 1) >> 1, nhy = (g.n[1] - cy + 1) >> 1;
    const int jlo = max(0, (ya - cy + 1) >> 1), jhi = min(nhy, (yb - cy + 1) >> 1);
    const int cnt = (jhi - jlo) * nhx;
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
        const int j = jlo + t / nhx, i = t % nhx;
        gs_update2<T, C, ZS, FIRST, REV>(g, ((k + 1) * g.h[1] + (j + 1)) * g.h[0] + (i + 1), Uin, Uout, F, vflag, K,
                                         B);
    }
}

// rows per tile. Measured at 511 x 256 x 256 elements (ms per sweep, against
// 1.06 for the colour-pair launches): 4 rows 1.41, 8 rows 1.38, 16 rows 1.44,
// 32 rows 1.56; tiles ordered y-fastest instead of plane-fastest 1.50. Fewer,
// longer-lived workgroups with three barriers each hide memory latency worse
// than the colour-pair launches' one row per workgroup, although they move
// less data, so this sweep is not the default.
#ifndef RT_GS_TILE_ROWS
#define RT_GS_TILE_ROWS 8
#endif

template <typename T, int CZ, bool REV, bool ZS>
__global__ void __launch_bounds__(256) k_fine_gs_tile(Geom g, Vec3<const T> Uin, Vec3<T> Uout, Vec3<const T> F,
                                                      const float* __restrict__ rhop,
                                                      const uint8_t* __restrict__ vflag, KeC<T> K, int nplanes) {
    constexpr bool FIRST = CZ == (REV ? 1 : 0);
    constexpr int TR = RT_GS_TILE_ROWS;
    static_assert(TR % 2 == 0, "tiles start on even rows");
    // planes fastest, so that the tiles sharing the other class's planes in
    // between run close together on one XCD
    const int b = xcd_row(blockIdx.x, gridDim.x);
    const int k = b % nplanes, ty = b / nplanes;
    const int ny = g.n[1];
    const int y0 = ty * TR, y1 = min(y0 + TR, ny);
    const int nbytes = g.N * (int)sizeof(T);
    const GsBufs2 B = {rsrc(Uin.c[0], nbytes),  rsrc(Uin.c[1], nbytes),  rsrc(Uin.c[2], nbytes),
                       rsrc(Uout.c[0], nbytes), rsrc(Uout.c[1], nbytes), rsrc(Uout.c[2], nbytes),
                       rsrc(rhop, g.N * 4),     rsrc(vflag, g.N)};
    // lead rows (the y parity relaxed first) with the one extra row the lag rows need
    const int la = REV ? y0 - 1 : y0, lb = REV ? y1 : y1 + 1;
    constexpr int Z = CZ << 2;
    if constexpr (!REV) {
        tile_phase<T, Z | 0, ZS, FIRST, REV>(g, k, la, lb, Uin, Uout, F, vflag, K, B);
        __syncthreads();
        tile_phase<T, Z | 1, ZS, FIRST, REV>(g, k, la, lb, Uin, Uout, F, vflag, K, B);
        __syncthreads();
        tile_phase<T, Z | 2, ZS, FIRST, REV>(g, k, y0, y1, Uin, Uout, F, vflag, K, B);
        __syncthreads();
        tile_phase<T, Z | 3, ZS, FIRST, REV>(g, k, y0, y1, Uin, Uout, F, vflag, K, B);
    } else {
        tile_phase<T, Z | 3, ZS, FIRST, REV>(g, k, la, lb, Uin, Uout, F, vflag, K, B);
        __syncthreads();
        tile_phase<T, Z | 2, ZS, FIRST, REV>(g, k, la, lb, Uin, Uout, F, vflag, K, B);
        __syncthreads();
        tile_phase<T, Z | 1, ZS, FIRST, REV>(g, k, y0, y1, Uin, Uout, F, vflag, K, B);
        __syncthreads();
        tile_phase<T, Z | 0, ZS, FIRST, REV>(g, k, y0, y1, Uin, Uout, F, vflag, K, B);
    }
}

// one full sweep (colours 0..7, or 7..0 when reverse) from Uin to Uout
template <typename T>
void launch_fine_gs_sweep2(const Geom& g, Vec3<const T> Uin, Vec3<T> Uout, Vec3<const T> F, const float* rhop,
                           const uint8_t* vflag, bool reverse, bool zero_support, hipStream_t s) {
    check_buffer_range<T>(g);
    const KeC<T> K = kec<T>();
    const int ntile = (g.n[1] + RT_GS_TILE_ROWS - 1) / RT_GS_TILE_ROWS;
    for (int pass = 0; pass < 2; pass++) {
        const int cz = reverse ? 1 - pass : pass;
        const int nplanes = (g.n[2] - cz + 1) >> 1;
        if (nplanes <= 0) continue;
        dim3 grid((unsigned)(nplanes * ntile));
        switch (cz * 4 + (reverse ? 2 : 0) + (zero_support ? 1 : 0)) {
#define RT_CASE(c, r, z) case c * 4 + r * 2 + z: hipLaunchKernelGGL((k_fine_gs_tile<T, c, (r) != 0, (z) != 0>), grid, dim3(256), 0, s, g, Uin, Uout, F, rhop, vflag, K, nplanes); break;
            RT_CASE(0, 0, 0) RT_CASE(0, 0, 1) RT_CASE(0, 1, 0) RT_CASE(0, 1, 1)
            RT_CASE(1, 0, 0) RT_CASE(1, 0, 1) RT_CASE(1, 1, 0) RT_CASE(1, 1, 1)
#undef RT_CASE
        }
        RT_HIP(hipGetLastError());
    }
}

template void launch_fine_gs_sweep2<float>(const Geom&, Vec3<const float>, Vec3<float>, Vec3<const float>,
                                           const float*, const uint8_t*, bool, bool, hipStream_t);
template void launch_fine_gs_sweep2<double>(const Geom&, Vec3<const double>, Vec3<double>, Vec3<const double>,
                                            const float*, const uint8_t*, bool, bool, hipStream_t);

// pair p = 0..3 relaxes colours (2p, 2p+1), or (2p+1, 2p) when reverse
// zero_start: the first forward sweep from u = 0 (with zero supports); U may
// hold anything on entry, and every slot the sweep visits is written
// workgroups of one colour-pair launch (the partials a fu launch writes)
int fine_gs_pair_blocks(const Geom& g, int pair) {
    const int cy = pair & 1, cz = pair >> 1;
    const long long rows = (long long)((g.n[1] - cy + 1) >> 1) * ((g.n[2] - cz + 1) >> 1);
    return rows <= 0 ? 0 : (int)((rows + RT_GS_ROWS - 1) / RT_GS_ROWS);
}

template <typename T>
void launch_fine_gs(const Geom& g, Vec3<T> U, Vec3<const T> F, const float* rhop, const uint8_t* vflag, int pair,
                    bool reverse, bool zero_support, hipStream_t s, bool zero_start, double* fu_part) {
    check_buffer_range<T>(g);
    const int cy = pair & 1, cz = pair >> 1;
    const long long rows = (long long)((g.n[1] - cy + 1) >> 1) * ((g.n[2] - cz + 1) >> 1);
    if (rows <= 0) return;
    dim3 grid((unsigned)((rows + RT_GS_ROWS - 1) / RT_GS_ROWS));
    const KeC<T> K = kec<T>();
    if (fu_part) {
        if (!reverse || !zero_support || zero_start)
            throw std::logic_error("f.u partials come from a reverse sweep with zero supports");
        switch (pair) {
#define RT_CASE(p) case p: hipLaunchKernelGGL((k_fine_gs_pair<T, ((p) & 1), ((p) >> 1), true, true, false, true>), grid, dim3(256), 0, s, g, U, F, rhop, vflag, K, (int)rows, fu_part); break;
            RT_CASE(0) RT_CASE(1) RT_CASE(2) RT_CASE(3)
#undef RT_CASE
        }
        RT_HIP(hipGetLastError());
        return;
    }
    if (zero_start) {
        if (reverse || !zero_support) throw std::logic_error("a zero-start sweep is forward with zero supports");
        switch (pair) {
#define RT_CASE(p) case p: hipLaunchKernelGGL((k_fine_gs_pair<T, ((p) & 1), ((p) >> 1), false, true, true>), grid, dim3(256), 0, s, g, U, F, rhop, vflag, K, (int)rows); break;
            RT_CASE(0) RT_CASE(1) RT_CASE(2) RT_CASE(3)
#undef RT_CASE
        }
        RT_HIP(hipGetLastError());
        return;
    }
    switch (pair * 4 + (reverse ? 2 : 0) + (zero_support ? 1 : 0)) {
#define RT_CASE(p, r, z) case p * 4 + r * 2 + z: hipLaunchKernelGGL((k_fine_gs_pair<T, ((p) & 1), ((p) >> 1), (r) != 0, (z) != 0>), grid, dim3(256), 0, s, g, U, F, rhop, vflag, K, (int)rows); break;
        RT_CASE(0, 0, 0) RT_CASE(0, 1, 0) RT_CASE(1, 0, 0) RT_CASE(1, 1, 0) RT_CASE(2, 0, 0) RT_CASE(2, 1, 0)
        RT_CASE(3, 0, 0) RT_CASE(3, 1, 0) RT_CASE(0, 0, 1) RT_CASE(0, 1, 1) RT_CASE(1, 0, 1) RT_CASE(1, 1, 1)
        RT_CASE(2, 0, 1) RT_CASE(2, 1, 1) RT_CASE(3, 0, 1) RT_CASE(3, 1, 1)
#undef RT_CASE
    }
    RT_HIP(hipGetLastError());
}

template void launch_fine_gs<float>(const Geom&, Vec3<float>, Vec3<const float>, const float*, const uint8_t*, int,
                                    bool, bool, hipStream_t, bool, double*);
template void launch_fine_gs<double>(const Geom&, Vec3<double>, Vec3<const double>, const float*, const uint8_t*, int,
                                     bool, bool, hipStream_t, bool, double*);

}  // namespace rt
