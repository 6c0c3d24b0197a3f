// Finest-level residual r = f - K u and apply y = K x, element-centric.
//
// Same operator as the node-centric form of Grid.cu:1450-1557 (WS rules:
// support nodes contribute no displacement, support rows give r = f / y = 0),
// evaluated element by element with the stiffness factorised in the
// tensor-product sum/difference basis: KE u = G^T (M (G u)), G = G1 (x) G1 (x) G1
// per displacement component, M with 45 nonzeros (modal_pattern.h). One element
// costs two 3-stage butterflies per component, 45 FMA and 21 scalings, against
// 576 FMA for KE u directly.
//
// Tiling: a workgroup owns a 31 x 7 column of nodes and marches through a
// chunk of Geom::zc z planes. Each step loads one node plane (33 x 9 with halo) into
// LDS, computes the 32 x 8 elements between the two resident planes (one per
// thread), stores their 24 vertex forces in LDS, and every node thread adds the
// four elements below and above it. A node plane is finished when both element
// planes around it are done; the partial from the lower element plane is kept
// in registers.
#include "rt_internal.h"
#include "modal_pattern.h"

namespace rt {

#ifndef RT_MODAL_XCD
#define RT_MODAL_XCD 1
#endif

#ifndef RT_MODAL_TX
#define RT_MODAL_TX 31
#endif
// z planes per workgroup (Geom::zc, default 24, chosen per hierarchy by the
// autotuner). Measured at 511 x 256 x 256 (ms per residual): 8 planes 0.81,
// 12 0.80, 16 0.81, 24 0.79, 32 0.86, 64 1.07; a 15 x 15 node tile instead of
// 31 x 7: 0.95.

namespace {
constexpr int TX = RT_MODAL_TX;            // node tile: TX x TY with (TX + 1)(TY + 1) = 256
constexpr int TY = 256 / (TX + 1) - 1;
constexpr int EX = TX + 1, EY = TY + 1;    // element tile (= 256 threads)
constexpr int PX = TX + 2, PY = TY + 2;    // node plane with halo
constexpr int PLANE = PX * PY;
static_assert(EX * EY == 256 && PLANE <= 512, "tile must fill 256 threads; a plane is staged by two points per thread");
}

template <typename T> struct ModalK { T m[RT_MODAL_NNZ]; };

template <typename T>
__device__ __forceinline__ void modal_apply(const T (&a)[24], T rp, const ModalK<T>& K, T (&y)[24]) {
    T t[24];
#pragma unroll
    for (int i = 0; i < 24; i++) t[i] = a[i];
    // forward butterflies on each component: bit b of the vertex index -> (sum, diff)
#pragma unroll
    for (int b = 1; b < 8; b <<= 1)
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (!(k & b)) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const T lo = t[3 * k + c], hi = t[3 * (k + b) + c];
                    t[3 * k + c] = lo + hi;
                    t[3 * (k + b) + c] = hi - lo;
                }
            }
    T w[24];
#pragma unroll
    for (int i = 0; i < 24; i++) w[i] = T(0);
#pragma unroll
    for (int n = 0; n < RT_MODAL_NNZ; n++) w[rt_modal_row[n]] += K.m[n] * t[rt_modal_col[n]];
#pragma unroll
    for (int i = 3; i < 24; i++) w[i] *= rp;
    // transposed butterflies: (w_s, w_d) -> (w_s - w_d, w_s + w_d)
#pragma unroll
    for (int b = 1; b < 8; b <<= 1)
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (!(k & b)) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const T s = w[3 * k + c], d = w[3 * (k + b) + c];
                    w[3 * k + c] = s - d;
                    w[3 * (k + b) + c] = s + d;
                }
            }
#pragma unroll
    for (int i = 0; i < 24; i++) y[i] = w[i];
}

__device__ __forceinline__ double blk_sum_m(double v) {
    __shared__ double red[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (threadIdx.x == 0) ? (red[0] + red[1]) + (red[2] + red[3]) : 0.0;
}

// MODE 0: R = F - K U.  MODE 1: R = K U and per-block partials of U.R.
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_fine_res_modal(Geom g, Vec3<const T> U, Vec3<const T> F, Vec3<T> R,
                                                        const float* __restrict__ rhop,
                                                        const uint8_t* __restrict__ vflag, ModalK<T> K,
                                                        double* __restrict__ partials) {
    constexpr bool APPLY = MODE != 0;
    __shared__ T su[2][3][PLANE];
    __shared__ uint8_t sf[2][PLANE];
    __shared__ T sy[24][256];

    const int tid = threadIdx.x;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    // Tile of this workgroup: the flattened block index is remapped so that
    // each XCD marches one contiguous range of tiles (x fastest), so the tiles
    // that share halo rows and partly used cache lines meet in one L2.
    const int gx = gridDim.x, gy = gridDim.y;
    const int bl = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int bm = RT_MODAL_XCD ? xcd_row(bl, gx * gy * (int)gridDim.z) : bl;
    if (g.tile_act && !g.tile_act[bm]) {   // no active node in this tile: nothing to write
        if (APPLY && partials && tid == 0) partials[bl] = 0.0;
        return;
    }
    const int x0 = (bm % gx) * TX, y0 = ((bm / gx) % gy) * TY, z0 = (bm / (gx * gy)) * g.zc;
    const int z1 = min(z0 + g.zc, nz);
    // element owned by this thread (tile-local i, j -> global x0 - 1 + i, y0 - 1 + j)
    const int ei = tid % EX, ej = tid / EX;
    const int gex = x0 - 1 + ei, gey = y0 - 1 + ej;
    const bool e_in = gex >= 0 && gey >= 0 && gex < nx - 1 && gey < ny - 1;
    // node owned by this thread
    const bool own = tid < TX * TY;
    const int na = tid % TX, nb = tid / TX;
    const int gnx = x0 + na, gny = y0 + nb;
    const bool n_in = own && gnx < nx && gny < ny;

    // Each thread stages plane points tid and tid + 256 (PLANE < 512). A plane's
    // flags and values are loaded together (absent and ghost slots hold zeros;
    // support values are zeroed once the flag is in), and the loads for plane
    // ez + 3 are issued at step ez, two steps before the plane is committed
    // to LDS, so their latency overlaps two steps of element work.
    struct Stage { T u[2][3]; uint8_t fl[2]; };
    auto fetch = [&](Stage& st, int z) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int p = tid + 256 * q;
            st.u[q][0] = st.u[q][1] = st.u[q][2] = T(0);
            st.fl[q] = 0;
            if (p >= PLANE) continue;
            const int px = x0 - 1 + p % PX, py = y0 - 1 + p / PX;
            if (px >= 0 && py >= 0 && z >= 0 && px < nx && py < ny && z < nz) {
                const int sl = oct_p(g, px, py, z);
                st.fl[q] = vflag[sl];
                st.u[q][0] = U.c[0][sl]; st.u[q][1] = U.c[1][sl]; st.u[q][2] = U.c[2][sl];
            }
        }
    };
    auto commit = [&](const Stage& st, int slot) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int p = tid + 256 * q;
            if (p >= PLANE) continue;
            const bool live = (st.fl[q] & NF_ACTIVE) && !(st.fl[q] & NF_FIXED);
            su[slot][0][p] = live ? st.u[q][0] : T(0);
            su[slot][1][p] = live ? st.u[q][1] : T(0);
            su[slot][2][p] = live ? st.u[q][2] : T(0);
            sf[slot][p] = st.fl[q];
        }
    };

    // rho^p of the thread's element and (residual) f of its output node are
    // loaded one step ahead of their use, like the planes above
    auto load_rp = [&](int z) -> float {
        return (e_in && z >= 0 && z < nz - 1) ? rhop[oct_p(g, gex, gey, z)] : 0.f;
    };
    auto load_f = [&](int z, T (&f)[3]) {
        f[0] = f[1] = f[2] = T(0);
        if (!APPLY && n_in && z >= z0 && z < z1) {
            const int s = oct_p(g, gnx, gny, z);
            f[0] = F.c[0][s]; f[1] = F.c[1][s]; f[2] = F.c[2][s];
        }
    };

    T part[3] = {T(0), T(0), T(0)};
    double dotacc = 0.0;
    int lo = 0;
    Stage sa, sb;
    fetch(sa, z0 - 1);
    commit(sa, lo);
    fetch(sa, z0);
    fetch(sb, z0 + 1);
    float rp_next = load_rp(z0 - 1);
    T f_next[3];
    load_f(z0 - 1, f_next);
    for (int ez = z0 - 1; ez < z1; ez++) {
        const int hi = lo ^ 1;
        commit(sa, hi);        // plane ez + 1
        sa = sb;
        fetch(sb, ez + 3);     // in flight during this step and the next
        const float rp = rp_next;
        rp_next = load_rp(ez + 1);
        const T fo[3] = {f_next[0], f_next[1], f_next[2]};   // f of node plane ez
        load_f(ez + 1, f_next);
        __syncthreads();
        // element (gex, gey, ez) between planes lo (kz = 0) and hi (kz = 1)
        T y[24];
        if (rp != 0.f) {
            T a[24];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int p = (ei + (k & 1)) + PX * (ej + ((k >> 1) & 1));
                const int sl = (k >> 2) ? hi : lo;
                a[3 * k + 0] = su[sl][0][p];
                a[3 * k + 1] = su[sl][1][p];
                a[3 * k + 2] = su[sl][2][p];
            }
            modal_apply<T>(a, (T)rp, K, y);
        } else {
#pragma unroll
            for (int i = 0; i < 24; i++) y[i] = T(0);
        }
#pragma unroll
        for (int i = 0; i < 24; i++) sy[i][tid] = y[i];
        __syncthreads();
        if (n_in) {
            // node (na, nb) is vertex (1-dx, 1-dy) of element (na + dx, nb + dy)
            T lower[3] = {T(0), T(0), T(0)}, upper[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int dx = d & 1, dy = d >> 1;
                const int e = (na + dx) + EX * (nb + dy);
                const int kl = (1 - dx) + 2 * (1 - dy);     // vertex with kz = 0
                const int ku = kl + 4;                       // vertex with kz = 1
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    lower[c] += sy[3 * kl + c][e];
                    upper[c] += sy[3 * ku + c][e];
                }
            }
            // node plane ez is complete: the element plane below contributed `part`
            if (ez >= z0) {
                const int p = (na + 1) + PX * (nb + 1);
                const uint8_t fl = sf[lo][p];
                if (fl & NF_ACTIVE) {
                    const int s = oct_p(g, gnx, gny, ez);
                    T KU[3];
#pragma unroll
                    for (int c = 0; c < 3; c++) KU[c] = part[c] + lower[c];
                    if (APPLY) {
                        if (fl & NF_FIXED) { KU[0] = KU[1] = KU[2] = T(0); }
                        R.c[0][s] = KU[0]; R.c[1][s] = KU[1]; R.c[2][s] = KU[2];
                        dotacc += (double)KU[0] * (double)su[lo][0][p] + (double)KU[1] * (double)su[lo][1][p] +
                                  (double)KU[2] * (double)su[lo][2][p];
                    } else {
                        if (fl & NF_FIXED) {
                            R.c[0][s] = fo[0]; R.c[1][s] = fo[1]; R.c[2][s] = fo[2];
                        } else {
                            R.c[0][s] = fo[0] - KU[0]; R.c[1][s] = fo[1] - KU[1]; R.c[2][s] = fo[2] - KU[2];
                        }
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 3; c++) part[c] = upper[c];
        }
        lo = hi;
        __syncthreads();
    }
    if (APPLY && partials) {
        const double t = blk_sum_m(dotacc);
        if (tid == 0) partials[bl] = t;
    }
}

template <typename T>
static ModalK<T> modal_values(const double* Mv) {
    ModalK<T> K;
    for (int i = 0; i < RT_MODAL_NNZ; i++) K.m[i] = (T)Mv[i];
    return K;
}

static double g_modal[RT_MODAL_NNZ];

// M = G^-T KE G^-1 on the fixed pattern
void set_modal_constants(const double* ke576) {
    double Gi1[2][2] = {{0.5, -0.5}, {0.5, 0.5}};   // inverse of [[1, 1], [-1, 1]]
    double Gi[8][8];
    for (int m = 0; m < 8; m++)
        for (int k = 0; k < 8; k++)
            Gi[m][k] = Gi1[m & 1][k & 1] * Gi1[(m >> 1) & 1][(k >> 1) & 1] * Gi1[m >> 2][k >> 2];
    // KE G^-1 then G^-T (.)  (component blocks are untouched)
    static double T1[24][24], M[24][24];
    for (int a = 0; a < 24; a++)
        for (int m = 0; m < 8; m++)
            for (int c = 0; c < 3; c++) {
                double s = 0;
                for (int k = 0; k < 8; k++) s += ke576[a * 24 + 3 * k + c] * Gi[k][m];
                T1[a][3 * m + c] = s;
            }
    for (int m = 0; m < 8; m++)
        for (int c = 0; c < 3; c++)
            for (int b = 0; b < 24; b++) {
                double s = 0;
                for (int k = 0; k < 8; k++) s += Gi[k][m] * T1[3 * k + c][b];
                M[3 * m + c][b] = s;
            }
    for (int i = 0; i < RT_MODAL_NNZ; i++) g_modal[i] = M[rt_modal_row[i]][rt_modal_col[i]];
}

static dim3 modal_grid(const Geom& g) {
    if (g.zc < 1) throw std::invalid_argument("z planes per workgroup must be positive");
    return dim3((g.n[0] + TX - 1) / TX, (g.n[1] + TY - 1) / TY, (g.n[2] + g.zc - 1) / g.zc);
}

void modal_tile_dims(int* tx, int* ty) { *tx = TX; *ty = TY; }

int fine_apply_partials_modal(const Geom& g) {
    dim3 gr = modal_grid(g);
    return (int)(gr.x * gr.y * gr.z);
}

template <typename T>
void launch_fine_residual_modal(const Geom& g, Vec3<const T> U, Vec3<const T> F, Vec3<T> R, const float* rhop,
                                const uint8_t* vflag, hipStream_t s) {
    hipLaunchKernelGGL((k_fine_res_modal<T, 0>), modal_grid(g), dim3(256), 0, s, g, U, F, R, rhop, vflag,
                       modal_values<T>(g_modal), (double*)nullptr);
    RT_HIP(hipGetLastError());
}

template <typename T>
void launch_fine_apply_modal(const Geom& g, Vec3<const T> X, Vec3<T> Y, const float* rhop, const uint8_t* vflag,
                             double* partials, int* nparts, hipStream_t s) {
    hipLaunchKernelGGL((k_fine_res_modal<T, 1>), modal_grid(g), dim3(256), 0, s, g, X, X, Y, rhop, vflag,
                       modal_values<T>(g_modal), partials);
    RT_HIP(hipGetLastError());
    if (nparts) *nparts = fine_apply_partials_modal(g);
}

template void launch_fine_residual_modal<float>(const Geom&, Vec3<const float>, Vec3<const float>, Vec3<float>,
                                                const float*, const uint8_t*, hipStream_t);
template void launch_fine_residual_modal<double>(const Geom&, Vec3<const double>, Vec3<const double>, Vec3<double>,
                                                 const float*, const uint8_t*, hipStream_t);
template void launch_fine_apply_modal<float>(const Geom&, Vec3<const float>, Vec3<float>, const float*, const uint8_t*,
                                             double*, int*, hipStream_t);
template void launch_fine_apply_modal<double>(const Geom&, Vec3<const double>, Vec3<double>, const float*,
                                              const uint8_t*, double*, int*, hipStream_t);

}  // namespace rt
