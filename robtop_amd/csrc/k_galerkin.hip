// Galerkin coarse operators built from element matrices.
//
// The reference assembles the x4 coarse stencil A2 = P4^T A~0 P4 node by node
// from rho and KE (restrict_stencil_nondyadic_OTFA_WS_kernel, Grid.cu:392-506)
// and the x2 stencils A_{l+1} = P2^T A_l P2 from stored stencils
// (restrict_stencil_dyadic_kernel, Grid.cu:120-177). Both are sums over
// elements, so the same matrices come out of
//   K_E = sum_{k in E} P_k^T A~_k P_k          (coarse element E, fine children k)
// followed by a stencil assembly over the 8 elements around each coarse node:
//   * x4: K_E = sum_k rho_k^p M_k with M_k = P_k^T KE P_k precomputed (a
//     [576 x 64] x [64 x n_E] product, one rocBLAS DGEMM), plus, for children touching a support,
//     P_k^T (A~_k - rho_k^p KE) P_k where A~_k has the support rows/columns
//     zeroed and 1e6 on their diagonal (Grid.cu:14, :483-488);
//   * x2: K_E' = sum_{c<8} W_c^T K_c W_c (W_c the child's trilinear weights).
// Coarse element matrices are stored entry-major, KEc[(24*row + col) * nE + e].
#include "rt_internal.h"

#include <algorithm>

namespace rt {

// The x4 product K_E = sum_k rho_k^p M_k for every coarse element is one GEMM,
// KEc[576 x nE] = M^T[576 x 64] R[64 x nE], run by rocBLAS (FP64 matrix cores)
// on the hierarchy's stream. This gathers its right operand: R[k][e] = rho^p of
// child k = (kx, ky, kz), k = kx + 4 ky + 16 kz, of coarse element e, in FP64.
__global__ void __launch_bounds__(256) k_children64(Geom gf, int ecx, int ecy, int ecz, const float* __restrict__ rhop,
                                                    double* __restrict__ R) {
    const int nE = ecx * ecy * ecz;
    const long long tot = 64ll * nE;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < tot;
         t += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(t / nE), e = (int)(t % nE);
        const int x = e % ecx, y = (e / ecx) % ecy, z = e / (ecx * ecy);
        R[t] = (double)rhop[oct_p(gf, 4 * x + (k & 3), 4 * y + ((k >> 2) & 3), 4 * z + (k >> 4))];
    }
}

void launch_children64(const Geom& gf, const int ec[3], const float* rhop, double* R, hipStream_t s) {
    const long long tot = 64ll * ec[0] * ec[1] * ec[2];
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_children64, dim3((unsigned)std::min<long long>((tot + 255) / 256, 65536)), dim3(256), 0, s, gf,
                       ec[0], ec[1], ec[2], rhop, R);
    RT_HIP(hipGetLastError());
}

// Support correction: one block per coarse element listed in `ce` (those with
// a child touching a support); one thread per output entry, children in fixed
// order, so the sum is deterministic.
__constant__ double c_KE64_g[576];

__global__ void __launch_bounds__(576) k_galerkin4_fix(Geom gf, int ecx, int ecy, int ecz,
                                                       const float* __restrict__ rhop,
                                                       const uint8_t* __restrict__ vflag,
                                                       const uint8_t* __restrict__ eflag, const int* __restrict__ ce,
                                                       double* __restrict__ KEc) {
    const int e = ce[blockIdx.x];
    const int nE = ecx * ecy * ecz;
    const int x = e % ecx, y = (e / ecx) % ecy, z = e / (ecx * ecy);
    const int t = threadIdx.x;             // output entry (24*row + col)
    const int A = t / 72, ra = (t / 24) % 3, B = (t % 24) / 3, sb = t % 3;   // row = 3A + ra, col = 3B + sb
    const int row = 3 * A + ra, col = 3 * B + sb;
    (void)row; (void)col;
    __shared__ double D[576];
    __shared__ int fixv[8];
    __shared__ float rp;
    __shared__ int touch;
    double acc = 0;
    for (int k = 0; k < 64; k++) {
        const int kx = k & 3, ky = (k >> 2) & 3, kz = k >> 4;
        const int fx = 4 * x + kx, fy = 4 * y + ky, fz = 4 * z + kz;
        __syncthreads();
        if (t == 0) {
            const int es = oct_p(gf, fx, fy, fz);
            touch = (eflag[es] & EF_ACTIVE) && (eflag[es] & EF_TOUCHFIX);
            rp = rhop[es];
        }
        if (t < 8) fixv[t] = (vflag[oct_p(gf, fx + (t & 1), fy + ((t >> 1) & 1), fz + (t >> 2))] & NF_FIXED) ? 1 : 0;
        __syncthreads();
        if (!touch) continue;
        // D = A~_k - rho^p KE (nonzero only in rows/cols of fixed vertices)
        {
            const int i = t / 72, j = (t % 24) / 3;
            const int a = t / 24, b = t % 24;
            double d = 0;
            if (fixv[i] || fixv[j]) {
                d = -(double)rp * c_KE64_g[a * 24 + b];
                if (a == b) d += 1e6;
            }
            D[t] = d;
        }
        __syncthreads();
        // acc += sum_{i,j} W[i][A] D[(i,ra),(j,sb)] W[j][B]
        for (int i = 0; i < 8; i++) {
            const int pix = kx + (i & 1), piy = ky + ((i >> 1) & 1), piz = kz + (i >> 2);
            const int wi = ((A & 1) ? pix : 4 - pix) * ((A & 2) ? piy : 4 - piy) * ((A & 4) ? piz : 4 - piz);
            if (wi == 0) continue;
            double s = 0;
            for (int j = 0; j < 8; j++) {
                const int pjx = kx + (j & 1), pjy = ky + ((j >> 1) & 1), pjz = kz + (j >> 2);
                const int wj = ((B & 1) ? pjx : 4 - pjx) * ((B & 2) ? pjy : 4 - pjy) * ((B & 4) ? pjz : 4 - pjz);
                if (wj == 0) continue;
                s += D[(3 * i + ra) * 24 + 3 * j + sb] * (double)wj;
            }
            acc += s * (double)wi;
        }
    }
    KEc[(size_t)t * nE + e] += acc * (1.0 / 4096.0);
}

void set_ke_galerkin(const double* ke576, hipStream_t s) {
    RT_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_KE64_g), ke576, sizeof(double) * 576, 0, hipMemcpyHostToDevice, s));
}

void launch_galerkin4_fix(const Geom& gf, const int ec[3], const float* rhop, const uint8_t* vflag,
                          const uint8_t* eflag, const int* ce, int nce, double* KEc, hipStream_t s) {
    if (nce <= 0) return;
    hipLaunchKernelGGL(k_galerkin4_fix, dim3(nce), dim3(576), 0, s, gf, ec[0], ec[1], ec[2], rhop, vflag, eflag, ce,
                       KEc);
    RT_HIP(hipGetLastError());
}

// x2 Galerkin on element matrices: one block per parent element, the 8 child
// matrices staged in LDS, one thread per output entry.
__global__ void __launch_bounds__(576) k_galerkin2(const double* __restrict__ KEf, int efx, int efy, int efz,
                                                   double* __restrict__ KEp, int epx, int epy, int epz) {
    const int e = blockIdx.x;
    const int nEp = epx * epy * epz, nEf = efx * efy * efz;
    const int x = e % epx, y = (e / epx) % epy, z = e / (epx * epy);
    const int t = threadIdx.x;
    __shared__ double Kc[8][576];
    for (int c = 0; c < 8; c++) {
        const int cx = 2 * x + (c & 1), cy = 2 * y + ((c >> 1) & 1), cz = 2 * z + (c >> 2);
        double v = 0;
        if (cx < efx && cy < efy && cz < efz) v = KEf[(size_t)t * nEf + cx + efx * (cy + efy * cz)];
        Kc[c][t] = v;
    }
    __syncthreads();
    const int A = t / 72, ra = (t / 24) % 3, B = (t % 24) / 3, sb = t % 3;
    double acc = 0;
    for (int c = 0; c < 8; c++) {
        const int ox = c & 1, oy = (c >> 1) & 1, oz = c >> 2;
        for (int i = 0; i < 8; i++) {
            const int pix = ox + (i & 1), piy = oy + ((i >> 1) & 1), piz = oz + (i >> 2);
            const int wi = ((A & 1) ? pix : 2 - pix) * ((A & 2) ? piy : 2 - piy) * ((A & 4) ? piz : 2 - piz);
            if (wi == 0) continue;
            double s = 0;
            for (int j = 0; j < 8; j++) {
                const int pjx = ox + (j & 1), pjy = oy + ((j >> 1) & 1), pjz = oz + (j >> 2);
                const int wj = ((B & 1) ? pjx : 2 - pjx) * ((B & 2) ? pjy : 2 - pjy) * ((B & 4) ? pjz : 2 - pjz);
                if (wj == 0) continue;
                s += Kc[c][(3 * i + ra) * 24 + 3 * j + sb] * (double)wj;
            }
            acc += s * (double)wi;
        }
    }
    KEp[(size_t)t * nEp + e] = acc * (1.0 / 64.0);
}

void launch_galerkin2(const double* KEf, const int ef[3], double* KEp, const int ep[3], hipStream_t s) {
    int nEp = ep[0] * ep[1] * ep[2];
    hipLaunchKernelGGL(k_galerkin2, dim3(nEp), dim3(576), 0, s, KEf, ef[0], ef[1], ef[2], KEp, ep[0], ep[1], ep[2]);
    RT_HIP(hipGetLastError());
}

// Stencil assembly from element matrices: block (v, v+d) = sum over the
// elements holding both nodes. One thread per (node, neighbour block).
template <typename TS>
__global__ void __launch_bounds__(256) k_assemble(Geom g, const double* __restrict__ KEc, int ex, int ey, int ez,
                                                  const uint8_t* __restrict__ vflag, TS* __restrict__ st) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int nn = nx * ny * nz;
    if (t >= nn * 27) return;
    const int nb = t / nn, vv = t % nn;
    const int x = vv % nx, y = (vv / nx) % ny, z = vv / (nx * ny);
    const int v = oct_p(g, x, y, z);
    if (!(vflag[v] & NF_ACTIVE)) return;
    const int dx = nb % 3 - 1, dy = nb / 3 % 3 - 1, dz = nb / 9 - 1;
    const int nE = ex * ey * ez;
    double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int w = 0; w < 8; w++) {
        const int ox = (w & 1) - 1, oy = ((w >> 1) & 1) - 1, oz = (w >> 2) - 1;   // element offset
        const int lvx = -ox, lvy = -oy, lvz = -oz;                                  // v's local vertex
        const int ljx = lvx + dx, ljy = lvy + dy, ljz = lvz + dz;
        if (ljx < 0 || ljx > 1 || ljy < 0 || ljy > 1 || ljz < 0 || ljz > 1) continue;
        const int exx = x + ox, eyy = y + oy, ezz = z + oz;
        if (exx < 0 || eyy < 0 || ezz < 0 || exx >= ex || eyy >= ey || ezz >= ez) continue;
        const int e = exx + ex * (eyy + ey * ezz);
        const int lv = lvx + 2 * lvy + 4 * lvz, lj = ljx + 2 * ljy + 4 * ljz;
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int q = 0; q < 3; q++) blk[3 * r + q] += KEc[(size_t)((3 * lv + r) * 24 + 3 * lj + q) * nE + e];
    }
    const size_t N = (size_t)g.N;
#pragma unroll
    for (int i = 0; i < 9; i++) st[((size_t)nb * 9 + i) * N + v] = (TS)blk[i];
}

template <typename TS>
void launch_assemble(const Geom& g, const double* KEc, const int e[3], const uint8_t* vflag, TS* st, hipStream_t s) {
    long long tot = (long long)g.n[0] * g.n[1] * g.n[2] * 27;
    hipLaunchKernelGGL((k_assemble<TS>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, KEc, e[0], e[1],
                       e[2], vflag, st);
    RT_HIP(hipGetLastError());
}
template void launch_assemble<double>(const Geom&, const double*, const int*, const uint8_t*, double*, hipStream_t);

}  // namespace rt
