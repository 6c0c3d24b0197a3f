// Kernels of the learned warm start's convolutional network (see
// net_internal.h for the tensor layout). The convolutions are implicit GEMMs on
// the matrix cores: one wave computes 16 consecutive positions (the MFMA's M)
// in each of 1 or 4 consecutive rows, times all output channels (N, in tiles of 16),
// accumulating over the 27 taps and the input channels (K, 16 per instruction).
//
// v_mfma_f32_16x16x16_bf16 operand layout (lane l, i = l % 16, q = l / 16):
//   A: lane l holds A[i][4q .. 4q+3]   B: lane l holds B[4q .. 4q+3][i]
//   D: lane l holds D[4q .. 4q+3][i]
// v_mfma_f32_16x16x32_bf16 (gfx950; layers with 32 or 64 input channels):
//   A: lane l holds A[i][8q .. 8q+7]   B: lane l holds B[8q .. 8q+7][i]
//   D: as above
#include "net_internal.h"

#include <algorithm>
#include <cstdlib>

namespace rt {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct Dims { int n[3]; };
static Dims mk(const int d[3]) { return Dims{{d[0], d[1], d[2]}}; }
static size_t npos(const int d[3]) { return (size_t)d[0] * d[1] * d[2]; }
static unsigned nblocks(size_t n, int per = 256, size_t cap = 1u << 20) {
    return (unsigned)std::max<size_t>(1, std::min<size_t>((n + per - 1) / per, cap));
}

// ---------------------------------------------------------------------------
// convolution (forward, and backward-data with flip = 1 and transposed taps)
// ---------------------------------------------------------------------------
// A wave computes MT row tiles: 16 consecutive x positions in each of MT
// consecutive rows (y) of one plane, so each weight fragment it loads feeds MT
// products. MT = 4 on boxes that still give at least kConvMinWaves waves, else 1
// (small boxes need the waves more than the reuse).
constexpr int kConvMinWaves = 1024;

template <int CIN, int COUT, int EPI, int MT>
__global__ void __launch_bounds__(256) k_conv(Dims D, int flip, const uint16_t* __restrict__ X,
                                              const uint16_t* __restrict__ W, const float* __restrict__ bias,
                                              const uint16_t* __restrict__ aux, void* __restrict__ Yv) {
    const int nx = D.n[0], ny = D.n[1], nz = D.n[2];
    const int lane = threadIdx.x & 63;
    const int segs = (nx + 15) >> 4, ygr = (ny + MT - 1) / MT;
    const int seg = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (seg >= segs * ygr * nz) return;   // whole waves leave; no barrier follows
    const int sx = seg % segs, row = seg / segs, y0 = (row % ygr) * MT, z = row / ygr;
    const int x0 = sx << 4, m = lane & 15, kq = lane >> 4;
    const int xm = x0 + m;
    f4 acc[MT][COUT / 16];
#pragma unroll
    for (int r = 0; r < MT; r++)
#pragma unroll
        for (int nb = 0; nb < COUT / 16; nb++) acc[r][nb] = f4{0.f, 0.f, 0.f, 0.f};
    const int sg = flip ? -1 : 1;
    for (int t = 0; t < 27; t++) {
        const int dx = sg * (t % 3 - 1), dy = sg * ((t / 3) % 3 - 1), dz = sg * (t / 9 - 1);
        const int xx = xm + dx, zz = z + dz;
        const bool okxz = xm < nx && xx >= 0 && xx < nx && zz >= 0 && zz < nz;
        // K per instruction: 32 input channels (8 per lane) when the layer has
        // them, else 16 (4 per lane)
        constexpr int KL = CIN >= 32 ? 8 : 4;
        const uint16_t* xp[MT];
        bool ok[MT];
#pragma unroll
        for (int r = 0; r < MT; r++) {
            const int yy = y0 + r + dy;
            ok[r] = okxz && y0 + r < ny && yy >= 0 && yy < ny;
            const size_t pn = ok[r] ? (size_t)xx + (size_t)nx * ((size_t)yy + (size_t)ny * zz) : 0;
            xp[r] = X + pn * CIN + kq * KL;
        }
        const uint16_t* wp = W + ((size_t)t * COUT + m) * CIN + kq * KL;
        if constexpr (CIN >= 32) {
#pragma unroll
            for (int kb = 0; kb < CIN / 32; kb++) {
                b8 a[MT];
#pragma unroll
                for (int r = 0; r < MT; r++) a[r] = ok[r] ? *(const b8*)(xp[r] + kb * 32) : b8{};
#pragma unroll
                for (int nb = 0; nb < COUT / 16; nb++) {
                    const b8 b = *(const b8*)(wp + (size_t)nb * 16 * CIN + kb * 32);
#pragma unroll
                    for (int r = 0; r < MT; r++)
                        acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b, acc[r][nb], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int kb = 0; kb < CIN / 16; kb++) {
                s4 a[MT];
#pragma unroll
                for (int r = 0; r < MT; r++) a[r] = ok[r] ? *(const s4*)(xp[r] + kb * 16) : s4{0, 0, 0, 0};
#pragma unroll
                for (int nb = 0; nb < COUT / 16; nb++) {
                    const s4 b = *(const s4*)(wp + (size_t)nb * 16 * CIN + kb * 16);
#pragma unroll
                    for (int r = 0; r < MT; r++)
                        acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[r], b, acc[r][nb], 0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < MT; r++) {
        if (y0 + r >= ny) break;
        const size_t prow = (size_t)nx * ((size_t)(y0 + r) + (size_t)ny * z);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int xo = x0 + 4 * kq + j;
            if (xo >= nx) continue;
            const size_t po = prow + xo;
#pragma unroll
            for (int nb = 0; nb < COUT / 16; nb++) {
                const int co = nb * 16 + m;
                float v = acc[r][nb][j];
                if (EPI == EPI_RELU || EPI == EPI_RELU_SKIP || EPI == EPI_F32) v += bias[co];
                if (EPI == EPI_RELU_SKIP) v += bf2f(aux[po * COUT + co]);
                if (EPI == EPI_RELU || EPI == EPI_RELU_SKIP) v = fmaxf(v, 0.f);
                if (EPI == EPI_MASK) v = bf2f(aux[po * COUT + co]) > 0.f ? v : 0.f;
                if (EPI == EPI_F32 || EPI == EPI_RAW_F32)
                    ((float*)Yv)[po * COUT + co] = v;
                else
                    ((uint16_t*)Yv)[po * COUT + co] = f2bf(v);
            }
        }
    }
}

// The same convolution with the input staged in LDS: a workgroup computes a
// 16 x 16 tile of one plane (wave w: rows 4w .. 4w + 3). For each of the three
// input planes the taps read (dz = -1, 0, 1, the taps' outer order) it stages
// the plane's 18 x 18 positions around the tile (zeros outside the box) for KS
// input channels at a time, and the waves read their A fragments from LDS
// instead of gathering them through L1 once per tap. With one slice
// (CIN <= 32) the products run in the order of k_conv<..., 4>, so the results
// are the same bits; with CIN = 64 each plane's nine taps run per 32-channel
// slice.
constexpr int kLdsTile = 16;
constexpr int kLdsHX = kLdsTile + 2, kLdsHP = kLdsHX * kLdsHX;
constexpr int kConvLdsMinTiles = 256;

template <int CIN, int COUT, int EPI>
__global__ void __launch_bounds__(256) k_conv_lds(Dims D, int flip, const uint16_t* __restrict__ X,
                                                  const uint16_t* __restrict__ W, const float* __restrict__ bias,
                                                  const uint16_t* __restrict__ aux, void* __restrict__ Yv) {
    constexpr int KS = CIN >= 32 ? 32 : 16;   // channels per slice: one MFMA K step
    constexpr int KL = CIN >= 32 ? 8 : 4;     // channels per lane and product
    constexpr int CH = KS / 8;                // 16-byte chunks per staged position
    __shared__ __attribute__((aligned(16))) uint16_t tile[kLdsHP * KS];
    const int nx = D.n[0], ny = D.n[1], nz = D.n[2];
    const int tx = (nx + kLdsTile - 1) / kLdsTile, ty = (ny + kLdsTile - 1) / kLdsTile;
    const int bt = blockIdx.x;
    const int x0 = (bt % tx) * kLdsTile, y0 = ((bt / tx) % ty) * kLdsTile, z = bt / (tx * ty);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    f4 acc[4][COUT / 16];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int nb = 0; nb < COUT / 16; nb++) acc[r][nb] = f4{0.f, 0.f, 0.f, 0.f};
    const int sg = flip ? -1 : 1;
    // stage st = (plane pl, slice sl), pl-major; the next stage's global loads
    // are issued into registers before this stage's products and written to
    // LDS after them, so their latency overlaps the matrix-core work
    constexpr int NSL = CIN / KS, NST = 3 * NSL;
    constexpr int NPT = (kLdsHP * CH + 255) / 256;   // staged chunks per thread
    uint4 pre[NPT];
    auto fetch = [&](int st) {
        const int gz = z + sg * (st / NSL - 1), sl = st % NSL;
        const bool zin = gz >= 0 && gz < nz;
#pragma unroll
        for (int j = 0; j < NPT; j++) {
            const int i = threadIdx.x + 256 * j;
            const int pos = i / CH, c = i % CH;
            const int gx = x0 - 1 + pos % kLdsHX, gy = y0 - 1 + pos / kLdsHX;
            pre[j] = uint4{0u, 0u, 0u, 0u};
            if (i < kLdsHP * CH && zin && gx >= 0 && gx < nx && gy >= 0 && gy < ny)
                pre[j] = *(const uint4*)(X + ((size_t)gx + (size_t)nx * ((size_t)gy + (size_t)ny * gz)) * CIN +
                                         sl * KS + c * 8);
        }
    };
    fetch(0);
    for (int st = 0; st < NST; st++) {
        const int pl = st / NSL, sl = st % NSL;
        {
            if (st) __syncthreads();   // every wave has read the previous stage
#pragma unroll
            for (int j = 0; j < NPT; j++) {
                const int i = threadIdx.x + 256 * j;
                if (i < kLdsHP * CH) *(uint4*)(tile + (i / CH) * KS + (i % CH) * 8) = pre[j];
            }
            __syncthreads();
            if (st + 1 < NST) fetch(st + 1);
            for (int t = 9 * pl; t < 9 * pl + 9; t++) {
                const int dx = sg * (t % 3 - 1), dy = sg * ((t / 3) % 3 - 1);
                const int hb = (4 * w + dy + 1) * kLdsHX + m + dx + 1;   // row 0 of the wave
                const uint16_t* wp = W + ((size_t)t * COUT + m) * CIN + sl * KS + kq * KL;
                if constexpr (CIN >= 32) {
                    b8 a[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) a[r] = *(const b8*)(tile + (hb + r * kLdsHX) * KS + kq * KL);
#pragma unroll
                    for (int nb = 0; nb < COUT / 16; nb++) {
                        const b8 b = *(const b8*)(wp + (size_t)nb * 16 * CIN);
#pragma unroll
                        for (int r = 0; r < 4; r++)
                            acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b, acc[r][nb], 0, 0, 0);
                    }
                } else {
                    s4 a[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) a[r] = *(const s4*)(tile + (hb + r * kLdsHX) * KS + kq * KL);
#pragma unroll
                    for (int nb = 0; nb < COUT / 16; nb++) {
                        const s4 b = *(const s4*)(wp + (size_t)nb * 16 * CIN);
#pragma unroll
                        for (int r = 0; r < 4; r++)
                            acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[r], b, acc[r][nb], 0, 0, 0);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int y = y0 + 4 * w + r;
        if (y >= ny) break;
        const size_t prow = (size_t)nx * ((size_t)y + (size_t)ny * z);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int xo = x0 + 4 * kq + j;
            if (xo >= nx) continue;
            const size_t po = prow + xo;
#pragma unroll
            for (int nb = 0; nb < COUT / 16; nb++) {
                const int co = nb * 16 + m;
                float v = acc[r][nb][j];
                if (EPI == EPI_RELU || EPI == EPI_RELU_SKIP || EPI == EPI_F32) v += bias[co];
                if (EPI == EPI_RELU_SKIP) v += bf2f(aux[po * COUT + co]);
                if (EPI == EPI_RELU || EPI == EPI_RELU_SKIP) v = fmaxf(v, 0.f);
                if (EPI == EPI_MASK) v = bf2f(aux[po * COUT + co]) > 0.f ? v : 0.f;
                if (EPI == EPI_F32 || EPI == EPI_RAW_F32)
                    ((float*)Yv)[po * COUT + co] = v;
                else
                    ((uint16_t*)Yv)[po * COUT + co] = f2bf(v);
            }
        }
    }
}

template <int CIN, int COUT>
static void conv_epi_lds(int epi, Dims D, int flip, const uint16_t* X, const uint16_t* W, const float* b,
                         const uint16_t* aux, void* Y, hipStream_t s) {
    const size_t tiles = (size_t)((D.n[0] + kLdsTile - 1) / kLdsTile) * ((D.n[1] + kLdsTile - 1) / kLdsTile) * D.n[2];
    const dim3 grid((unsigned)tiles), blk(256);
    switch (epi) {
        case EPI_RELU: hipLaunchKernelGGL((k_conv_lds<CIN, COUT, EPI_RELU>), grid, blk, 0, s, D, flip, X, W, b, aux, Y); break;
        case EPI_RELU_SKIP: hipLaunchKernelGGL((k_conv_lds<CIN, COUT, EPI_RELU_SKIP>), grid, blk, 0, s, D, flip, X, W, b, aux, Y); break;
        case EPI_F32: hipLaunchKernelGGL((k_conv_lds<CIN, COUT, EPI_F32>), grid, blk, 0, s, D, flip, X, W, b, aux, Y); break;
        case EPI_MASK: hipLaunchKernelGGL((k_conv_lds<CIN, COUT, EPI_MASK>), grid, blk, 0, s, D, flip, X, W, b, aux, Y); break;
        case EPI_RAW_F32: hipLaunchKernelGGL((k_conv_lds<CIN, COUT, EPI_RAW_F32>), grid, blk, 0, s, D, flip, X, W, b, aux, Y); break;
        default: throw std::invalid_argument("conv: unknown epilogue");
    }
}

// the LDS-staged convolution on the boxes large enough for it (on by default;
// RT_CONV_LDS=0 in the environment keeps k_conv, for comparison)
static bool conv_lds_on() {
    static const bool on = [] {
        const char* e = std::getenv("RT_CONV_LDS");
        return !(e && e[0] == '0');
    }();
    return on;
}

template <int CIN, int COUT, int MT>
static void conv_epi_mt(int epi, Dims D, int flip, const uint16_t* X, const uint16_t* W, const float* b,
                        const uint16_t* aux, void* Y, hipStream_t s) {
    const size_t waves = (size_t)((D.n[0] + 15) / 16) * ((D.n[1] + MT - 1) / MT) * D.n[2];
    const dim3 grid((unsigned)((waves + 3) / 4)), blk(256);
    switch (epi) {
        case EPI_RELU: hipLaunchKernelGGL((k_conv<CIN, COUT, EPI_RELU, MT>), grid, blk, 0, s, D, flip, X, W, b, aux, Y); break;
        case EPI_RELU_SKIP: hipLaunchKernelGGL((k_conv<CIN, COUT, EPI_RELU_SKIP, MT>), grid, blk, 0, s, D, flip, X, W, b, aux, Y); break;
        case EPI_F32: hipLaunchKernelGGL((k_conv<CIN, COUT, EPI_F32, MT>), grid, blk, 0, s, D, flip, X, W, b, aux, Y); break;
        case EPI_MASK: hipLaunchKernelGGL((k_conv<CIN, COUT, EPI_MASK, MT>), grid, blk, 0, s, D, flip, X, W, b, aux, Y); break;
        case EPI_RAW_F32: hipLaunchKernelGGL((k_conv<CIN, COUT, EPI_RAW_F32, MT>), grid, blk, 0, s, D, flip, X, W, b, aux, Y); break;
        default: throw std::invalid_argument("conv: unknown epilogue");
    }
}
template <int CIN, int COUT>
static void conv_epi(int epi, Dims D, int flip, const uint16_t* X, const uint16_t* W, const float* b,
                     const uint16_t* aux, void* Y, hipStream_t s) {
    const size_t waves4 = (size_t)((D.n[0] + 15) / 16) * ((D.n[1] + 3) / 4) * D.n[2];
    const size_t tiles = (size_t)((D.n[0] + kLdsTile - 1) / kLdsTile) * ((D.n[1] + kLdsTile - 1) / kLdsTile) * D.n[2];
    if (conv_lds_on() && tiles >= (size_t)kConvLdsMinTiles) conv_epi_lds<CIN, COUT>(epi, D, flip, X, W, b, aux, Y, s);
    else if (waves4 >= (size_t)kConvMinWaves) conv_epi_mt<CIN, COUT, 4>(epi, D, flip, X, W, b, aux, Y, s);
    else conv_epi_mt<CIN, COUT, 1>(epi, D, flip, X, W, b, aux, Y, s);
}

#define RT_CONV_SHAPES(X_) X_(16, 16) X_(16, 32) X_(32, 16) X_(32, 32) X_(32, 64) X_(64, 32) X_(64, 64)

void launch_conv(int cin, int cout, int epi, const int dims[3], int flip, const uint16_t* X, const uint16_t* W,
                 const float* bias, const uint16_t* aux, void* Y, hipStream_t s) {
    const Dims D = mk(dims);
#define RT_C(a, b) if (cin == a && cout == b) { conv_epi<a, b>(epi, D, flip, X, W, bias, aux, Y, s); RT_HIP(hipGetLastError()); return; }
    RT_CONV_SHAPES(RT_C)
#undef RT_C
    throw std::invalid_argument("conv: unsupported channel counts");
}

// ---------------------------------------------------------------------------
// weight gradient: dW[t][co][ci] = sum_p G[p][co] X[p + d_t][ci]. MFMA with
// M = output channel, N = input channel, K = 16 positions of a row.
// ---------------------------------------------------------------------------
constexpr int kWgradBlocks = 32;

template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv_wgrad(Dims D, const uint16_t* __restrict__ X,
                                                    const uint16_t* __restrict__ G, float* __restrict__ part) {
    const int nx = D.n[0], ny = D.n[1], nz = D.n[2];
    const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
    const int t = blockIdx.y;
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    const int segs = (nx + 15) >> 4, total = segs * ny * nz;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6), nw = gridDim.x * 4;
    const long long shift = dx + (long long)nx * (dy + (long long)ny * dz);
    f4 acc[COUT / 16][CIN / 16];
#pragma unroll
    for (int nb = 0; nb < COUT / 16; nb++)
#pragma unroll
        for (int kb = 0; kb < CIN / 16; kb++) acc[nb][kb] = f4{0.f, 0.f, 0.f, 0.f};
    for (int seg = w; seg < total; seg += nw) {
        const int sx = seg % segs, row = seg / segs, y = row % ny, z = row / ny;
        const bool rowok = y + dy >= 0 && y + dy < ny && z + dz >= 0 && z + dz < nz;
        const size_t prow = (size_t)nx * ((size_t)y + (size_t)ny * z);
        s4 a[COUT / 16], b[CIN / 16];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int xk = (sx << 4) + 4 * kq + j;
            const bool in = xk < nx;
            const bool nok = in && rowok && xk + dx >= 0 && xk + dx < nx;
            const size_t pk = prow + (in ? xk : 0);
            const size_t pn = nok ? (size_t)((long long)pk + shift) : 0;
#pragma unroll
            for (int nb = 0; nb < COUT / 16; nb++) a[nb][j] = in ? (short)G[pk * COUT + nb * 16 + m] : (short)0;
#pragma unroll
            for (int kb = 0; kb < CIN / 16; kb++) b[kb][j] = nok ? (short)X[pn * CIN + kb * 16 + m] : (short)0;
        }
#pragma unroll
        for (int nb = 0; nb < COUT / 16; nb++)
#pragma unroll
            for (int kb = 0; kb < CIN / 16; kb++)
                acc[nb][kb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[nb], b[kb], acc[nb][kb], 0, 0, 0);
    }
    // lane holds D[4 kq + j][m]: co = nb*16 + 4 kq + j, ci = kb*16 + m
    float* out = part + ((size_t)w * 27 + t) * COUT * CIN;
#pragma unroll
    for (int nb = 0; nb < COUT / 16; nb++)
#pragma unroll
        for (int kb = 0; kb < CIN / 16; kb++)
#pragma unroll
            for (int j = 0; j < 4; j++) out[(size_t)(nb * 16 + 4 * kq + j) * CIN + kb * 16 + m] = acc[nb][kb][j];
}

size_t conv_wgrad_part_elems(int cin, int cout) { return (size_t)kWgradBlocks * 4 * 27 * cin * cout; }

int launch_conv_wgrad(int cin, int cout, const int dims[3], const uint16_t* X, const uint16_t* G, float* part,
                      hipStream_t s) {
    const Dims D = mk(dims);
    const long long total = (long long)((dims[0] + 15) / 16) * dims[1] * dims[2];
    const int nblk = (int)std::max<long long>(1, std::min<long long>(kWgradBlocks, total / 64));
    const dim3 grid((unsigned)nblk, 27), blk(256);
#define RT_C(a_, b_) if (cin == a_ && cout == b_) { hipLaunchKernelGGL((k_conv_wgrad<a_, b_>), grid, blk, 0, s, D, X, G, part); RT_HIP(hipGetLastError()); return nblk * 4; }
    RT_CONV_SHAPES(RT_C)
#undef RT_C
    throw std::invalid_argument("conv wgrad: unsupported channel counts");
}

__global__ void k_sum_parts(const float* __restrict__ part, int nparts, size_t n, float* __restrict__ out, int acc) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float v = acc ? out[i] : 0.f;
        for (int w = 0; w < nparts; w++) v += part[(size_t)w * n + i];
        out[i] = v;
    }
}
void launch_sum_parts(const float* part, int nparts, size_t n, float* out, int accumulate, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_parts, dim3(nblocks(n, 256, 4096)), dim3(256), 0, s, part, nparts, n, out, accumulate);
    RT_HIP(hipGetLastError());
}

// bias gradient: block b sums its positions per channel into part[b][c]
constexpr int kBiasBlocks = 64;
__global__ void __launch_bounds__(256) k_bias_grad(int C, size_t P, const uint16_t* __restrict__ G,
                                                   float* __restrict__ part) {
    __shared__ float red[256];
    const int c = threadIdx.x % C, r = threadIdx.x / C, rows = 256 / C;
    float v = 0.f;
    for (size_t p = (size_t)blockIdx.x * rows + r; p < P; p += (size_t)gridDim.x * rows) v += bf2f(G[p * C + c]);
    red[threadIdx.x] = v;
    __syncthreads();
    if (r == 0) {
        for (int k = 1; k < rows; k++) v += red[k * C + c];
        part[(size_t)blockIdx.x * C + c] = v;
    }
}
void launch_bias_grad(int c, size_t P, const uint16_t* G, float* part, float* db, hipStream_t s) {
    hipLaunchKernelGGL(k_bias_grad, dim3(kBiasBlocks), dim3(256), 0, s, c, P, G, part);
    RT_HIP(hipGetLastError());
    launch_sum_parts(part, kBiasBlocks, (size_t)c, db, 1, s);
}

// ---------------------------------------------------------------------------
// pooling and upsampling between a box n and its half (n + 1) / 2
// ---------------------------------------------------------------------------
__device__ __forceinline__ void decode(size_t p, const Dims& D, int& x, int& y, int& z) {
    x = (int)(p % D.n[0]);
    const size_t r = p / D.n[0];
    y = (int)(r % D.n[1]);
    z = (int)(r / D.n[1]);
}
__device__ __forceinline__ Dims half(const Dims& D) {
    return Dims{{(D.n[0] + 1) / 2, (D.n[1] + 1) / 2, (D.n[2] + 1) / 2}};
}
__device__ __forceinline__ size_t lin(const Dims& D, int x, int y, int z) {
    return (size_t)x + (size_t)D.n[0] * ((size_t)y + (size_t)D.n[1] * z);
}
__device__ __forceinline__ int children(const Dims& F, int qx, int qy, int qz) {
    return (2 * qx + 1 < F.n[0] ? 2 : 1) * (2 * qy + 1 < F.n[1] ? 2 : 1) * (2 * qz + 1 < F.n[2] ? 2 : 1);
}

// 2x2x2 average pool, eight channels (16 bytes) per thread: C is a multiple of 16
__global__ void k_pool(int C, Dims F, const uint16_t* __restrict__ X, uint16_t* __restrict__ Y) {
    const Dims Hd = half(F);
    const int c8 = C >> 3;
    const size_t n = (size_t)Hd.n[0] * Hd.n[1] * Hd.n[2] * c8;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % c8) * 8;
        const size_t q = i / c8;
        int qx, qy, qz;
        decode(q, Hd, qx, qy, qz);
        float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < 8; k++) {
            const int x = 2 * qx + (k & 1), y = 2 * qy + ((k >> 1) & 1), z = 2 * qz + (k >> 2);
            if (x < F.n[0] && y < F.n[1] && z < F.n[2]) {
                const uint4 w = *(const uint4*)(X + lin(F, x, y, z) * C + c);
                const uint32_t u[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    v[2 * j] += bf2f((uint16_t)(u[j] & 0xffffu));
                    v[2 * j + 1] += bf2f((uint16_t)(u[j] >> 16));
                }
            }
        }
        const float nc = (float)children(F, qx, qy, qz);
        uint32_t o[4];
#pragma unroll
        for (int j = 0; j < 4; j++) o[j] = (uint32_t)f2bf(v[2 * j] / nc) | ((uint32_t)f2bf(v[2 * j + 1] / nc) << 16);
        *(uint4*)(Y + q * C + c) = uint4{o[0], o[1], o[2], o[3]};
    }
}
// nearest upsampling, eight channels (16 bytes) per thread: C is a multiple of 16
__global__ void k_upsample(int C, Dims F, const uint16_t* __restrict__ X, uint16_t* __restrict__ Y) {
    const Dims Hd = half(F);
    const int c8 = C >> 3;
    const size_t n = (size_t)F.n[0] * F.n[1] * F.n[2] * c8;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % c8) * 8;
        const size_t p = i / c8;
        int x, y, z;
        decode(p, F, x, y, z);
        *(uint4*)(Y + p * C + c) = *(const uint4*)(X + lin(Hd, x >> 1, y >> 1, z >> 1) * C + c);
    }
}
__global__ void k_upsample_t(int C, Dims F, const float* __restrict__ Hf, const uint16_t* __restrict__ act,
                             uint16_t* __restrict__ G) {
    const Dims Hd = half(F);
    const size_t n = (size_t)Hd.n[0] * Hd.n[1] * Hd.n[2] * C;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        int qx, qy, qz;
        decode(i / C, Hd, qx, qy, qz);
        float v = 0.f;
        for (int k = 0; k < 8; k++) {
            const int x = 2 * qx + (k & 1), y = 2 * qy + ((k >> 1) & 1), z = 2 * qz + (k >> 2);
            if (x < F.n[0] && y < F.n[1] && z < F.n[2]) v += Hf[lin(F, x, y, z) * C + c];
        }
        G[i] = f2bf(bf2f(act[i]) > 0.f ? v : 0.f);
    }
}
__global__ void k_pool_t(int C, Dims F, const float* __restrict__ Hf, const uint16_t* __restrict__ skip,
                         const uint16_t* __restrict__ act, uint16_t* __restrict__ G) {
    const Dims Hd = half(F);
    const size_t n = (size_t)F.n[0] * F.n[1] * F.n[2] * C;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        int x, y, z;
        decode(i / C, F, x, y, z);
        const int qx = x >> 1, qy = y >> 1, qz = z >> 1;
        float v = Hf[lin(Hd, qx, qy, qz) * C + c] / (float)children(F, qx, qy, qz);
        if (skip) v += bf2f(skip[i]);
        G[i] = f2bf(bf2f(act[i]) > 0.f ? v : 0.f);
    }
}

static size_t half_pos(const int d[3]) { return (size_t)((d[0] + 1) / 2) * ((d[1] + 1) / 2) * ((d[2] + 1) / 2); }

void launch_pool(int c, const int dims[3], const uint16_t* X, uint16_t* Y, hipStream_t s) {
    if (c % 16) throw std::invalid_argument("pool: channels must be a multiple of 16");
    hipLaunchKernelGGL(k_pool, dim3(nblocks(half_pos(dims) * (c / 8))), dim3(256), 0, s, c, mk(dims), X, Y);
    RT_HIP(hipGetLastError());
}
void launch_upsample(int c, const int dims[3], const uint16_t* X, uint16_t* Y, hipStream_t s) {
    if (c % 16) throw std::invalid_argument("upsample: channels must be a multiple of 16");
    hipLaunchKernelGGL(k_upsample, dim3(nblocks(npos(dims) * (c / 8))), dim3(256), 0, s, c, mk(dims), X, Y);
    RT_HIP(hipGetLastError());
}
void launch_upsample_t(int c, const int dims[3], const float* H, const uint16_t* act, uint16_t* G, hipStream_t s) {
    hipLaunchKernelGGL(k_upsample_t, dim3(nblocks(half_pos(dims) * c)), dim3(256), 0, s, c, mk(dims), H, act, G);
    RT_HIP(hipGetLastError());
}
void launch_pool_t(int c, const int dims[3], const float* H, const uint16_t* skip, const uint16_t* act, uint16_t* G,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_pool_t, dim3(nblocks(npos(dims) * c)), dim3(256), 0, s, c, mk(dims), H, skip, act, G);
    RT_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// loss, optimiser, weight packing, host tensor conversion
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_loss(size_t P, const float* __restrict__ Y, const float* __restrict__ T,
                                              float tscale, float gscale, uint16_t* __restrict__ G,
                                              double* __restrict__ partials) {
    __shared__ double red[256];
    double acc = 0;
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 16; c++) {
            float g = 0.f;
            if (c < 3) {
                const float d = Y[p * 16 + c] - tscale * T[p * 3 + c];
                acc += (double)d * d;
                g = 2.f * d * gscale;
            }
            G[p * 16 + c] = f2bf(g);
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}
void launch_loss(size_t P, const float* Y, const float* T, float tscale, float gscale, uint16_t* G, double* partials,
                 int nparts, hipStream_t s) {
    hipLaunchKernelGGL(k_loss, dim3(nparts), dim3(256), 0, s, P, Y, T, tscale, gscale, G, partials);
    RT_HIP(hipGetLastError());
}

__global__ void k_adam(size_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
                       float gscale) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i] * gscale;
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}
void launch_adam(size_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                 float c1, float c2, float gscale, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3(nblocks(n, 256, 4096)), dim3(256), 0, s, n, w, g, m, v, lr, b1, b2, eps, c1, c2,
                       gscale);
    RT_HIP(hipGetLastError());
}

__global__ void k_pack_weights(int cin, int cout, const float* __restrict__ W, uint16_t* __restrict__ Wf,
                               uint16_t* __restrict__ Wb) {
    const size_t n = (size_t)27 * cin * cout;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int ci = (int)(i % cin);
        const size_t r = i / cin;
        const int co = (int)(r % cout), t = (int)(r / cout);
        const uint16_t h = f2bf(W[i]);
        Wf[i] = h;
        Wb[((size_t)t * cin + ci) * cout + co] = h;
    }
}
void launch_pack_weights(int cin, int cout, const float* W, uint16_t* Wf, uint16_t* Wb, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_weights, dim3(nblocks((size_t)27 * cin * cout)), dim3(256), 0, s, cin, cout, W, Wf, Wb);
    RT_HIP(hipGetLastError());
}

__global__ void k_pack_input(size_t P, const float* __restrict__ x4, uint16_t* __restrict__ X) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < P * 16; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i & 15);
        X[i] = c < 4 ? f2bf(x4[(i >> 4) * 4 + c]) : (uint16_t)0;
    }
}
void launch_pack_input(size_t P, const float* x4, uint16_t* X, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_input, dim3(nblocks(P * 16)), dim3(256), 0, s, P, x4, X);
    RT_HIP(hipGetLastError());
}
__global__ void k_unpack_output(size_t P, const float* __restrict__ Y, float* __restrict__ y3) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < P * 3; i += (size_t)gridDim.x * blockDim.x)
        y3[i] = Y[(i / 3) * 16 + i % 3];
}
void launch_unpack_output(size_t P, const float* Y, float* y3, hipStream_t s) {
    hipLaunchKernelGGL(k_unpack_output, dim3(nblocks(P * 3)), dim3(256), 0, s, P, Y, y3);
    RT_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// hierarchy <-> network tensors on the x4 level's node box
// ---------------------------------------------------------------------------
__global__ void k_net_features(Geom g0, int ex, int ey, int ez, Geom gc, const float* __restrict__ rho,
                               Vec3<const double> Fc, float fscale, uint16_t* __restrict__ X) {
    const size_t P = (size_t)gc.n[0] * gc.n[1] * gc.n[2];
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(p % gc.n[0]);
        const size_t r = p / gc.n[0];
        const int j = (int)(r % gc.n[1]), k = (int)(r / gc.n[1]);
        float d = 0.f;
        for (int z = std::max(4 * k - 4, 0); z < std::min(4 * k + 4, ez); z++)
            for (int y = std::max(4 * j - 4, 0); y < std::min(4 * j + 4, ey); y++)
                for (int x = std::max(4 * i - 4, 0); x < std::min(4 * i + 4, ex); x++) d += rho[oct_p(g0, x, y, z)];
        const int sc = oct_p(gc, i, j, k);
        uint16_t* o = X + p * 16;
        o[0] = f2bf(d * (1.f / 512.f));
        for (int c = 0; c < 3; c++) o[1 + c] = f2bf((float)Fc.c[c][sc] * fscale);
        for (int c = 4; c < 16; c++) o[c] = 0;
    }
}
void launch_net_features(const Geom& g0, const int e0[3], const float* rho, const Geom& gc, Vec3<const double> Fc,
                         float fscale, uint16_t* X, hipStream_t s) {
    const size_t P = (size_t)gc.n[0] * gc.n[1] * gc.n[2];
    hipLaunchKernelGGL(k_net_features, dim3(nblocks(P)), dim3(256), 0, s, g0, e0[0], e0[1], e0[2], gc, rho, Fc,
                       fscale, X);
    RT_HIP(hipGetLastError());
}

__global__ void __launch_bounds__(256) k_net_target(Geom g0, Geom gc, Vec3<const double> U, float* __restrict__ T,
                                                    double* __restrict__ partials) {
    __shared__ double red[256];
    const size_t P = (size_t)gc.n[0] * gc.n[1] * gc.n[2];
    double acc = 0;
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(p % gc.n[0]);
        const size_t r = p / gc.n[0];
        const int j = (int)(r % gc.n[1]), k = (int)(r / gc.n[1]);
        const int sf = oct_p(g0, 4 * i, 4 * j, 4 * k);
        for (int c = 0; c < 3; c++) {
            const double v = U.c[c][sf];
            T[p * 3 + c] = (float)v;
            acc += v * v;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}
void launch_net_target(const Geom& g0, const Geom& gc, Vec3<const double> U, float* T, double* partials, int nparts,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_net_target, dim3(nparts), dim3(256), 0, s, g0, gc, U, T, partials);
    RT_HIP(hipGetLastError());
}

__global__ void k_net_to_level(Geom gc, const float* __restrict__ Y, const uint8_t* __restrict__ vflag,
                               Vec3<double> Uc) {
    const size_t P = (size_t)gc.n[0] * gc.n[1] * gc.n[2];
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(p % gc.n[0]);
        const size_t r = p / gc.n[0];
        const int j = (int)(r % gc.n[1]), k = (int)(r / gc.n[1]);
        const int sc = oct_p(gc, i, j, k);
        const bool act = (vflag[sc] & NF_ACTIVE) != 0;
        for (int c = 0; c < 3; c++) Uc.c[c][sc] = act ? (double)Y[p * 16 + c] : 0.0;
    }
}
void launch_net_to_level(const Geom& gc, const float* Y, const uint8_t* vflag_c, Vec3<double> Uc, hipStream_t s) {
    const size_t P = (size_t)gc.n[0] * gc.n[1] * gc.n[2];
    hipLaunchKernelGGL(k_net_to_level, dim3(nblocks(P)), dim3(256), 0, s, gc, Y, vflag_c, Uc);
    RT_HIP(hipGetLastError());
}

__global__ void __launch_bounds__(256) k_absmax3(int N, Vec3<const double> F, unsigned* __restrict__ out) {
    __shared__ float red[256];
    float v = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        for (int c = 0; c < 3; c++) v = fmaxf(v, fabsf((float)F.c[c][i]));
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, __float_as_uint(red[0]));
}
void launch_absmax3(const Geom& gc, Vec3<const double> Fc, unsigned* out, hipStream_t s) {
    hipLaunchKernelGGL(k_absmax3, dim3(nblocks((size_t)gc.N, 256, 1024)), dim3(256), 0, s, gc.N, Fc, out);
    RT_HIP(hipGetLastError());
}

}  // namespace rt
