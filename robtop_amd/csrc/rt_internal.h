// Internal declarations shared by the HIP kernels and the host library.
//
// Layout ("octant" layout). Every level stores nodes, and the finest level
// stores elements, on a box padded by two ghost layers below and at least one
// above, split into 8 colour blocks by coordinate parity:
//   q = p + 2 (padded coords), colour(q) = (qx&1) | (qy&1)<<1 | (qz&1)<<2,
//   index(q) = colour*cs + ((qz>>1)*hy + (qy>>1))*hx + (qx>>1).
// The box origin is a multiple of the coarsest spacing, so parity in the box is
// the absolute parity the reference colours by (Grid.cu:579). A colour pass of
// the Gauss-Seidel smoother therefore reads and writes contiguous runs, and each
// of its 27 neighbour streams is a contiguous run of another colour block.
// Ghost slots and absent nodes/elements hold zeros and are never written.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>
#include <stdexcept>

namespace rt {

enum NodeFlag : uint8_t {
    NF_ACTIVE = 1,      // node exists (some incident element is solid)
    NF_FIXED = 2,       // support node (all three components fixed)
    NF_NEARFIX = 4,     // some other node of the 27-neighbourhood is fixed
    NF_LOAD = 8,        // load node (information only)
    // bits 4..7: number of solid incident elements (1..8)
};
enum ElemFlag : uint8_t {
    EF_ACTIVE = 1,
    EF_SHELL = 2,       // density forced to 1 by the OC update (optimization.cu:159)
    EF_TOUCHFIX = 4,    // some vertex is fixed
};

struct Geom {
    int n[3];           // interior node box
    int h[3];           // padded half dims, h = (n + 5) / 2
    int cs;             // colour stride h0*h1*h2
    int N;              // 8*cs slots
    // launch shapes (chosen per level by the autotuner, rt_autotune):
    int zc = 24;        // finest level: z planes marched per residual / apply workgroup
    int coarse_gs = 1;  // coarse smoother: 1 = nine waves per 64 nodes (sliced), 0 = a thread per node
    int coarse_res = -1;// coarse residual: 1 sliced, 0 a thread per node, -1 = by level size
    // finest level: device mask of the residual / apply tiles (of the current
    // zc) that hold an active node; null = all
    const uint8_t* tile_act = nullptr;
};

inline Geom make_geom(int nx, int ny, int nz) {
    Geom g;
    g.n[0] = nx; g.n[1] = ny; g.n[2] = nz;
    for (int k = 0; k < 3; k++) g.h[k] = (g.n[k] + 5) / 2;
    long long cs = (long long)g.h[0] * g.h[1] * g.h[2];
    if (cs * 8 >= (1ll << 31)) throw std::runtime_error("level too large for 32-bit slot indices");
    g.cs = (int)cs;
    g.N = 8 * g.cs;
    return g;
}

// padded coords -> slot
__host__ __device__ inline int oct_index(const Geom& g, int qx, int qy, int qz) {
    int c = (qx & 1) | ((qy & 1) << 1) | ((qz & 1) << 2);
    return c * g.cs + ((qz >> 1) * g.h[1] + (qy >> 1)) * g.h[0] + (qx >> 1);
}
// interior node/element coords -> slot
__host__ __device__ inline int oct_p(const Geom& g, int x, int y, int z) { return oct_index(g, x + 2, y + 2, z + 2); }

template <typename T> struct Vec3 { T* c[3]; };

// Workgroups are handed to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8), and each XCD has its own L2. This maps workgroup b to a work index
// so that XCD x processes one contiguous range of work items: neighbouring
// items (rows that share halo data) then meet in the same L2. A bijection
// on [0, n).
__device__ __forceinline__ int xcd_row(int b, int n) {
    const int x = b & 7, loc = b >> 3, q = n >> 3, r = n & 7;
    return x * q + (x < r ? x : r) + loc;
}

// Raw buffer loads: voff is the per-lane byte offset, soff a wave-uniform byte
// offset held in a scalar register, so a kernel's many neighbour loads need no
// per-load address arithmetic. Both offsets must be non-negative.
template <typename T> __device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
// bfloat16 storage (upper half of an IEEE binary32); widened on load
struct bf16 {
    uint16_t v;
    __host__ __device__ operator float() const { return __builtin_bit_cast(float, (uint32_t)v << 16); }
};
template <>
__device__ __forceinline__ bf16 bload<bf16>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return bf16{(uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0)};
}
__device__ __forceinline__ uint8_t bload_u8(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_amdgcn_raw_buffer_load_b8(r, voff, soff, 0);
}
// buffer resource over `bytes` bytes from p (loads beyond them return 0)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, bytes, 0x00020000);
}

// ---------------------------------------------------------------------------
// kernel launchers (defined in the .hip files); all launch on stream `s`
// ---------------------------------------------------------------------------
// k_fine.hip
void set_ke_constants(const double* ke576, hipStream_t s);
// zero_start: the first forward sweep of a cycle from u = 0 with zero supports;
// U's old contents are ignored, and every node slot of the pair is written.
// fu_part (a reverse sweep with zero supports): each workgroup also writes the
// sum of f . u_new over the nodes it relaxes, fine_gs_pair_blocks(g, pair) of them.
template <typename T>
void launch_fine_gs(const Geom& g, Vec3<T> U, Vec3<const T> F, const float* rhop, const uint8_t* vflag, int pair,
                    bool reverse, bool zero_support, hipStream_t s, bool zero_start = false,
                    double* fu_part = nullptr);
int fine_gs_pair_blocks(const Geom& g, int pair);
// one full fine sweep in two launches, old values from Uin, new ones to Uout
// (Uin is not written; the two must not overlap)
template <typename T>
void launch_fine_gs_sweep2(const Geom& g, Vec3<const T> Uin, Vec3<T> Uout, Vec3<const T> F, const float* rhop,
                           const uint8_t* vflag, bool reverse, bool zero_support, hipStream_t s);

// k_fine_modal.hip (element-centric residual / apply in the sum-difference basis)
void set_modal_constants(const double* ke576);
template <typename T>
void launch_fine_residual_modal(const Geom& g, Vec3<const T> U, Vec3<const T> F, Vec3<T> R, const float* rhop,
                                const uint8_t* vflag, hipStream_t s);
template <typename T>
void launch_fine_apply_modal(const Geom& g, Vec3<const T> X, Vec3<T> Y, const float* rhop, const uint8_t* vflag,
                             double* partials, int* nparts, hipStream_t s);
int fine_apply_partials_modal(const Geom& g);
// node tile of the residual / apply kernels (x by y; z by Geom::zc)
void modal_tile_dims(int* tx, int* ty);

// k_coarse.hip
template <typename T, typename TS>
void launch_coarse_gs_pair(const Geom& g, Vec3<T> U, Vec3<const T> F, const TS* st, const uint8_t* vflag, int pair,
                           bool reverse, hipStream_t s);
// whether launch_coarse_gs_pair applies to this level (rows of <= 64 nodes of a
// colour, sliced smoother selected)
bool coarse_gs_pairs(const Geom& g);
template <typename T, typename TS>
void launch_coarse_gs(const Geom& g, Vec3<T> U, Vec3<const T> F, const TS* st, const uint8_t* vflag, int colour,
                      bool reverse, hipStream_t s);
template <typename T, typename TS>
void launch_coarse_residual(const Geom& g, Vec3<const T> U, Vec3<const T> F, Vec3<T> R, const TS* st,
                            const uint8_t* vflag, hipStream_t s);
// batched coarse-level kernels: nb <= kBatchMax vectors, vector b at base + b 3N
constexpr int kBatchMax = 8;
template <typename T, typename TS>
void launch_coarse_gs_batch(const Geom& g, int nb, T* U, const T* F, const TS* st, const uint8_t* vflag, int colour,
                            bool reverse, hipStream_t s);
template <typename T, typename TS>
void launch_coarse_residual_batch(const Geom& g, int nb, const T* U, const T* F, T* R, const TS* st,
                                  const uint8_t* vflag, hipStream_t s);
// fbuf must hold nb * n3 doubles
template <typename T>
void launch_coarsest_batch(const double* Ainv, int n3, const int* rowslot, int nb, int N, const T* F, T* U,
                           double* fbuf, hipStream_t s);
template <typename T>
void launch_restrict(const Geom& gf, const int of[3], Vec3<const T> Rf, const Geom& gc, const int oc[3], Vec3<T> Fc,
                     const uint8_t* vflag_c, int ratio, T* scratch, hipStream_t s);
// elements of T the two-stage x4 restriction needs as scratch
size_t restrict_scratch_elems(const Geom& gf, const Geom& gc);
template <typename T>
void launch_prolong(const Geom& gf, const int of[3], Vec3<T> Uf, const uint8_t* vflag_f, const Geom& gc,
                    const int oc[3], Vec3<const T> Uc, int ratio, bool skip_fixed, hipStream_t s);
template <typename T>
void launch_coarsest(const double* Ainv, int n3, const int* rowslot, int unused, Vec3<const T> F, Vec3<T> U,
                     double* fbuf, hipStream_t s);
void launch_mirror_upper(double* A, int n, hipStream_t s);
void launch_dense_from_stencil(const double* st, size_t N, const int* slots, const int* nbcol, int nr, double* A,
                               hipStream_t s);
// x = sum_i c[i] h[i], k <= 8 FP64 vectors of length n
void launch_lincomb(double* x, const double* const* h, const double* c, int k, size_t n, hipStream_t s);
// floating-cluster deflation (k_island.hip): builds the islands of stiff
// elements (rho >= theta, at least min_size elements, at most max_isl) that
// touch no support, their centroids and pseudo-inverted rigid-body energies;
// returns the island count
int island_build(const Geom& g, const float* rho, const float* rhop, const uint8_t* eflag, const uint8_t* vflag,
                 const double* KE_dev, float theta, int min_size, int max_isl, int* lab, uint8_t* anchored, int* count,
                 int* root_isl, int* node_isl, double* cent, double* E, double* Einv, int* flag, hipStream_t s);
// Z += sum_i Z_i E_i^+ Z_i^T R (zr must be zero on entry and is left zero)
template <typename T>
void island_correct(const Geom& g, int nisl, const int* node_isl, const double* cent, const double* Einv, double* zr,
                    double* coef, Vec3<const T> R, Vec3<T> Zv, hipStream_t s);
// b = a rounded to bfloat16 (round to nearest even, through binary32)
void launch_to_bf16(const double* a, bf16* b, size_t n, hipStream_t s);
template <typename TA, typename TB>
void launch_convert(const TA* a, TB* b, size_t n, hipStream_t s);
template <typename T>
void launch_fill(T* a, T v, size_t n, hipStream_t s);
void launch_dot_partials(const double* a, const double* b, size_t n, double* partials, int nparts, hipStream_t s);
template <typename TA, typename TB>
void launch_dot_partials_t(const TA* a, const TB* b, size_t n, double* partials, int nparts, hipStream_t s,
                           const uint8_t* chunk = nullptr);
void launch_sum_partials(const double* partials, int nparts, double* out, hipStream_t s);
template <typename T, typename TZ>
void launch_pcg_update_x_r(T* x, T* r, const T* p, const T* q, double alpha, size_t n, double* partials, int nparts,
                           TZ* rz_out, hipStream_t s);
// reliable-update PCG (FP32 vectors): r -= alpha q with partials of r.r, and
// dx += alpha p unless dx is null
void launch_pcg_ru_dxr(float* r, const float* q, float* dx, const float* p, double alpha, size_t n, double* partials,
                       int nparts, hipStream_t s, const uint8_t* chunk = nullptr);
// x64 += dx (+ alpha p, a pending increment, when p is given), dx = 0
void launch_pcg_flush_x(double* x, float* dx, size_t n, hipStream_t s, const float* p = nullptr, double alpha = 0.0);
// dx += alpha p, then p = z + beta p (FP32, n a multiple of 4)
void launch_pcg_p_dx(float* p, const float* z, double beta, float* dx, double alpha, size_t n, hipStream_t s,
                     const uint8_t* chunk = nullptr);
// streaming kernels' grid size (partials written by the reducing ones)
#ifndef RT_STREAM_BLOCKS
#define RT_STREAM_BLOCKS 2048
#endif
constexpr int kStreamBlocks = RT_STREAM_BLOCKS;
// chunk mask of the finest level's vectors for the streaming PCG kernels (one
// byte per 4096 floats: 0 where no active node's value lies)
void launch_vec_chunks(const uint8_t* vflag, int N, uint8_t* chunk, hipStream_t s);
template <typename T, typename TZ>
void launch_pcg_update_p(T* p, const TZ* z, double beta, size_t n, hipStream_t s, const uint8_t* chunk = nullptr);
template <typename T, typename TZ>
void launch_sub_conv(const T* b, const T* q, T* r, TZ* rz, size_t n, double* partials, int nparts, hipStream_t s);
template <typename T, typename TE>
void launch_add_conv(T* x, const TE* e, size_t n, hipStream_t s);
template <typename T>
void launch_scale(T* x, double a, size_t n, hipStream_t s);
void launch_ratio(const double* a, const double* b, double* out, hipStream_t s);
template <typename T>
void launch_box_to_oct(const Geom& g, const T* box, T* oct, int node_or_elem, hipStream_t s);
template <typename T>
void launch_oct_to_box(const Geom& g, const T* oct, T* box, int node_or_elem, hipStream_t s);

// k_galerkin.hip
void set_ke_galerkin(const double* ke576, hipStream_t s);
// R[64][nE]: FP64 rho^p of the 64 children of every x4 coarse element (the
// right operand of the Galerkin GEMM KEc = M^T R)
void launch_children64(const Geom& gf, const int ec[3], const float* rhop, double* R, hipStream_t s);
void launch_galerkin4_fix(const Geom& gf, const int ec[3], const float* rhop, const uint8_t* vflag,
                          const uint8_t* eflag, const int* ce, int nce, double* KEc, hipStream_t s);
void launch_galerkin2(const double* KEf, const int ef[3], double* KEp, const int ep[3], hipStream_t s);
template <typename TS>
void launch_assemble(const Geom& g, const double* KEc, const int e[3], const uint8_t* vflag, TS* st, hipStream_t s);

// k_simp.hip
void set_ke_simp(const double* ke576, hipStream_t s);
void launch_rhop(const Geom& g, const float* rho, const uint8_t* eflag, float p, float* rhop, hipStream_t s);
template <typename T>
void launch_sensitivity(const Geom& g, Vec3<const T> U, const float* rho, const uint8_t* eflag, float p, float* sens,
                        hipStream_t s);
void launch_filter(const Geom& g, const float* gin, const uint8_t* eflag, float R, float* gout, hipStream_t s);
void launch_oc_trial(const Geom& g, const float* rho, const float* sens, const uint8_t* eflag, float thr, float step,
                     float damp, float rhomin, float* rho_out, double* partials, int nparts, hipStream_t s);
void launch_maxabs(const Geom& g, const float* a, const uint8_t* eflag, double* partials, int nparts, hipStream_t s);
void launch_sum_active(const Geom& g, const float* a, const uint8_t* eflag, double* partials, int nparts,
                       hipStream_t s);

// k_worst.hip
void launch_rand_force(const Geom& g, const uint8_t* vflag, uint64_t seed, Vec3<double> F, hipStream_t s);
void launch_pm_gather(int nl, const int* slots, Vec3<const double> X, const double* nrm, double* fs,
                      double* partials, int nparts, hipStream_t s);
void launch_pm_commit(int nl, const int* slots, const double* fs, double scale, Vec3<double> F, double* fold,
                      double* partials, int nparts, hipStream_t s);

// k_voxel.hip
void launch_voxelize(const float* tris, int ntris, const float bmin[3], const float unit[3], int R, uint32_t* bits,
                     hipStream_t s);

// The padded cube of a mesh with bounding box [lo, hi] (padd2SatifyResolution,
// voxelize_main.cpp:312-340, and voxinfo, util.h:57-68): side R = the smallest
// power of two above prefer_reso, the box grown about its centre to
// maxlen R / prefer_reso per axis, unit = box / R. Returns R.
inline int voxel_frame(const float lo[3], const float hi[3], int prefer_reso, float bmin[3], float bmax[3],
                       float unit[3]) {
    unsigned padd = 1;
    for (int i = 0; i < 32; i++) {
        if ((unsigned)prefer_reso < padd) break;
        padd <<= 1;
    }
    const float len[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    float maxlen = len[0];
    if (len[1] > maxlen) maxlen = len[1];
    if (len[2] > maxlen) maxlen = len[2];
    for (int j = 0; j < 3; j++) {
        const float pl = maxlen * (float)padd / (float)prefer_reso;
        bmin[j] = lo[j] - (pl - len[j]) / 2;
        bmax[j] = hi[j] + (pl - len[j]) / 2;
        unit[j] = (bmax[j] - bmin[j]) / float(padd);
    }
    return (int)padd;
}

// maps the exception being handled to an RT_E* status and records its message
// for rt_last_error (call only inside a catch block)
int status_of_current_exception();

#define RT_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " __FILE__ \
                                     ":" + std::to_string(__LINE__));                                  \
    } while (0)

// ---------------------------------------------------------------------------
// device buffer
// ---------------------------------------------------------------------------
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b) {
        if (b == bytes && p) { RT_HIP(hipMemset(p, 0, b)); return; }
        release();
        if (b == 0) return;
        RT_HIP(hipMalloc(&p, b));
        RT_HIP(hipMemset(p, 0, b));
        bytes = b;
    }
    void ensure(size_t b) { if (bytes < b) alloc(b); }
    template <class T> T* as() const { return (T*)p; }
};

}  // namespace rt
