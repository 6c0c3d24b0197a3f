// Host side of robtop_amd: hierarchy construction, device memory, the V-cycle
// and MG-PCG drivers, the SIMP steps and the C ABI declared in
// include/robtop_amd.h. Compiled by hipcc as HIP host code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "../../include/robtop_amd.h"
#include "rt_internal.h"
#include "net_internal.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

// dense symmetric eigensolver (defined with the block Lanczos search below)
static void jacobi_eigen(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& V);

namespace rt {

thread_local std::string g_err;

// coarsening stops at the first layer with at most this many solid elements
// (Grid.cpp:229: 400)
#ifndef RT_COARSEST_ELEMS
#define RT_COARSEST_ELEMS 400
#endif

// a 3-component vector: one contiguous buffer of 3N
template <typename T>
struct V3 {
    DBuf b;
    int N = 0;
    bool ok() const { return b.p != nullptr; }
    void alloc(int n) { N = n; b.alloc(sizeof(T) * 3 * (size_t)n); }
    T* c(int k) const { return b.as<T>() + (size_t)k * N; }
    T* data() const { return b.as<T>(); }
    Vec3<T> v() const { return Vec3<T>{{c(0), c(1), c(2)}}; }
    Vec3<const T> cv() const { return Vec3<const T>{{c(0), c(1), c(2)}}; }
};

template <typename T>
struct LevelVecs {
    V3<T> U, F, R;
    bool ok() const { return U.ok(); }
    void alloc(int N) { U.alloc(N); F.alloc(N); R.alloc(N); }
};

// nb vectors of 3N values, vector b at data() + b * 3N (components as in V3)
template <typename T>
struct BVec {
    DBuf b;
    int N = 0, nb = 0;
    bool fits(int n, int k) const { return b.p && N == n && nb >= k; }
    void alloc(int n, int k) { N = n; nb = k; b.alloc(sizeof(T) * 3 * (size_t)n * k); }
    T* data() const { return b.as<T>(); }
    T* at(int i) const { return data() + (size_t)i * 3 * N; }
    Vec3<T> v(int i) const { T* p = at(i); return Vec3<T>{{p, p + N, p + 2 * (size_t)N}}; }
    Vec3<const T> cv(int i) const { const T* p = at(i); return Vec3<const T>{{p, p + N, p + 2 * (size_t)N}}; }
};

struct Level {
    int layer = 0, scale = 1;
    int o[3] = {0, 0, 0};
    int e[3] = {0, 0, 0};
    Geom g;
    std::vector<uint8_t> eact, vact;      // box masks
    std::vector<uint8_t> vflag_box;       // flags in box layout (host copy)
    size_t n_eact = 0, n_vact = 0;
    DBuf vflag;
    LevelVecs<double> v64;
    LevelVecs<float> v32;
    DBuf st64, st32, KEc;
    DBuf st16;                            // bfloat16 copy for FP32 preconditioner cycles (optional)
    DBuf tile_act;                        // finest level: mask of residual / apply tiles holding an active node
    BVec<float> bU, bF, bR;               // batched preconditioner cycle (worst-case block solves)
    size_t nn_box() const { return (size_t)g.n[0] * g.n[1] * g.n[2]; }
    size_t ne_box() const { return (size_t)e[0] * e[1] * e[2]; }
};

struct Hier {
    int R = 0, nlayers = 0;
    std::vector<int> layer_active;
    std::vector<Level> lv;
    hipStream_t s = nullptr;
    double KE[576];
    double E = 1e6, nu = 0.4;
    float p = 3.f;
    DBuf dM;                 // galerkin x4 M_k [64][576]
    DBuf rh64;               // galerkin x4 GEMM operand: rho^p of the 64 children per coarse element, FP64 [64][nE]
    DBuf xfer;               // scratch of the two-stage x4 restriction (sized for FP64)
    // finest level
    DBuf rho, rhop, eflag, sens, sens_tmp;
    std::vector<uint8_t> eflag_box, fixed_box;
    V3<double> x64, f64;     // displacement and force
    // PCG work
    V3<double> r64, p64, q64;
    V3<float> x32, b32, r32, p32, q32;
    V3<float> dx32;          // reliable-update PCG: FP32 solution increment since the last replacement
    LevelVecs<double> pc64;  // preconditioner level-0 vectors
    LevelVecs<float> pc32;
    DBuf partials, scal, fbuf, stage;
    int nparts_max = 0;
    // coarsest: explicit inverse of the dense system (FP64, n3 x n3), factored
    // with rocSOLVER on the handle's stream
    DBuf Ainv, rowslot, dinfo;
    rocblas_handle blas = nullptr;
    int n3 = 0;
    DBuf fixce;
    int nfixce = 0;
    // worst-case load search: load-node slots, unit normals (constrained
    // direction), projected force and the previous step's force, all [3][nload]
    DBuf load_slots, load_nrm, pm_fs, pm_old;
    // no-support modes: the admissible rigid motions of the load nodes
    // (compact [k][c * nload + i], nrd of them) and the finest nodes' rigid
    // motions (6 vectors, built on first use)
    DBuf load_rd;
    int nrd = 0;
    std::vector<V3<double>> rigid_f;
    int nload = 0;
    bool load_normals = false;
    int vprec = 64;
    int pc_mult = 1, pc_gamma = 1;   // preconditioner cycle: coarse-level sweep multiplier, cycle index
    std::vector<int> pc_lv_mult;     // per-level overrides of pc_mult (0: pc_mult)
    // f . u partials of the preconditioner's last fine sweep (its r . z), one
    // run per colour pair at rzpart_off[pair]
    DBuf rzpart;
    int rzpart_n = 0, rzpart_off[4] = {0, 0, 0, 0};
    DBuf vchunk;                     // finest-level vector chunks holding an active node (launch_vec_chunks)
    // The preconditioner cycle may stop one level early and solve that level
    // exactly when it has at most pc_direct_dofs unknowns (pc_levels = the
    // number of levels its cycle visits); its dense inverse lives here.
    int pc_direct_dofs = 6000, pc_levels = 0;
    DBuf pcAinv, pcrowslot, pcfbuf, pcbfbuf;
    int pcn3 = 0;
    int coarse_bits = 32;            // stencil storage of FP32 coarse-level cycles: 32 or 16 (bfloat16)
    int fine_smoother = 0;           // 0: four colour-pair launches in place; 1: two tiled launches, two buffers
    int ws_m = 0;                    // solves start from the best combination of the last ws_m solutions (0: off)
    // floating-cluster deflation of the preconditioner (theta <= 0: off)
    float isl_theta = 0.f;
    int isl_min = 8, isl_max = 4096, nisl = 0;
    DBuf isl_lab, isl_anch, isl_cnt, isl_root, isl_node, isl_cent, isl_E, isl_Einv, isl_zr, isl_coef, isl_flag,
        isl_KE;
    std::vector<V3<double>> ws_hist; // ring of previous solutions
    // learned warm start: network (not owned), its prolonged prediction, and
    // {taken, E(plain start), E(learned start), prediction coefficient} of the last solve
    Net* net = nullptr;
    V3<double> nn_u;
    // CG coefficients of the last reliable-update solve (Lanczos tridiagonal of
    // the preconditioned operator): alpha_j, and beta_j forming p_{j+1}
    std::vector<double> cg_alpha, cg_beta;
    double nn_info[4] = {0, 0, 0, 0};
    int ws_count = 0, ws_head = 0;
    V3<float> gs_tmp32;              // second buffer of the tiled fine smoother
    V3<double> gs_tmp64;
    // the preconditioner cycle replayed as a HIP graph (pc_key identifies its shape)
    bool use_graphs = true;
    hipGraphExec_t pc_graph = nullptr;
    long long pc_key = -1;
    // batched solves: FP64 Krylov vectors per right-hand side, the batched
    // cycle's graph, coarsest gather buffer, per-vector partial sums
    BVec<double> bx, bb, br, bp, bq;
    BVec<float> bp32, bq32, bdx32;   // reliable-update batched solves: FP32 Krylov vectors and increment
    DBuf bfbuf, bpart, bscal;
    hipGraphExec_t bgraph = nullptr;
    long long bkey = -1;
    bool batch_solves = true;        // block Lanczos solves its k right-hand sides together
    bool stencil_ready = false, material_set = false;
    void drop_graph() {
        if (pc_graph) (void)hipGraphExecDestroy(pc_graph);
        pc_graph = nullptr;
        pc_key = -1;
        if (bgraph) (void)hipGraphExecDestroy(bgraph);
        bgraph = nullptr;
        bkey = -1;
    }
    ~Hier() {
        drop_graph();
        if (blas) (void)rocblas_destroy_handle(blas);
        if (s) (void)hipStreamDestroy(s);
    }
};

// ---------------------------------------------------------------------------
// KE for an isotropic unit cube (trilinear element, exact 2-point Gauss):
// a(u,v) = int lam div u div v + mu (grad u : grad v + grad u : grad v^T).
// ---------------------------------------------------------------------------
static void element_matrix(double E, double nu, double KE[576]) {
    const double lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    const double mu = E / (2 * (1 + nu));
    const double gp[2] = {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)};
    std::fill(KE, KE + 576, 0.0);
    for (int gz = 0; gz < 2; gz++)
        for (int gy = 0; gy < 2; gy++)
            for (int gx = 0; gx < 2; gx++) {
                const double xq[3] = {gp[gx], gp[gy], gp[gz]};
                double gr[8][3];
                for (int i = 0; i < 8; i++) {
                    const int b[3] = {i & 1, (i >> 1) & 1, i >> 2};
                    double f[3], d[3];
                    for (int k = 0; k < 3; k++) {
                        f[k] = b[k] ? xq[k] : 1 - xq[k];
                        d[k] = b[k] ? 1.0 : -1.0;
                    }
                    gr[i][0] = d[0] * f[1] * f[2];
                    gr[i][1] = f[0] * d[1] * f[2];
                    gr[i][2] = f[0] * f[1] * d[2];
                }
                for (int i = 0; i < 8; i++)
                    for (int a = 0; a < 3; a++)
                        for (int j = 0; j < 8; j++)
                            for (int b = 0; b < 3; b++) {
                                double v = lam * gr[i][a] * gr[j][b] + mu * gr[i][b] * gr[j][a];
                                if (a == b) v += mu * (gr[i][0] * gr[j][0] + gr[i][1] * gr[j][1] + gr[i][2] * gr[j][2]);
                                KE[(3 * i + a) * 24 + 3 * j + b] += 0.125 * v;
                            }
            }
}

// M_k = P_k^T KE P_k for the 64 children of a x4 coarse element
static void galerkin4_matrices(const double KE[576], std::vector<double>& M) {
    M.assign(64 * 576, 0.0);
    for (int k = 0; k < 64; k++) {
        const int kx = k & 3, ky = (k >> 2) & 3, kz = k >> 4;
        double W[8][8];
        for (int i = 0; i < 8; i++) {
            const int px = kx + (i & 1), py = ky + ((i >> 1) & 1), pz = kz + (i >> 2);
            for (int A = 0; A < 8; A++)
                W[i][A] = ((A & 1) ? px : 4 - px) * ((A & 2) ? py : 4 - py) * ((A & 4) ? pz : 4 - pz) / 64.0;
        }
        double T[24][24];   // KE P
        for (int a = 0; a < 24; a++)
            for (int B = 0; B < 8; B++)
                for (int s = 0; s < 3; s++) {
                    double v = 0;
                    for (int j = 0; j < 8; j++) v += KE[a * 24 + 3 * j + s] * W[j][B];
                    T[a][3 * B + s] = v;
                }
        for (int A = 0; A < 8; A++)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 24; c++) {
                    double v = 0;
                    for (int i = 0; i < 8; i++) v += W[i][A] * T[3 * i + r][c];
                    M[(size_t)k * 576 + (3 * A + r) * 24 + c] = v;
                }
    }
}

// ---------------------------------------------------------------------------
// hierarchy construction
// ---------------------------------------------------------------------------
// The smoothers and coarse kernels load through buffer resources whose byte
// ranges are 32-bit. A level whose vectors (at `value_bytes` per entry) or whose
// coarse stencil (nine FP64 3x3 blocks per slot, the larger precision) would pass
// 2 GiB is refused here rather than read as zeros past the limit.
static void check_load_range(const Level& L, int value_bytes) {
    const long long lim = 1ll << 31;
    const long long vec = (long long)L.g.N * value_bytes;
    const long long st = L.layer > 0 ? 9ll * L.g.N * (long long)sizeof(double) : 0;
    if (vec >= lim || st >= lim)
        throw std::invalid_argument("level " + std::to_string(L.layer) + " has " + std::to_string(L.g.N) +
                                    " slots: its " + (vec >= lim ? std::to_string(value_bytes * 8) + "-bit vectors"
                                                                  : std::string("coarse stencil")) +
                                    " pass the 2 GiB range of 32-bit buffer offsets");
}

static Hier* build(int R, const uint32_t* bits, size_t nwords) {
    if (R < 4 || (R & (R - 1))) throw std::invalid_argument("R must be a power of two >= 4");
    const size_t n0 = (size_t)R * R * R;
    if (nwords * 32 < n0) throw std::invalid_argument("solid bit array too short for R^3 bits");
    std::vector<std::vector<uint8_t>> act;
    {
        std::vector<uint8_t> a(n0);
        for (size_t i = 0; i < n0; i++) a[i] = (bits[i >> 5] >> (i & 31)) & 1u;
        act.push_back(std::move(a));
    }
    std::vector<int> counts;
    auto count = [](const std::vector<uint8_t>& a) {
        size_t c = 0;
        for (uint8_t v : a) c += v;
        return (int)c;
    };
    counts.push_back(count(act[0]));
    if (counts[0] == 0) throw std::invalid_argument("no solid element");
    int reso = R;
    while (counts.back() > RT_COARSEST_ELEMS && reso > 1) {
        const int fr = reso;
        reso >>= 1;
        std::vector<uint8_t> c((size_t)reso * reso * reso, 0);
        const std::vector<uint8_t>& f = act.back();
        for (int z = 0; z < fr; z++)
            for (int y = 0; y < fr; y++) {
                const uint8_t* row = f.data() + (size_t)fr * (y + (size_t)fr * z);
                uint8_t* crow = c.data() + (size_t)reso * ((y >> 1) + (size_t)reso * (z >> 1));
                for (int x = 0; x < fr; x++)
                    if (row[x]) crow[x >> 1] = 1;
            }
        act.push_back(std::move(c));
        counts.push_back(count(act.back()));
    }
    const int nl = (int)act.size();
    if (nl < 3) throw std::invalid_argument("fewer than three layers: the x4 first coarsening needs layers 0 and 2");
    const int K = nl - 1, A = 1 << K;
    int lo[3] = {R, R, R}, hi[3] = {0, 0, 0};
    for (int z = 0; z < R; z++)
        for (int y = 0; y < R; y++) {
            const uint8_t* row = act[0].data() + (size_t)R * (y + (size_t)R * z);
            int xa = -1, xb = -1;
            for (int x = 0; x < R; x++)
                if (row[x]) { if (xa < 0) xa = x; xb = x; }
            if (xa < 0) continue;
            lo[0] = std::min(lo[0], xa); hi[0] = std::max(hi[0], xb + 1);
            lo[1] = std::min(lo[1], y); hi[1] = std::max(hi[1], y + 1);
            lo[2] = std::min(lo[2], z); hi[2] = std::max(hi[2], z + 1);
        }
    std::unique_ptr<Hier> h(new Hier());
    h->R = R;
    h->nlayers = nl;
    h->layer_active = counts;
    for (int layer = 0; layer < nl; layer++) {
        if (layer == 1) continue;
        Level L;
        L.layer = layer;
        L.scale = 1 << layer;
        const int r = R >> layer;
        for (int k = 0; k < 3; k++) {
            const int l0 = (lo[k] / A) * A, h0 = std::min(R, ((hi[k] + A - 1) / A) * A);
            L.o[k] = l0 >> layer;
            L.e[k] = (h0 - l0) >> layer;
        }
        L.g = make_geom(L.e[0] + 1, L.e[1] + 1, L.e[2] + 1);
        check_load_range(L, 4);
        L.eact.assign(L.ne_box(), 0);
        for (int z = 0; z < L.e[2]; z++)
            for (int y = 0; y < L.e[1]; y++)
                for (int x = 0; x < L.e[0]; x++) {
                    const size_t gi = (size_t)(x + L.o[0]) + (size_t)r * ((size_t)(y + L.o[1]) + (size_t)r * (z + L.o[2]));
                    if (act[layer][gi]) { L.eact[x + (size_t)L.e[0] * (y + (size_t)L.e[1] * z)] = 1; L.n_eact++; }
                }
        const int nx = L.g.n[0], ny = L.g.n[1], nz = L.g.n[2];
        L.vact.assign(L.nn_box(), 0);
        L.vflag_box.assign(L.nn_box(), 0);
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    int ne = 0;
                    for (int w = 0; w < 8; w++) {
                        const int ex = x + (w & 1) - 1, ey = y + ((w >> 1) & 1) - 1, ez = z + (w >> 2) - 1;
                        if (ex < 0 || ey < 0 || ez < 0 || ex >= L.e[0] || ey >= L.e[1] || ez >= L.e[2]) continue;
                        ne += L.eact[ex + (size_t)L.e[0] * (ey + (size_t)L.e[1] * ez)];
                    }
                    const size_t v = x + (size_t)nx * (y + (size_t)ny * z);
                    if (ne) {
                        L.vact[v] = 1;
                        L.vflag_box[v] = (uint8_t)(NF_ACTIVE | (ne << 4));
                        L.n_vact++;
                    }
                }
        h->lv.push_back(std::move(L));
    }
    // the first device call, after every check of the input
    RT_HIP(hipStreamCreate(&h->s));
    return h.release();
}

// ---------------------------------------------------------------------------
// transfers between host box arrays and device octant arrays
// ---------------------------------------------------------------------------
template <typename T>
static void upload_box(Hier& H, const Geom& g, int elem, const T* host, T* dev) {
    const int d = elem ? 1 : 0;
    const size_t cnt = (size_t)(g.n[0] - d) * (g.n[1] - d) * (g.n[2] - d);
    H.stage.ensure(cnt * sizeof(T));
    RT_HIP(hipMemcpyAsync(H.stage.p, host, cnt * sizeof(T), hipMemcpyHostToDevice, H.s));
    launch_box_to_oct<T>(g, H.stage.as<T>(), dev, elem, H.s);
    RT_HIP(hipStreamSynchronize(H.s));
}

template <typename T>
static void download_box(Hier& H, const Geom& g, int elem, const T* dev, T* host) {
    const int d = elem ? 1 : 0;
    const size_t cnt = (size_t)(g.n[0] - d) * (g.n[1] - d) * (g.n[2] - d);
    H.stage.ensure(cnt * sizeof(T));
    launch_oct_to_box<T>(g, dev, H.stage.as<T>(), elem, H.s);
    RT_HIP(hipMemcpyAsync(host, H.stage.p, cnt * sizeof(T), hipMemcpyDeviceToHost, H.s));
    RT_HIP(hipStreamSynchronize(H.s));
}

// 3-vector of FP64 box data -> device T vector
template <typename T>
static void upload_vec(Hier& H, const Geom& g, const double* box, V3<T>& v) {
    const size_t nb = (size_t)g.n[0] * g.n[1] * g.n[2];
    std::vector<T> tmp(nb);
    for (int c = 0; c < 3; c++) {
        for (size_t i = 0; i < nb; i++) tmp[i] = (T)box[c * nb + i];
        upload_box<T>(H, g, 0, tmp.data(), v.c(c));
    }
}
template <typename T>
static void download_vec(Hier& H, const Geom& g, const V3<T>& v, double* box) {
    const size_t nb = (size_t)g.n[0] * g.n[1] * g.n[2];
    std::vector<T> tmp(nb);
    for (int c = 0; c < 3; c++) {
        download_box<T>(H, g, 0, v.c(c), tmp.data());
        for (size_t i = 0; i < nb; i++) box[c * nb + i] = (double)tmp[i];
    }
}

static void upload_flags(Hier& H, Level& L) {
    L.vflag.alloc(L.g.N);
    upload_box<uint8_t>(H, L.g, 0, L.vflag_box.data(), L.vflag.as<uint8_t>());
}

static void ensure_partials(Hier& H) {
    if (H.partials.p) return;
    int need = 65536;
    need = std::max(need, fine_apply_partials_modal(H.lv[0].g));
    for (auto& L : H.lv) need = std::max(need, (int)((L.g.N + 255) / 256));
    H.nparts_max = need;
    H.partials.alloc(sizeof(double) * (size_t)need);
    H.scal.alloc(sizeof(double) * 16);
}

// Finest-level work mask (Geom::tile_act) from the active nodes: the residual /
// apply kernels skip tiles that hold no active node (their outputs are absent
// slots, never written). Rebuilt whenever Geom::zc changes.
static void build_tile_mask(Level& L) {
    int tx, ty;
    modal_tile_dims(&tx, &ty);
    const int nx = L.g.n[0], ny = L.g.n[1], nz = L.g.n[2], tz = L.g.zc;
    const int gx = (nx + tx - 1) / tx, gy = (ny + ty - 1) / ty, gz = (nz + tz - 1) / tz;
    std::vector<uint8_t> m((size_t)gx * gy * gz, 0);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++) {
            const uint8_t* row = L.vact.data() + (size_t)nx * ((size_t)y + (size_t)ny * z);
            for (int x = 0; x < nx; x++)
                if (row[x]) m[(size_t)(x / tx) + (size_t)gx * ((size_t)(y / ty) + (size_t)gy * (z / tz))] = 1;
        }
    L.tile_act.alloc(m.size());
    RT_HIP(hipMemcpy(L.tile_act.p, m.data(), m.size(), hipMemcpyHostToDevice));
    L.g.tile_act = L.tile_act.as<uint8_t>();
}

static void setup_device(Hier& H) {
    for (auto& L : H.lv) upload_flags(H, L);
    Level& L0 = H.lv[0];
    const int N0 = L0.g.N;
    H.eflag_box.assign(L0.ne_box(), 0);
    for (size_t i = 0; i < L0.ne_box(); i++) H.eflag_box[i] = L0.eact[i] ? EF_ACTIVE : 0;
    H.eflag.alloc(N0);
    upload_box<uint8_t>(H, L0.g, 1, H.eflag_box.data(), H.eflag.as<uint8_t>());
    H.rho.alloc(sizeof(float) * N0);
    H.rhop.alloc(sizeof(float) * N0);
    H.sens.alloc(sizeof(float) * N0);
    H.sens_tmp.alloc(sizeof(float) * N0);
    H.x64.alloc(N0);
    H.f64.alloc(N0);
    H.fixed_box.assign(L0.nn_box(), 0);
    for (size_t li = 1; li < H.lv.size(); li++) {
        Level& L = H.lv[li];
        L.KEc.alloc(sizeof(double) * 576 * L.ne_box());
        L.st64.alloc(sizeof(double) * 243 * (size_t)L.g.N);
        L.st32.alloc(sizeof(float) * 243 * (size_t)L.g.N);
        if (L.scale / H.lv[li - 1].scale == 4)
            H.xfer.ensure(sizeof(double) * restrict_scratch_elems(H.lv[li - 1].g, L.g));
    }
    ensure_partials(H);
    build_tile_mask(L0);
}

template <typename T> LevelVecs<T>& vecs(Level& L);
template <> LevelVecs<double>& vecs<double>(Level& L) { if (!L.v64.ok()) L.v64.alloc(L.g.N); return L.v64; }
template <> LevelVecs<float>& vecs<float>(Level& L) { if (!L.v32.ok()) L.v32.alloc(L.g.N); return L.v32; }
template <typename T> const T* stencil(const Level& L);
template <> const double* stencil<double>(const Level& L) { return L.st64.as<double>(); }
template <> const float* stencil<float>(const Level& L) { return L.st32.as<float>(); }
template <typename T> LevelVecs<T>& pcvecs(Hier& H);
template <> LevelVecs<double>& pcvecs<double>(Hier& H) { if (!H.pc64.ok()) H.pc64.alloc(H.lv[0].g.N); return H.pc64; }
template <> LevelVecs<float>& pcvecs<float>(Hier& H) { if (!H.pc32.ok()) H.pc32.alloc(H.lv[0].g.N); return H.pc32; }

template <typename T> V3<T>& gs_tmp(Hier& H);
template <> V3<double>& gs_tmp<double>(Hier& H) { if (!H.gs_tmp64.ok()) H.gs_tmp64.alloc(H.lv[0].g.N); return H.gs_tmp64; }
template <> V3<float>& gs_tmp<float>(Hier& H) { if (!H.gs_tmp32.ok()) H.gs_tmp32.alloc(H.lv[0].g.N); return H.gs_tmp32; }

// ---------------------------------------------------------------------------
// multigrid operations (level index li into H.lv; the finest level is 0)
// ---------------------------------------------------------------------------
template <typename T>
// zero_support: u and f are 0 at every support node of the finest level (the
// preconditioner's case), so the smoother skips neighbour support checks.
// zero_start: U holds no valid values and the smoothing starts from u = 0; the
// colour-pair fine smoother's first forward sweep takes that start itself
// (reading no neighbour that is still 0), otherwise U is cleared first.
// fu_part: the last sweep also leaves the partial sums of f . u_new (the
// colour-pair fine smoother only; see rz_fused)
static void gs(Hier& H, int li, Vec3<T> U, Vec3<const T> F, int n, bool reverse, bool zero_support = false,
               bool zero_start = false, double* fu_part = nullptr) {
    Level& L = H.lv[li];
    const bool z0 = zero_start && li == 0 && H.fine_smoother == 0 && n > 0 && !reverse && zero_support;
    if (zero_start && !z0)
        for (int c = 0; c < 3; c++) launch_fill<T>(U.c[c], T(0), (size_t)L.g.N, H.s);
    for (int it = 0; it < n; it++) {
        if (li == 0 && H.fine_smoother == 1) {
            // two launches per sweep, ping-ponging between U and the second buffer
            V3<T>& tmp = gs_tmp<T>(H);
            const bool in_u = it % 2 == 0;
            const Vec3<T> Tv = tmp.v();
            launch_fine_gs_sweep2<T>(L.g, in_u ? Vec3<const T>{{U.c[0], U.c[1], U.c[2]}} : tmp.cv(), in_u ? Tv : U, F,
                                     H.rhop.as<float>(), L.vflag.as<uint8_t>(), reverse, zero_support, H.s);
            if (in_u && it == n - 1)
                for (int c = 0; c < 3; c++) launch_convert<T, T>(Tv.c[c], U.c[c], (size_t)L.g.N, H.s);
        } else if (li == 0) {
            // four launches, each relaxing the colour pair (2p, 2p+1) row by row
            for (int pp = 0; pp < 4; pp++) {
                const int p = reverse ? 3 - pp : pp;
                launch_fine_gs<T>(L.g, U, F, H.rhop.as<float>(), L.vflag.as<uint8_t>(), p, reverse, zero_support,
                                  H.s, z0 && it == 0, fu_part && it == n - 1 ? fu_part + H.rzpart_off[p] : nullptr);
            }
        } else if (coarse_gs_pairs(L.g)) {
            // four launches, each relaxing the colour pair (2p, 2p+1) row by row
            for (int pp = 0; pp < 4; pp++) {
                const int p = reverse ? 3 - pp : pp;
                if constexpr (std::is_same<T, float>::value) {
                    if (H.coarse_bits == 16) {
                        launch_coarse_gs_pair<float, bf16>(L.g, U, F, L.st16.as<bf16>(), L.vflag.as<uint8_t>(), p,
                                                           reverse, H.s);
                        continue;
                    }
                }
                launch_coarse_gs_pair<T, T>(L.g, U, F, stencil<T>(L), L.vflag.as<uint8_t>(), p, reverse, H.s);
            }
        } else {
            for (int cc = 0; cc < 8; cc++) {
                const int c = reverse ? 7 - cc : cc;
                if constexpr (std::is_same<T, float>::value) {
                    if (H.coarse_bits == 16) {
                        launch_coarse_gs<float, bf16>(L.g, U, F, L.st16.as<bf16>(), L.vflag.as<uint8_t>(), c, reverse,
                                                      H.s);
                        continue;
                    }
                }
                launch_coarse_gs<T, T>(L.g, U, F, stencil<T>(L), L.vflag.as<uint8_t>(), c, reverse, H.s);
            }
        }
    }
}

template <typename T>
static void residual(Hier& H, int li, Vec3<const T> U, Vec3<const T> F, Vec3<T> R) {
    Level& L = H.lv[li];
    if (li == 0)
        launch_fine_residual_modal<T>(L.g, U, F, R, H.rhop.as<float>(), L.vflag.as<uint8_t>(), H.s);
    else {
        if constexpr (std::is_same<T, float>::value) {
            if (H.coarse_bits == 16) {
                launch_coarse_residual<float, bf16>(L.g, U, F, R, L.st16.as<bf16>(), L.vflag.as<uint8_t>(), H.s);
                return;
            }
        }
        launch_coarse_residual<T, T>(L.g, U, F, R, stencil<T>(L), L.vflag.as<uint8_t>(), H.s);
    }
}

template <typename T>
static void restrict_to(Hier& H, int li, Vec3<const T> Rf, Vec3<T> Fc) {
    Level& C = H.lv[li];
    Level& Fl = H.lv[li - 1];
    const int ratio = C.scale / Fl.scale;
    T* scratch = nullptr;
    if (ratio == 4 && H.xfer.bytes >= sizeof(T) * restrict_scratch_elems(Fl.g, C.g)) scratch = H.xfer.as<T>();
    launch_restrict<T>(Fl.g, Fl.o, Rf, C.g, C.o, Fc, C.vflag.as<uint8_t>(), ratio, scratch, H.s);
}

template <typename T>
// skip_fixed leaves support nodes untouched; the smoother that follows resets
// them to f / n_e and their neighbours ignore them, so the cycle's result is
// the same either way (Grid.cu:1818-1895 corrects them too)
static void prolong_from(Hier& H, int li, Vec3<T> Uf, Vec3<const T> Uc, bool skip_fixed = false) {
    Level& Fl = H.lv[li];
    Level& C = H.lv[li + 1];
    launch_prolong<T>(Fl.g, Fl.o, Uf, Fl.vflag.as<uint8_t>(), C.g, C.o, Uc, C.scale / Fl.scale, skip_fixed, H.s);
}

template <typename T>
static void coarsest(Hier& H, Vec3<const T> F, Vec3<T> U) {
    if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    launch_coarsest<T>(H.Ainv.as<double>(), H.n3, H.rowslot.as<int>(), 0, F, U, H.fbuf.as<double>(), H.s);
}

// whether the preconditioner's last fine sweep yields r . z itself (the
// colour-pair smoother and at least one post-smoothing sweep)
static bool rz_fused(const Hier& H, int post) { return H.fine_smoother == 0 && post > 0; }

static double* ensure_rzpart(Hier& H) {
    if (!H.rzpart.p) {
        int off = 0;
        for (int p = 0; p < 4; p++) {
            H.rzpart_off[p] = off;
            off += fine_gs_pair_blocks(H.lv[0].g, p);
        }
        H.rzpart_n = off;
        H.rzpart.alloc(sizeof(double) * (size_t)std::max(off, 1));
    }
    return H.rzpart.as<double>();
}

// the chunk mask of the finest level's vectors (built once: node activity is
// fixed with the hierarchy); the reliable-update PCG's streaming kernels skip
// the chunks whose values are all 0
static const uint8_t* vec_chunks(Hier& H) {
    if (!H.vchunk.p) {
        const Level& L0 = H.lv[0];
        const size_t nch = ((size_t)3 * L0.g.N + 4095) / 4096;
        H.vchunk.alloc(nch);
        launch_vec_chunks(L0.vflag.as<uint8_t>(), L0.g.N, H.vchunk.as<uint8_t>(), H.s);
    }
    return H.vchunk.as<uint8_t>();
}

// sweep multiplier of coarse level i in the preconditioner's cycle
static int level_mult(const Hier& H, int i) {
    return i < (int)H.pc_lv_mult.size() && H.pc_lv_mult[i] > 0 ? H.pc_lv_mult[i] : H.pc_mult;
}

// V(pre, post) with caller-owned finest vectors (Grid.cpp:1068-1113 ordering).
template <typename T>
// zero_support: F0 and U0 vanish at support nodes (the PCG preconditioner), so
// the finest smoother runs without neighbour support checks and the
// prolongation keeps supports at 0.
// islands: the floating-cluster correction U0 += sum_i Z_i E_i^+ Z_i^T R0 joins
// the coarse-grid correction of the finest level, from the same residual R0
// (an additive pair of coarse spaces inside the multiplicative cycle, so the
// cycle stays symmetric when its smoothing is)
// u0_unset: U0 holds no valid values and the cycle starts from U0 = 0 (the
// first smoothing pass clears it or starts from zero itself)
static void vcycle_core(Hier& H, V3<T>& U0, V3<T>& F0, V3<T>& R0, int pre, int post, bool symmetric,
                        int coarse_mult = 1, int gamma = 1, bool zero_support = false, bool islands = false,
                        bool u0_unset = false, double* fu_part = nullptr) {
    // the preconditioner (zero_support) may stop a level early and solve it
    // exactly (pc_levels); the reference-semantics cycle visits every level
    const int n = zero_support && H.pc_levels > 0 ? H.pc_levels : (int)H.lv.size();
    const bool early = n < (int)H.lv.size();
    auto Uv = [&](int i) { return i == 0 ? U0.v() : vecs<T>(H.lv[i]).U.v(); };
    auto Fv = [&](int i) { return i == 0 ? F0.cv() : vecs<T>(H.lv[i]).F.cv(); };
    auto Rv = [&](int i) { return i == 0 ? R0.v() : vecs<T>(H.lv[i]).R.v(); };
    auto cvt = [](Vec3<T> a) { return Vec3<const T>{{a.c[0], a.c[1], a.c[2]}}; };
    // level i: pre-smooth, restrict the residual, gamma coarse cycles from zero,
    // prolong, post-smooth (Grid.cpp:1068-1113 for gamma = 1, no multiplier);
    // levels below the finest smooth coarse_mult times as often
    auto cyc = [&](auto&& self, int i) -> void {
        if (i == n - 1) {
            if (early)
                launch_coarsest<T>(H.pcAinv.as<double>(), H.pcn3, H.pcrowslot.as<int>(), 0, Fv(i), Uv(i),
                                   H.pcfbuf.as<double>(), H.s);
            else
                coarsest<T>(H, Fv(i), Uv(i));
            return;
        }
        const int m = i == 0 ? 1 : zero_support ? level_mult(H, i) : coarse_mult;
        const bool zs = i == 0 && zero_support;
        gs<T>(H, i, Uv(i), Fv(i), pre * m, false, zs, i == 0 && u0_unset);
        residual<T>(H, i, cvt(Uv(i)), Fv(i), Rv(i));
        LevelVecs<T>& lvv = vecs<T>(H.lv[i + 1]);
        restrict_to<T>(H, i + 1, cvt(Rv(i)), lvv.F.v());
        launch_fill<T>(lvv.U.data(), T(0), (size_t)3 * lvv.U.N, H.s);
        const int visits = i + 1 == n - 1 ? 1 : gamma;
        for (int v = 0; v < visits; v++) self(self, i + 1);
        prolong_from<T>(H, i, Uv(i), cvt(Uv(i + 1)), zs);
        if (i == 0 && islands)
            island_correct<T>(H.lv[0].g, H.nisl, H.isl_node.as<int>(), H.isl_cent.as<double>(),
                              H.isl_Einv.as<double>(), H.isl_zr.as<double>(), H.isl_coef.as<double>(),
                              cvt(Rv(0)), Uv(0), H.s);
        gs<T>(H, i, Uv(i), Fv(i), post * m, symmetric, zs, false, i == 0 ? fu_part : nullptr);
    };
    cyc(cyc, 0);
}

// z = M^-1 r for the PCG: one preconditioner cycle on pc.F from pc.U = 0 (the
// first fine sweep starts from zero without reading pc.U, or pc.U is cleared
// by a fill kernel). After an eager first application (which also performs
// lazy allocations) the cycle is captured once into a HIP graph and replayed;
// the graph holds no memset nodes and is dropped whenever stencils, material,
// supports or the cycle shape change.
template <typename TC>
static void precondition(Hier& H, LevelVecs<TC>& pc, int pre, int post) {
    const long long key = ((((long long)sizeof(TC) * 64 + pre) * 64 + post) * 64 + H.pc_mult) * 4 + H.pc_gamma;
    if (H.use_graphs && H.pc_graph && H.pc_key == key) {
        RT_HIP(hipGraphLaunch(H.pc_graph, H.s));
        return;
    }
    double* fu = std::is_same<TC, float>::value && rz_fused(H, post) ? ensure_rzpart(H) : nullptr;
    vcycle_core<TC>(H, pc.U, pc.F, pc.R, pre, post, true, H.pc_mult, H.pc_gamma, true, true, true, fu);
    if (!H.use_graphs) return;
    H.drop_graph();
    RT_HIP(hipStreamBeginCapture(H.s, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
        vcycle_core<TC>(H, pc.U, pc.F, pc.R, pre, post, true, H.pc_mult, H.pc_gamma, true, true, true, fu);
    } catch (...) {
        (void)hipStreamEndCapture(H.s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    RT_HIP(hipStreamEndCapture(H.s, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    RT_HIP(e);
    H.pc_graph = ex;
    H.pc_key = key;
}

template <typename TA, typename TB>
static double dot(Hier& H, const TA* a, const TB* b, size_t n) {
    const int np = kStreamBlocks;
    launch_dot_partials_t<TA, TB>(a, b, n, H.partials.as<double>(), np, H.s);
    launch_sum_partials(H.partials.as<double>(), np, H.scal.as<double>(), H.s);
    double v;
    RT_HIP(hipMemcpyAsync(&v, H.scal.p, sizeof(double), hipMemcpyDeviceToHost, H.s));
    RT_HIP(hipStreamSynchronize(H.s));
    return v;
}

static double sum_partials_at(Hier& H, const double* p, int np) {
    launch_sum_partials(p, np, H.scal.as<double>(), H.s);
    double v;
    RT_HIP(hipMemcpyAsync(&v, H.scal.p, sizeof(double), hipMemcpyDeviceToHost, H.s));
    RT_HIP(hipStreamSynchronize(H.s));
    return v;
}
static double sum_partials_host(Hier& H, int np) {
    launch_sum_partials(H.partials.as<double>(), np, H.scal.as<double>(), H.s);
    double v;
    RT_HIP(hipMemcpyAsync(&v, H.scal.p, sizeof(double), hipMemcpyDeviceToHost, H.s));
    RT_HIP(hipStreamSynchronize(H.s));
    return v;
}

// r . z after precondition<float>(H, pc, pre, post): from the last fine
// sweep's partials when it wrote them, else a separate pass over both vectors
static double precond_rz(Hier& H, LevelVecs<float>& pc, int post, size_t n) {
    if (!rz_fused(H, post)) {
        launch_dot_partials_t<float, float>(pc.F.data(), pc.U.data(), n, H.partials.as<double>(), kStreamBlocks, H.s,
                                            vec_chunks(H));
        return sum_partials_host(H, kStreamBlocks);
    }
    launch_sum_partials(H.rzpart.as<double>(), H.rzpart_n, H.scal.as<double>(), H.s);
    double v;
    RT_HIP(hipMemcpyAsync(&v, H.scal.p, sizeof(double), hipMemcpyDeviceToHost, H.s));
    RT_HIP(hipStreamSynchronize(H.s));
    return v;
}

// ---------------------------------------------------------------------------
// Galerkin + coarsest
// ---------------------------------------------------------------------------
// the hierarchy's rocBLAS handle, on its stream
static void ensure_blas(Hier& H) {
    if (!H.blas) {
        if (rocblas_create_handle(&H.blas) != rocblas_status_success) throw std::runtime_error("rocblas_create_handle failed");
    }
    if (rocblas_set_stream(H.blas, H.s) != rocblas_status_success) throw std::runtime_error("rocblas_set_stream failed");
}

static void gemm(Hier& H, rocblas_operation ta, rocblas_operation tb, int M, int N, int K, double alpha,
                 const double* A, int lda, const double* B, int ldb, double beta, double* C, int ldc) {
    if (rocblas_dgemm(H.blas, ta, tb, M, N, K, &alpha, A, lda, B, ldb, &beta, C, ldc) != rocblas_status_success)
        throw std::runtime_error("rocblas_dgemm failed");
}

// whether any finest node is a support (the with-support modes); without
// supports the stiffness has the rigid motions as its null space
static bool has_supports(const Hier& H) {
    for (uint8_t f : H.fixed_box) if (f) return true;
    return false;
}

// rigid motions at points p (3 per point) as 6 columns of 3 n dofs (dof
// 3 i + c, column-major, ld = 3 n): e_x, e_y, e_z and the rotations [phat(p)]
// of projection.cpp:46-112; then modified Gram-Schmidt in that order, a
// column dropped when it falls below 1e-6 of its norm (projection.cpp:232-257).
// Returns the number of columns kept, compacted to the front.
static int rigid_basis(const std::vector<double>& p, std::vector<double>& R) {
    const size_t n = p.size() / 3, m = 3 * n;
    R.assign(m * 6, 0.0);
    for (size_t i = 0; i < n; i++) {
        const double x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
        const double ph[3][3] = {{0, -z, y}, {z, 0, -x}, {-y, x, 0}};
        for (int c = 0; c < 3; c++) {
            R[c * m + 3 * i + c] = 1.0;
            for (int k = 0; k < 3; k++) R[(3 + k) * m + 3 * i + c] = ph[c][k];
        }
    }
    bool keep[6];
    for (int i = 0; i < 6; i++) {
        double* ci = R.data() + i * m;
        double old = 0;
        for (size_t r = 0; r < m; r++) old += ci[r] * ci[r];
        old = std::sqrt(old);
        for (int j = 0; j < i; j++) {
            if (!keep[j]) continue;
            const double* cj = R.data() + j * m;
            double w = 0;
            for (size_t r = 0; r < m; r++) w += ci[r] * cj[r];
            for (size_t r = 0; r < m; r++) ci[r] -= w * cj[r];
        }
        double res = 0;
        for (size_t r = 0; r < m; r++) res += ci[r] * ci[r];
        res = std::sqrt(res);
        keep[i] = !(res < 1e-6 * old || old < 1e-6);
        if (keep[i])
            for (size_t r = 0; r < m; r++) ci[r] /= res;
    }
    int k = 0;
    for (int i = 0; i < 6; i++) {
        if (!keep[i]) continue;
        if (k != i) std::copy(R.begin() + i * m, R.begin() + (i + 1) * m, R.begin() + k * m);
        k++;
    }
    R.resize((size_t)k * m);
    return k;
}

// dense exact solve of level li: the matrix assembled on the device from its
// FP64 stencils, Cholesky and inverse there too (rocSOLVER, FP64); Ainv
// n3 x n3, rowslot[r] = slot of the level's r-th active node
static void factor_level(Hier& H, int li, DBuf& Ainv, DBuf& rowslot, DBuf& fbuf, int& n3_out) {
    Level& L = H.lv[li];
    const size_t N = (size_t)L.g.N;
    const int nx = L.g.n[0], ny = L.g.n[1], nz = L.g.n[2];
    std::vector<int> slots, rowof(N, -1);
    std::vector<double> pos;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++)
                if (L.vact[x + (size_t)nx * (y + (size_t)ny * z)]) {
                    const int s = oct_p(L.g, x, y, z);
                    rowof[s] = (int)slots.size();
                    slots.push_back(s);
                    pos.push_back(x); pos.push_back(y); pos.push_back(z);
                }
    const int nr = (int)slots.size(), n = 3 * nr;
    std::vector<int> nbcol((size_t)27 * nr, -1);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const int r = rowof[oct_p(L.g, x, y, z)];
                if (r < 0) continue;
                for (int nb = 0; nb < 27; nb++) {
                    const int jx = x + nb % 3 - 1, jy = y + nb / 3 % 3 - 1, jz = z + nb / 9 - 1;
                    if (jx < 0 || jy < 0 || jz < 0 || jx >= nx || jy >= ny || jz >= nz) continue;
                    nbcol[(size_t)27 * r + nb] = rowof[oct_p(L.g, jx, jy, jz)];
                }
            }
    rowslot.ensure(sizeof(int) * (size_t)std::max(nr, 1));
    DBuf nbc;
    nbc.alloc(sizeof(int) * std::max<size_t>(nbcol.size(), 1));
    RT_HIP(hipMemcpyAsync(rowslot.p, slots.data(), sizeof(int) * (size_t)nr, hipMemcpyHostToDevice, H.s));
    RT_HIP(hipMemcpyAsync(nbc.p, nbcol.data(), sizeof(int) * nbcol.size(), hipMemcpyHostToDevice, H.s));
    Ainv.ensure(sizeof(double) * std::max<size_t>((size_t)n * n, 1));
    launch_dense_from_stencil(L.st64.as<double>(), N, rowslot.as<int>(), nbc.as<int>(), nr, Ainv.as<double>(), H.s);
    ensure_blas(H);
    // No supports (the no-support modes): the level's matrix has the rigid
    // motions R of its nodes as null space. The reference's coarsest solve is
    // BDCSVD's minimum-norm solution (Grid.cpp:1981-2043); the same is
    // P (A + a R R^T)^-1 P b with P = I - R R^T and a the mean diagonal, so
    // A + a R R^T is factored and the stored inverse is P (A + a R R^T)^-1 P.
    const bool singular = !has_supports(H);
    DBuf Rd_;
    if (singular) {
        std::vector<double> R;
        if (rigid_basis(pos, R) != 6) throw std::domain_error("level has fewer than 6 rigid motions");
        Rd_.alloc(sizeof(double) * R.size());
        RT_HIP(hipMemcpyAsync(Rd_.p, R.data(), sizeof(double) * R.size(), hipMemcpyHostToDevice, H.s));
        std::vector<double> diag(n);
        RT_HIP(hipMemcpy2DAsync(diag.data(), sizeof(double), Ainv.p, sizeof(double) * (size_t)(n + 1), sizeof(double),
                                n, hipMemcpyDeviceToHost, H.s));
        RT_HIP(hipStreamSynchronize(H.s));
        double a = 0;
        for (double d : diag) a += d;
        a /= n;
        gemm(H, rocblas_operation_none, rocblas_operation_transpose, n, n, 6, a, Rd_.as<double>(), n,
             Rd_.as<double>(), n, 1.0, Ainv.as<double>(), n);
    }
    // Cholesky and inverse (rocSOLVER, FP64), then mirror the triangle the
    // inverse was written to so the GEMV can read whole rows
    H.dinfo.ensure(sizeof(int));
    auto info_ok = [&](const char* what) {
        int info = 0;
        RT_HIP(hipMemcpyAsync(&info, H.dinfo.p, sizeof(int), hipMemcpyDeviceToHost, H.s));
        RT_HIP(hipStreamSynchronize(H.s));
        if (info != 0)
            throw std::domain_error(std::string("coarsest system: ") + what + " failed (not positive definite; no supports?)");
    };
    if (rocsolver_dpotrf(H.blas, rocblas_fill_lower, n, Ainv.as<double>(), n, H.dinfo.as<int>()) !=
        rocblas_status_success)
        throw std::runtime_error("rocsolver_dpotrf failed");
    info_ok("Cholesky factorisation");
    if (rocsolver_dpotri(H.blas, rocblas_fill_lower, n, Ainv.as<double>(), n, H.dinfo.as<int>()) !=
        rocblas_status_success)
        throw std::runtime_error("rocsolver_dpotri failed");
    info_ok("inverse");
    launch_mirror_upper(Ainv.as<double>(), n, H.s);
    if (singular) {
        // M <- P M P = M - G R^T - R G^T + R (R^T G) R^T, G = M R
        const double* R = Rd_.as<double>();
        DBuf G, T, S;
        G.alloc(sizeof(double) * (size_t)n * 6);
        T.alloc(sizeof(double) * (size_t)n * 6);
        S.alloc(sizeof(double) * 36);
        double* M = Ainv.as<double>();
        gemm(H, rocblas_operation_none, rocblas_operation_none, n, 6, n, 1.0, M, n, R, n, 0.0, G.as<double>(), n);
        gemm(H, rocblas_operation_transpose, rocblas_operation_none, 6, 6, n, 1.0, R, n, G.as<double>(), n, 0.0,
             S.as<double>(), 6);
        gemm(H, rocblas_operation_none, rocblas_operation_none, n, 6, 6, 1.0, R, n, S.as<double>(), 6, 0.0,
             T.as<double>(), n);
        gemm(H, rocblas_operation_none, rocblas_operation_transpose, n, n, 6, -1.0, G.as<double>(), n, R, n, 1.0, M, n);
        gemm(H, rocblas_operation_none, rocblas_operation_transpose, n, n, 6, -1.0, R, n, G.as<double>(), n, 1.0, M, n);
        gemm(H, rocblas_operation_none, rocblas_operation_transpose, n, n, 6, 1.0, T.as<double>(), n, R, n, 1.0, M, n);
        RT_HIP(hipStreamSynchronize(H.s));
    }
    n3_out = n;
    fbuf.ensure(sizeof(double) * (size_t)std::max(n, 1));
    RT_HIP(hipStreamSynchronize(H.s));
}

static void factor_coarsest(Hier& H) {
    factor_level(H, (int)H.lv.size() - 1, H.Ainv, H.rowslot, H.fbuf, H.n3);
    // the preconditioner's cycle stops one level early when that level is
    // small enough to solve exactly (its smoothing passes cost launches, not
    // bandwidth); the reference-semantics V-cycle keeps the whole hierarchy
    const int n = (int)H.lv.size();
    H.pc_levels = n;
    if (n >= 3 && 3 * (long long)H.lv[n - 2].n_vact <= H.pc_direct_dofs) {
        factor_level(H, n - 2, H.pcAinv, H.pcrowslot, H.pcfbuf, H.pcn3);
        H.pc_levels = n - 1;
    }
}

static void build_islands(Hier& H) {
    H.nisl = 0;
    if (!(H.isl_theta > 0)) return;
    Level& L0 = H.lv[0];
    const size_t N = (size_t)L0.g.N, mx = (size_t)H.isl_max;
    H.isl_lab.ensure(sizeof(int) * N);
    H.isl_anch.ensure(N);
    H.isl_cnt.ensure(sizeof(int) * N);
    H.isl_root.ensure(sizeof(int) * N);
    H.isl_node.ensure(sizeof(int) * N);
    H.isl_cent.ensure(sizeof(double) * 4 * mx);
    H.isl_E.ensure(sizeof(double) * 36 * mx);
    H.isl_Einv.ensure(sizeof(double) * 36 * mx);
    if (!H.isl_zr.p) H.isl_zr.alloc(sizeof(double) * 6 * mx);     // zeroed once; the correction leaves it zero
    H.isl_coef.ensure(sizeof(double) * 6 * mx);
    H.isl_flag.ensure(sizeof(int));
    H.isl_KE.ensure(sizeof(double) * 576);
    RT_HIP(hipMemcpyAsync(H.isl_KE.p, H.KE, sizeof(double) * 576, hipMemcpyHostToDevice, H.s));
    H.nisl = island_build(L0.g, H.rho.as<float>(), H.rhop.as<float>(), H.eflag.as<uint8_t>(), L0.vflag.as<uint8_t>(),
                          H.isl_KE.as<double>(), H.isl_theta, H.isl_min, H.isl_max, H.isl_lab.as<int>(),
                          H.isl_anch.as<uint8_t>(), H.isl_cnt.as<int>(), H.isl_root.as<int>(), H.isl_node.as<int>(),
                          H.isl_cent.as<double>(), H.isl_E.as<double>(), H.isl_Einv.as<double>(),
                          H.isl_flag.as<int>(), H.s);
}

static void update_stencil(Hier& H) {
    H.drop_graph();
    if (!H.material_set) throw std::logic_error("rt_set_material has not been called");
    Level& L0 = H.lv[0];
    launch_rhop(L0.g, H.rho.as<float>(), H.eflag.as<uint8_t>(), H.p, H.rhop.as<float>(), H.s);
    Level& L2 = H.lv[1];
    {
        // KEc[576 x nE] (row stride nE) = M^T R: column-major C(nE x 576) = R(nE x 64) M(64 x 576), where M is
        // stored [64][576], i.e. column-major 576 x 64 read transposed
        const int nE = L2.e[0] * L2.e[1] * L2.e[2];
        H.rh64.ensure(sizeof(double) * 64 * (size_t)nE);
        launch_children64(L0.g, L2.e, H.rhop.as<float>(), H.rh64.as<double>(), H.s);
        ensure_blas(H);
        gemm(H, rocblas_operation_none, rocblas_operation_transpose, nE, 576, 64, 1.0, H.rh64.as<double>(), nE,
             H.dM.as<double>(), 576, 0.0, L2.KEc.as<double>(), nE);
    }
    launch_galerkin4_fix(L0.g, L2.e, H.rhop.as<float>(), L0.vflag.as<uint8_t>(), H.eflag.as<uint8_t>(),
                         H.fixce.as<int>(), H.nfixce, L2.KEc.as<double>(), H.s);
    for (size_t li = 2; li < H.lv.size(); li++)
        launch_galerkin2(H.lv[li - 1].KEc.as<double>(), H.lv[li - 1].e, H.lv[li].KEc.as<double>(), H.lv[li].e, H.s);
    for (size_t li = 1; li < H.lv.size(); li++) {
        Level& L = H.lv[li];
        launch_assemble<double>(L.g, L.KEc.as<double>(), L.e, L.vflag.as<uint8_t>(), L.st64.as<double>(), H.s);
        launch_convert<double, float>(L.st64.as<double>(), L.st32.as<float>(), (size_t)243 * L.g.N, H.s);
        if (H.coarse_bits == 16) {
            L.st16.ensure(sizeof(uint16_t) * 243 * (size_t)L.g.N);
            launch_to_bf16(L.st64.as<double>(), L.st16.as<bf16>(), (size_t)243 * L.g.N, H.s);
        }
    }
    factor_coarsest(H);
    build_islands(H);
    H.stencil_ready = true;
}

// ---------------------------------------------------------------------------
// MG-PCG. TV: Krylov vectors, TC: V-cycle. Dots in FP64.
// ---------------------------------------------------------------------------
template <typename TV> struct KV;
template <> struct KV<double> {
    static V3<double>& x(Hier& H) { return H.x64; }
    static V3<double>& b(Hier& H) { return H.f64; }
    static V3<double>& r(Hier& H) { return H.r64; }
    static V3<double>& p(Hier& H) { return H.p64; }
    static V3<double>& q(Hier& H) { return H.q64; }
};
template <> struct KV<float> {
    static V3<float>& x(Hier& H) { return H.x32; }
    static V3<float>& b(Hier& H) { return H.b32; }
    static V3<float>& r(Hier& H) { return H.r32; }
    static V3<float>& p(Hier& H) { return H.p32; }
    static V3<float>& q(Hier& H) { return H.q32; }
};

template <typename TV>
static double apply_dot(Hier& H, V3<TV>& X, V3<TV>& Y) {
    Level& L0 = H.lv[0];
    int np = 0;
    launch_fine_apply_modal<TV>(L0.g, X.cv(), Y.v(), H.rhop.as<float>(), L0.vflag.as<uint8_t>(), H.partials.as<double>(), &np,
                          H.s);
    return sum_partials_host(H, np);
}

static double true_rel_residual(Hier& H) {
    Level& L0 = H.lv[0];
    if (!H.r64.ok()) H.r64.alloc(L0.g.N);
    if (!H.q64.ok()) H.q64.alloc(L0.g.N);
    const size_t n = (size_t)3 * L0.g.N;
    launch_fine_apply_modal<double>(L0.g, H.x64.cv(), H.q64.v(), H.rhop.as<float>(), L0.vflag.as<uint8_t>(), nullptr,
                              nullptr, H.s);
    launch_sub_conv<double, double>(H.f64.data(), H.q64.data(), H.r64.data(), (double*)nullptr, n,
                                    H.partials.as<double>(), 1024, H.s);
    const double rr = sum_partials_host(H, 1024);
    const double ff = dot<double, double>(H, H.f64.data(), H.f64.data(), n);
    return ff > 0 ? std::sqrt(rr / ff) : 0.0;
}

// PCG on caller-given vectors x (initial guess, updated) and b; stops when the
// recurrence residual reaches tol * |b|. Returns the iteration count; `rel`
// receives the final recurrence residual ratio.
// Subspace warm start: x64 = sum_i c_i h_i over the stored previous solutions
// h_i, with c the Galerkin (energy-norm optimal) coefficients, G c = b,
// G_ij = h_i . K h_j, b_i = h_i . f, solved through the eigen-decomposition of
// G with near-dependent directions dropped. With one stored solution this is
// the energy-optimal scaling applied in any case.
static void subspace_warm_start(Hier& H) {
    if (H.ws_m < 1 || H.ws_count < 1) return;
    const int k = H.ws_count;
    Level& L0 = H.lv[0];
    const size_t n = (size_t)3 * L0.g.N;
    if (!H.q64.ok()) H.q64.alloc(L0.g.N);
    std::vector<double> G((size_t)k * k), b(k);
    for (int j = 0; j < k; j++) {
        launch_fine_apply_modal<double>(L0.g, H.ws_hist[j].cv(), H.q64.v(), H.rhop.as<float>(),
                                        L0.vflag.as<uint8_t>(), nullptr, nullptr, H.s);
        for (int i = 0; i <= j; i++)
            G[(size_t)i * k + j] = G[(size_t)j * k + i] = dot<double, double>(H, H.ws_hist[i].data(), H.q64.data(), n);
        b[j] = dot<double, double>(H, H.ws_hist[j].data(), H.f64.data(), n);
    }
    std::vector<double> A = G, w, V;
    jacobi_eigen(k, A, w, V);
    double wmax = 0;
    for (double x : w) wmax = std::max(wmax, x);
    if (!(wmax > 0)) return;
    std::vector<double> c(k, 0.0);
    for (int e = 0; e < k; e++) {
        if (!(w[e] > 1e-10 * wmax)) continue;
        double vb = 0;
        for (int i = 0; i < k; i++) vb += V[(size_t)i * k + e] * b[i];
        for (int i = 0; i < k; i++) c[i] += V[(size_t)i * k + e] * vb / w[e];
    }
    std::vector<const double*> h(k);
    for (int i = 0; i < k; i++) h[i] = H.ws_hist[i].data();
    launch_lincomb(H.x64.data(), h.data(), c.data(), k, n, H.s);
}

// store the current solution in the ring of previous solutions
static void subspace_push(Hier& H) {
    if (H.ws_m < 1) return;
    Level& L0 = H.lv[0];
    if ((int)H.ws_hist.size() != H.ws_m) {
        H.ws_hist.clear();
        H.ws_hist.resize(H.ws_m);
        H.ws_count = H.ws_head = 0;
    }
    V3<double>& d = H.ws_hist[H.ws_head];
    if (!d.ok()) d.alloc(L0.g.N);
    RT_HIP(hipMemcpyAsync(d.data(), H.x64.data(), sizeof(double) * 3 * (size_t)L0.g.N, hipMemcpyDeviceToDevice, H.s));
    H.ws_head = (H.ws_head + 1) % H.ws_m;
    H.ws_count = std::min(H.ws_count + 1, H.ws_m);
}

// Learned warm start. The network predicts the displacement on the x4 level
// from the density and the restricted load; its prolongation u_nn joins the
// current start x (the previous solution, or the subspace combination) in a
// Galerkin combination x' = c0 x + c1 u_nn minimising the energy
// E(y) = y.Ky / 2 - f.y, i.e. the K-norm error. x' replaces x only when
// E(x') < E(a x), the energy of the best multiple of x alone (the PCG scales
// its start that way); otherwise the plain start stays (the plain MG-PCG
// fallback).
static void learned_warm_start(Hier& H) {
    if (!H.net) return;
    Level& L0 = H.lv[0];
    Level& C = H.lv[1];
    const size_t n = (size_t)3 * L0.g.N;
    LevelVecs<double>& cv = vecs<double>(C);
    restrict_to<double>(H, 1, H.f64.cv(), cv.F.v());
    RT_HIP(hipMemsetAsync(cv.U.data(), 0, sizeof(double) * 3 * (size_t)C.g.N, H.s));
    net_predict(*H.net, L0.g, L0.e, H.rho.as<float>(), C.g, cv.F.cv(), C.vflag.as<uint8_t>(), cv.U.v(), H.s);
    if (!H.nn_u.ok()) H.nn_u.alloc(L0.g.N);
    RT_HIP(hipMemsetAsync(H.nn_u.data(), 0, sizeof(double) * n, H.s));
    prolong_from<double>(H, 0, H.nn_u.v(), cv.U.cv(), true);
    if (!H.q64.ok()) H.q64.alloc(L0.g.N);
    const V3<double>* h[2] = {&H.x64, &H.nn_u};
    double G[4], b[2];
    for (int j = 0; j < 2; j++) {
        launch_fine_apply_modal<double>(L0.g, h[j]->cv(), H.q64.v(), H.rhop.as<float>(), L0.vflag.as<uint8_t>(),
                                        nullptr, nullptr, H.s);
        for (int i = 0; i <= j; i++) G[i * 2 + j] = G[j * 2 + i] = dot<double, double>(H, h[i]->data(), H.q64.data(), n);
        b[j] = dot<double, double>(H, h[j]->data(), H.f64.data(), n);
    }
    std::vector<double> A(G, G + 4), w, V;
    jacobi_eigen(2, A, w, V);
    const double wmax = std::max(w[0], w[1]);
    double c[2] = {0, 0};
    for (int e = 0; e < 2 && wmax > 0; e++) {
        if (!(w[e] > 1e-12 * wmax)) continue;
        const double vb = V[e] * b[0] + V[2 + e] * b[1];
        c[0] += V[e] * vb / w[e];
        c[1] += V[2 + e] * vb / w[e];
    }
    // the plain start is scaled by its own energy-optimal factor b0 / G0 before
    // the PCG begins, so the learned start must beat that: E(a x) = -b0^2 / (2 G0)
    const double e_plain = G[0] > 0 ? -0.5 * b[0] * b[0] / G[0] : 0.0;
    const double e_learn = 0.5 * (c[0] * c[0] * G[0] + 2 * c[0] * c[1] * G[1] + c[1] * c[1] * G[3]) - c[0] * b[0] -
                           c[1] * b[1];
    const bool take = std::isfinite(e_learn) && e_learn < e_plain;
    H.nn_info[0] = take ? 1.0 : 0.0;
    H.nn_info[1] = e_plain;
    H.nn_info[2] = e_learn;
    H.nn_info[3] = c[1];
    if (!take) return;
    const double* hp[2] = {H.x64.data(), H.nn_u.data()};
    launch_lincomb(H.q64.data(), hp, c, 2, n, H.s);
    RT_HIP(hipMemcpyAsync(H.x64.data(), H.q64.data(), sizeof(double) * n, hipMemcpyDeviceToDevice, H.s));
}

static void warm_start(Hier& H) {
    subspace_warm_start(H);
    learned_warm_start(H);
}

template <typename TV, typename TC>
static int pcg_core(Hier& H, V3<TV>& x, V3<TV>& b, double tol, int maxit, int pre, int post, double* rel_out) {
    Level& L0 = H.lv[0];
    const int N = L0.g.N;
    const size_t n = (size_t)3 * N;
    V3<TV>& r = KV<TV>::r(H);
    V3<TV>& p = KV<TV>::p(H);
    V3<TV>& q = KV<TV>::q(H);
    if (!r.ok()) r.alloc(N);
    if (!p.ok()) p.alloc(N);
    if (!q.ok()) q.alloc(N);
    LevelVecs<TC>& pc = pcvecs<TC>(H);
    const int np = 1024;
    // r = b - A x, with the initial guess scaled by the energy-optimal factor
    // (x.b)/(x.Ax): never worse than zero in the energy norm, and much better
    // than the raw previous solution after a density update.
    const double xAx = apply_dot<TV>(H, x, q);
    if (xAx > 0) {
        const double xb = dot<TV, TV>(H, x.data(), b.data(), n);
        const double alpha = xb > 0 ? xb / xAx : 0.0;
        launch_scale<TV>(x.data(), alpha, n, H.s);
        launch_scale<TV>(q.data(), alpha, n, H.s);
    }
    launch_sub_conv<TV, TC>(b.data(), q.data(), r.data(), pc.F.data(), n, H.partials.as<double>(), np, H.s);
    double rr = sum_partials_host(H, np);
    const double nb = std::sqrt(dot<TV, TV>(H, b.data(), b.data(), n));
    int it = 0;
    double rel = nb > 0 ? std::sqrt(rr) / nb : 0.0;
    if (nb > 0 && rel > tol) {
        precondition<TC>(H, pc, pre, post);
        double rz = dot<TV, TC>(H, r.data(), pc.U.data(), n);
        launch_convert<TC, TV>(pc.U.data(), p.data(), n, H.s);
        for (it = 1; it <= maxit; it++) {
            const double pq = apply_dot<TV>(H, p, q);
            if (!(pq > 0)) throw std::domain_error("PCG breakdown: p.Ap <= 0");
            const double alpha = rz / pq;
            launch_pcg_update_x_r<TV, TC>(x.data(), r.data(), p.data(), q.data(), alpha, n, H.partials.as<double>(), np,
                                          pc.F.data(), H.s);
            rr = sum_partials_host(H, np);
            rel = std::sqrt(rr) / nb;
            if (rel <= tol) break;
            precondition<TC>(H, pc, pre, post);
            const double rz_new = dot<TV, TC>(H, r.data(), pc.U.data(), n);
            const double beta = rz_new / rz;
            rz = rz_new;
            launch_pcg_update_p<TV, TC>(p.data(), pc.U.data(), beta, n, H.s);
        }
        if (it > maxit) it = maxit;
    }
    if (rel_out) *rel_out = rel;
    return it;
}

// warm: apply the warm starts first (false when continuing from the current x64)
template <typename TV, typename TC>
static void pcg(Hier& H, double tol, int maxit, int pre, int post, rt_solve_stats* st, bool warm = true) {
    if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    auto t0 = std::chrono::steady_clock::now();
    Level& L0 = H.lv[0];
    const int N = L0.g.N;
    const size_t n = (size_t)3 * N;
    V3<TV>& x = KV<TV>::x(H);
    V3<TV>& b = KV<TV>::b(H);
    if (!x.ok()) x.alloc(N);
    if (!b.ok()) b.alloc(N);
    if (warm) warm_start(H);
    if ((void*)x.data() != (void*)H.x64.data()) {
        launch_convert<double, TV>(H.x64.data(), x.data(), n, H.s);
        launch_convert<double, TV>(H.f64.data(), b.data(), n, H.s);
    }
    double rel = 0;
    const int it = pcg_core<TV, TC>(H, x, b, tol, maxit, pre, post, &rel);
    if ((void*)x.data() != (void*)H.x64.data()) launch_convert<TV, double>(x.data(), H.x64.data(), n, H.s);
    subspace_push(H);
    RT_HIP(hipStreamSynchronize(H.s));
    const double tr = true_rel_residual(H);
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        st->iterations = it;
        st->rel_res = tr;
        st->rel_res_pcg = rel;
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->converged = tr <= tol;
    }
}

// Iterative refinement: FP64 true residual r = f - K x outside, FP32 MG-PCG on
// K e = r inside (each inner solve reduces its residual by `inner`), x += e.
static void pcg_ir(Hier& H, double tol, int maxit, int pre, int post, double inner, rt_solve_stats* st) {
    if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    auto t0 = std::chrono::steady_clock::now();
    Level& L0 = H.lv[0];
    const int N = L0.g.N;
    const size_t n = (size_t)3 * N;
    if (!H.x32.ok()) H.x32.alloc(N);
    if (!H.b32.ok()) H.b32.alloc(N);
    int total = 0;
    warm_start(H);
    double rel = true_rel_residual(H), rel_in = 0;   // leaves r64 = f - K x64
    for (int outer = 0; outer < 50 && rel > tol && total < maxit; outer++) {
        launch_fill<float>(H.x32.data(), 0.f, n, H.s);
        launch_convert<double, float>(H.r64.data(), H.b32.data(), n, H.s);
        const double want = std::max(inner, 0.5 * tol / rel);
        total += pcg_core<float, float>(H, H.x32, H.b32, want, maxit - total, pre, post, &rel_in);
        launch_add_conv<double, float>(H.x64.data(), H.x32.data(), n, H.s);
        rel = true_rel_residual(H);
    }
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        st->iterations = total;
        st->rel_res = rel;
        st->rel_res_pcg = rel_in;
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->converged = rel <= tol;
    }
}

// PCG with FP32 Krylov vectors and FP32 operator, the solution accumulated in
// FP64, and reliable updates (residual replacement): whenever the FP32
// recurrence residual has dropped by 1/RT_REPLACE_DROP since the last
// replacement, and before stopping, r is recomputed as f - K x in FP64 and the
// recurrence continues from it. The loop stops only on a replaced (true)
// residual. Converges like the FP64 PCG while streaming half the bytes per
// iteration outside the V-cycle. Measured (511 x 256 x 256 cantilever / 384^3
// torus worst case): replacing every 10x drop 0.732 s / 5.80 s per SIMP
// iteration, every 100x 0.728 s / 5.56 s, every 1000x 0.899 s / 5.93 s (the
// FP32 recurrence drifts and PCG needs 99.6 instead of 79.6 iterations). At
// every 100x a 128 x 64 x 64 solve to 1e-8 stalled at a true residual of 3e-4,
// so replacements stay at every 10x.
#ifndef RT_REPLACE_DROP
#define RT_REPLACE_DROP 0.1
#endif
static void pcg_reliable(Hier& H, double tol, int maxit, int pre, int post, rt_solve_stats* st) {
    if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    auto t0 = std::chrono::steady_clock::now();
    Level& L0 = H.lv[0];
    const int N = L0.g.N;
    const size_t n = (size_t)3 * N;
    if (!H.r64.ok()) H.r64.alloc(N);
    if (!H.q64.ok()) H.q64.alloc(N);
    if (!H.p32.ok()) H.p32.alloc(N);
    if (!H.q32.ok()) H.q32.alloc(N);
    if (!H.dx32.ok()) H.dx32.alloc(N);
    LevelVecs<float>& pc = pcvecs<float>(H);   // pc.F holds the FP32 residual, pc.U the preconditioned one
    const int np = 1024;
    warm_start(H);
    // energy-optimal scaling of the start; a zero start (the worst-case
    // search's solves) needs no operator application: K x = 0
    if (dot<double, double>(H, H.x64.data(), H.x64.data(), n) > 0) {
        const double xAx = apply_dot<double>(H, H.x64, H.q64);
        if (xAx > 0) {
            const double xb = dot<double, double>(H, H.x64.data(), H.f64.data(), n);
            const double alpha = xb > 0 ? xb / xAx : 0.0;
            launch_scale<double>(H.x64.data(), alpha, n, H.s);
            launch_scale<double>(H.q64.data(), alpha, n, H.s);
        }
    } else {
        launch_fill<double>(H.q64.data(), 0.0, n, H.s);
    }
    launch_sub_conv<double, float>(H.f64.data(), H.q64.data(), H.r64.data(), pc.F.data(), n, H.partials.as<double>(),
                                   np, H.s);
    double rr = sum_partials_host(H, np);
    const double nb = std::sqrt(dot<double, double>(H, H.f64.data(), H.f64.data(), n));
    double rel = nb > 0 ? std::sqrt(rr) / nb : 0.0, rel_true = rel, rmax = std::sqrt(rr);
    int it = 0;
    bool diverged = false;
    bool have_true = false;   // rel_true is the true residual of the returned x64
    if (nb > 0 && rel > tol) {
        // The solution increment since the last residual replacement is kept
        // in FP32 (dx) and flushed into x64 at each replacement and at exit.
        // Each step's alpha p is added to dx by the next direction update,
        // which reads p anyway, or by the flush (apend: the step pending).
        const int ns = kStreamBlocks;
        double best_true = rel_true;
        H.cg_alpha.clear();
        H.cg_beta.clear();
        launch_fill<float>(H.dx32.data(), 0.f, n, H.s);
        precondition<float>(H, pc, pre, post);
        double rz = precond_rz(H, pc, post, n);
        launch_convert<float, float>(pc.U.data(), H.p32.data(), n, H.s);
        bool flushed = false;
        double apend = 0.0;
        for (it = 1; it <= maxit; it++) {
            const double pq = apply_dot<float>(H, H.p32, H.q32);
            if (!(pq > 0)) throw std::domain_error("PCG breakdown: p.Ap <= 0");
            const double alpha = rz / pq;
            H.cg_alpha.push_back(alpha);
            launch_pcg_ru_dxr(pc.F.data(), H.q32.data(), nullptr, nullptr, alpha, n, H.partials.as<double>(), ns,
                              H.s, vec_chunks(H));
            apend = alpha;
            flushed = false;
            const double rn = std::sqrt(sum_partials_host(H, ns));
            rel = rn / nb;
            if (rel <= tol || rn < RT_REPLACE_DROP * rmax) {
                launch_pcg_flush_x(H.x64.data(), H.dx32.data(), n, H.s, H.p32.data(), apend);
                apend = 0.0;
                flushed = true;
                rel_true = true_rel_residual(H);      // r64 = f - K x64
                if (rel_true <= tol) { rel = rel_true; have_true = true; break; }
                if (!(rel_true <= 10 * best_true)) { diverged = true; break; }
                best_true = std::min(best_true, rel_true);
                launch_convert<double, float>(H.r64.data(), pc.F.data(), n, H.s);
                rmax = rel_true * nb;
                rel = rel_true;
            }
            if (it == maxit) break;
            precondition<float>(H, pc, pre, post);
            const double rz_new = precond_rz(H, pc, post, n);
            const double beta = rz_new / rz;
            H.cg_beta.push_back(beta);
            rz = rz_new;
            if (apend != 0.0)
                launch_pcg_p_dx(H.p32.data(), pc.U.data(), beta, H.dx32.data(), apend, n, H.s, vec_chunks(H));
            else
                launch_pcg_update_p<float, float>(H.p32.data(), pc.U.data(), beta, n, H.s, vec_chunks(H));
            apend = 0.0;
        }
        if (!flushed) launch_pcg_flush_x(H.x64.data(), H.dx32.data(), n, H.s, H.p32.data(), apend);
        if (it > maxit) it = maxit;
    }
    if (diverged) {
        // The FP32 recurrence has lost the true residual (a replacement found it
        // 10x above the best one so far; seen with 0/1 densities at a 1e-9
        // stiffness contrast). Finish with FP64 Krylov vectors from the current
        // x, whose energy-optimal scaling is never worse than a zero start.
        rt_solve_stats s2{};
        pcg<double, float>(H, tol, std::max(0, maxit - it), pre, post, &s2, false);
        if (st) {
            st->iterations = it + s2.iterations;
            st->rel_res = s2.rel_res;
            st->rel_res_pcg = s2.rel_res_pcg;
            st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            st->converged = s2.converged;
        }
        return;
    }
    subspace_push(H);
    RT_HIP(hipStreamSynchronize(H.s));
    const double tr = have_true ? rel_true : true_rel_residual(H);
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        st->iterations = it;
        st->rel_res = tr;
        st->rel_res_pcg = rel;
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->converged = tr <= tol;
    }
}

// reference-semantics V-cycle on the finest displacement / force
template <typename T>
static double vcycle_on_solution(Hier& H, int pre, int post) {
    if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    Level& L0 = H.lv[0];
    const size_t n = (size_t)3 * L0.g.N;
    LevelVecs<T>& v = vecs<T>(L0);
    launch_convert<double, T>(H.x64.data(), v.U.data(), n, H.s);
    launch_convert<double, T>(H.f64.data(), v.F.data(), n, H.s);
    vcycle_core<T>(H, v.U, v.F, v.R, pre, post, false);
    residual<T>(H, 0, v.U.cv(), v.F.cv(), v.R.v());
    launch_convert<T, double>(v.U.data(), H.x64.data(), n, H.s);
    const double rr = dot<T, T>(H, v.R.data(), v.R.data(), n);
    const double ff = dot<T, T>(H, v.F.data(), v.F.data(), n);
    return std::sqrt(rr) / std::sqrt(ff);
}

// ---------------------------------------------------------------------------
// boundary conditions
// ---------------------------------------------------------------------------
static void set_support(Hier& H, const uint8_t* fixed_box) {
    H.drop_graph();
    Level& L0 = H.lv[0];
    const int nx = L0.g.n[0], ny = L0.g.n[1], nz = L0.g.n[2];
    const int ex = L0.e[0], ey = L0.e[1], ez = L0.e[2];
    for (size_t v = 0; v < L0.nn_box(); v++) {
        H.fixed_box[v] = (fixed_box[v] && L0.vact[v]) ? 1 : 0;
        uint8_t f = L0.vflag_box[v] & ~(NF_FIXED | NF_NEARFIX);
        if (H.fixed_box[v]) f |= NF_FIXED;
        L0.vflag_box[v] = f;
    }
    // mark the other 26 neighbours of every fixed node
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                if (!H.fixed_box[x + (size_t)nx * (y + (size_t)ny * z)]) continue;
                for (int nb = 0; nb < 27; nb++) {
                    if (nb == 13) continue;
                    const int jx = x + nb % 3 - 1, jy = y + nb / 3 % 3 - 1, jz = z + nb / 9 - 1;
                    if (jx < 0 || jy < 0 || jz < 0 || jx >= nx || jy >= ny || jz >= nz) continue;
                    const size_t j = jx + (size_t)nx * (jy + (size_t)ny * jz);
                    if (L0.vact[j]) L0.vflag_box[j] |= NF_NEARFIX;
                }
            }
    // elements touching a support, and the x4 coarse elements holding them
    Level& L2 = H.lv[1];
    std::vector<uint8_t> cmark(L2.ne_box(), 0);
    for (int z = 0; z < ez; z++)
        for (int y = 0; y < ey; y++)
            for (int x = 0; x < ex; x++) {
                const size_t e = x + (size_t)ex * (y + (size_t)ey * z);
                uint8_t f = H.eflag_box[e] & ~EF_TOUCHFIX;
                if (L0.eact[e]) {
                    bool t = false;
                    for (int k = 0; k < 8 && !t; k++)
                        t = H.fixed_box[(x + (k & 1)) + (size_t)nx * ((y + ((k >> 1) & 1)) + (size_t)ny * (z + (k >> 2)))];
                    if (t) {
                        f |= EF_TOUCHFIX;
                        cmark[(x >> 2) + (size_t)L2.e[0] * ((y >> 2) + (size_t)L2.e[1] * (z >> 2))] = 1;
                    }
                }
                H.eflag_box[e] = f;
            }
    std::vector<int> ce;
    for (size_t i = 0; i < cmark.size(); i++)
        if (cmark[i]) ce.push_back((int)i);
    H.nfixce = (int)ce.size();
    H.fixce.release();
    if (!ce.empty()) {
        H.fixce.alloc(sizeof(int) * ce.size());
        RT_HIP(hipMemcpy(H.fixce.p, ce.data(), sizeof(int) * ce.size(), hipMemcpyHostToDevice));
    }
    upload_box<uint8_t>(H, L0.g, 0, L0.vflag_box.data(), L0.vflag.as<uint8_t>());
    upload_box<uint8_t>(H, L0.g, 1, H.eflag_box.data(), H.eflag.as<uint8_t>());
    H.stencil_ready = false;
}

static void zero_fixed_force(Hier& H) {
    // F is zero at supports (load nodes exclude supports, Grid.cpp:1333-1345)
    Level& L0 = H.lv[0];
    const size_t nb = L0.nn_box();
    std::vector<double> box(3 * nb);
    download_vec<double>(H, L0.g, H.f64, box.data());
    bool any = false;
    for (size_t v = 0; v < nb; v++)
        if (H.fixed_box[v])
            for (int c = 0; c < 3; c++)
                if (box[c * nb + v] != 0) { box[c * nb + v] = 0; any = true; }
    if (any) upload_vec<double>(H, L0.g, box.data(), H.f64);
}

// active (lexicographic) <-> box
template <typename T>
static void expand_active(const std::vector<uint8_t>& mask, const T* act, T* box, int ncomp) {
    const size_t nb = mask.size();
    size_t na = 0;
    for (uint8_t m : mask) na += m;
    for (int c = 0; c < ncomp; c++) {
        size_t k = 0;
        for (size_t i = 0; i < nb; i++) box[c * nb + i] = mask[i] ? act[c * na + k++] : T(0);
    }
}
template <typename T>
static void compress_active(const std::vector<uint8_t>& mask, const T* box, T* act, int ncomp) {
    const size_t nb = mask.size();
    size_t na = 0;
    for (uint8_t m : mask) na += m;
    for (int c = 0; c < ncomp; c++) {
        size_t k = 0;
        for (size_t i = 0; i < nb; i++)
            if (mask[i]) act[c * na + k++] = box[c * nb + i];
    }
}

}  // namespace rt

// ===========================================================================
// C ABI
// ===========================================================================
using namespace rt;

struct rt_hier {
    Hier* h;
};

#define RT_TRY(...)                                                                  \
    try {                                                                            \
        __VA_ARGS__;                                                                 \
        return RT_OK;                                                                \
    } catch (const std::invalid_argument& e) {                                       \
        g_err = e.what();                                                            \
        return RT_EARG;                                                              \
    } catch (const std::domain_error& e) {                                           \
        g_err = e.what();                                                            \
        return RT_ENUM;                                                              \
    } catch (const std::logic_error& e) {                                            \
        g_err = e.what();                                                            \
        return RT_ESTATE;                                                            \
    } catch (const std::runtime_error& e) {                                          \
        g_err = e.what();                                                            \
        return std::string(e.what()).rfind("HIP", 0) == 0 ? RT_EHIP : RT_EINTERNAL;  \
    } catch (const std::exception& e) {                                              \
        g_err = e.what();                                                            \
        return RT_EINTERNAL;                                                         \
    }

int rt::status_of_current_exception() {
    try {
        throw;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return RT_EARG;
    } catch (const std::domain_error& e) {
        g_err = e.what();
        return RT_ENUM;
    } catch (const std::logic_error& e) {
        g_err = e.what();
        return RT_ESTATE;
    } catch (const std::runtime_error& e) {
        g_err = e.what();
        return std::string(e.what()).rfind("HIP", 0) == 0 ? RT_EHIP : RT_EINTERNAL;
    } catch (const std::exception& e) {
        g_err = e.what();
        return RT_EINTERNAL;
    } catch (...) {
        g_err = "unknown error";
        return RT_EINTERNAL;
    }
}

static Hier& H_(rt_hier* h) {
    if (!h || !h->h) throw std::invalid_argument("null handle");
    return *h->h;
}
static Hier& H_(const rt_hier* h) {
    if (!h || !h->h) throw std::invalid_argument("null handle");
    return *h->h;
}
static Level& LV(Hier& H, int level) {
    if (level < 0 || level >= (int)H.lv.size()) throw std::invalid_argument("level out of range");
    return H.lv[level];
}
static void need(const void* p) {
    if (!p) throw std::invalid_argument("null pointer argument");
}

extern "C" {

const char* rt_last_error(void) { return g_err.c_str(); }
const char* rt_version(void) { return "robtop_amd 0.1 (gfx950)"; }

int rt_create_from_voxels(int R, const uint32_t* solid_bits, size_t nwords, rt_hier** out) {
    RT_TRY({
        need(solid_bits); need(out);
        *out = nullptr;
        std::unique_ptr<Hier> h(build(R, solid_bits, nwords));
        setup_device(*h);
        *out = new rt_hier{h.release()};
    })
}

int rt_destroy(rt_hier* h) {
    RT_TRY({
        if (h) { delete h->h; delete h; }
    })
}

int rt_num_levels(const rt_hier* h) {
    try { return (int)H_(h).lv.size(); } catch (const std::exception& e) { g_err = e.what(); return RT_EARG; }
}
int rt_num_layers(const rt_hier* h) {
    try { return H_(h).nlayers; } catch (const std::exception& e) { g_err = e.what(); return RT_EARG; }
}
int rt_layer_active_elements(const rt_hier* h, int layer) {
    try {
        Hier& H = H_(h);
        if (layer < 0 || layer >= H.nlayers) throw std::invalid_argument("layer out of range");
        return H.layer_active[layer];
    } catch (const std::exception& e) { g_err = e.what(); return RT_EARG; }
}

int rt_level_info(const rt_hier* h, int level, int* info) {
    RT_TRY({
        need(info);
        Hier& H = H_(h);
        Level& L = LV(H, level);
        info[0] = L.layer; info[1] = L.scale;
        for (int k = 0; k < 3; k++) { info[2 + k] = L.o[k]; info[5 + k] = L.e[k]; info[8 + k] = L.g.n[k]; }
    })
}

int rt_level_masks(const rt_hier* h, int level, uint8_t* eact_box, uint8_t* vact_box) {
    RT_TRY({
        Hier& H = H_(h);
        Level& L = LV(H, level);
        if (eact_box) std::memcpy(eact_box, L.eact.data(), L.eact.size());
        if (vact_box) std::memcpy(vact_box, L.vact.data(), L.vact.size());
    })
}

// mark_surface_nodes_kernel (Grid.cu:2005-2039): a node is on the surface when
// along some axis no pair of solid elements faces each other across it
int rt_surface_nodes_box(const rt_hier* h, int level, uint8_t* surf_box) {
    RT_TRY({
        need(surf_box);
        Hier& H = H_(h);
        Level& L = LV(H, level);
        const int nx = L.g.n[0], ny = L.g.n[1], nz = L.g.n[2];
        const int ex = L.e[0], ey = L.e[1], ez = L.e[2];
        auto solid = [&](int x, int y, int z) {
            return x >= 0 && y >= 0 && z >= 0 && x < ex && y < ey && z < ez &&
                   L.eact[((size_t)z * ey + y) * ex + x] != 0;
        };
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    const size_t v = ((size_t)z * ny + y) * nx + x;
                    if (!L.vact[v]) { surf_box[v] = 0; continue; }
                    bool s[2][2][2];
                    for (int w = 0; w < 8; w++)
                        s[w & 1][(w >> 1) & 1][w >> 2] = solid(x + (w & 1) - 1, y + ((w >> 1) & 1) - 1, z + (w >> 2) - 1);
                    bool ax[3] = {false, false, false};
                    for (int i = 0; i < 2; i++)
                        for (int j = 0; j < 2; j++) {
                            ax[0] |= s[0][i][j] && s[1][i][j];
                            ax[1] |= s[i][0][j] && s[i][1][j];
                            ax[2] |= s[i][j][0] && s[i][j][1];
                        }
                    surf_box[v] = (!ax[0] || !ax[1] || !ax[2]) ? 1 : 0;
                }
    })
}

int rt_num_active(const rt_hier* h, int level, size_t* ne, size_t* nn) {
    RT_TRY({
        Hier& H = H_(h);
        Level& L = LV(H, level);
        if (ne) *ne = L.n_eact;
        if (nn) *nn = L.n_vact;
    })
}

int rt_set_material(rt_hier* h, double youngs, double poisson, float power_penalty) {
    RT_TRY({
        Hier& H = H_(h);
        H.drop_graph();
        if (!(youngs > 0) || !(poisson > -1 && poisson < 0.5)) throw std::invalid_argument("invalid material");
        H.E = youngs; H.nu = poisson; H.p = power_penalty;
        element_matrix(youngs, poisson, H.KE);
        set_ke_constants(H.KE, H.s);
        set_ke_galerkin(H.KE, H.s);
        set_ke_simp(H.KE, H.s);
        set_modal_constants(H.KE);
        std::vector<double> M;
        galerkin4_matrices(H.KE, M);
        H.dM.alloc(sizeof(double) * M.size());
        RT_HIP(hipMemcpy(H.dM.p, M.data(), sizeof(double) * M.size(), hipMemcpyHostToDevice));
        launch_rhop(H.lv[0].g, H.rho.as<float>(), H.eflag.as<uint8_t>(), H.p, H.rhop.as<float>(), H.s);
        RT_HIP(hipStreamSynchronize(H.s));
        H.material_set = true;
        H.stencil_ready = false;
    })
}

int rt_get_template_matrix(const rt_hier* h, double* ke) {
    RT_TRY({ need(ke); std::memcpy(ke, H_(h).KE, sizeof(double) * 576); })
}

int rt_set_support_box(rt_hier* h, const uint8_t* fixed_box) {
    RT_TRY({
        need(fixed_box);
        Hier& H = H_(h);
        set_support(H, fixed_box);
        zero_fixed_force(H);
    })
}

int rt_set_support_active(rt_hier* h, const uint8_t* fixed_active) {
    RT_TRY({
        need(fixed_active);
        Hier& H = H_(h);
        std::vector<uint8_t> box(H.lv[0].nn_box());
        expand_active<uint8_t>(H.lv[0].vact, fixed_active, box.data(), 1);
        set_support(H, box.data());
        zero_fixed_force(H);
    })
}

int rt_set_force_box(rt_hier* h, const double* f_box) {
    RT_TRY({
        need(f_box);
        Hier& H = H_(h);
        Level& L0 = H.lv[0];
        const size_t nb = L0.nn_box();
        std::vector<double> box(f_box, f_box + 3 * nb);
        for (size_t v = 0; v < nb; v++)
            if (!L0.vact[v] || H.fixed_box[v])
                for (int c = 0; c < 3; c++) box[c * nb + v] = 0;
        upload_vec<double>(H, L0.g, box.data(), H.f64);
    })
}

int rt_set_force_active(rt_hier* h, const double* f_active) {
    RT_TRY({
        need(f_active);
        Hier& H = H_(h);
        std::vector<double> box(3 * H.lv[0].nn_box());
        expand_active<double>(H.lv[0].vact, f_active, box.data(), 3);
        for (size_t v = 0; v < H.lv[0].nn_box(); v++)
            if (H.fixed_box[v])
                for (int c = 0; c < 3; c++) box[c * H.lv[0].nn_box() + v] = 0;
        upload_vec<double>(H, H.lv[0].g, box.data(), H.f64);
    })
}

int rt_get_force_box(const rt_hier* h, double* f_box) {
    RT_TRY({ need(f_box); Hier& H = H_(h); download_vec<double>(H, H.lv[0].g, H.f64, f_box); })
}

int rt_set_shell_box(rt_hier* h, const uint8_t* shell_box) {
    RT_TRY({
        need(shell_box);
        Hier& H = H_(h);
        for (size_t e = 0; e < H.eflag_box.size(); e++) {
            uint8_t f = H.eflag_box[e] & ~EF_SHELL;
            if (shell_box[e] && H.lv[0].eact[e]) f |= EF_SHELL;
            H.eflag_box[e] = f;
        }
        upload_box<uint8_t>(H, H.lv[0].g, 1, H.eflag_box.data(), H.eflag.as<uint8_t>());
    })
}

int rt_init_density(rt_hier* h, float rho) {
    RT_TRY({
        Hier& H = H_(h);
        std::vector<float> box(H.lv[0].ne_box());
        for (size_t e = 0; e < box.size(); e++) box[e] = H.lv[0].eact[e] ? rho : 0.f;
        upload_box<float>(H, H.lv[0].g, 1, box.data(), H.rho.as<float>());
        launch_rhop(H.lv[0].g, H.rho.as<float>(), H.eflag.as<uint8_t>(), H.p, H.rhop.as<float>(), H.s);
        RT_HIP(hipStreamSynchronize(H.s));
        H.stencil_ready = false;
    })
}

int rt_set_density_box(rt_hier* h, const float* rho_box) {
    RT_TRY({
        need(rho_box);
        Hier& H = H_(h);
        std::vector<float> box(rho_box, rho_box + H.lv[0].ne_box());
        for (size_t e = 0; e < box.size(); e++) if (!H.lv[0].eact[e]) box[e] = 0.f;
        upload_box<float>(H, H.lv[0].g, 1, box.data(), H.rho.as<float>());
        launch_rhop(H.lv[0].g, H.rho.as<float>(), H.eflag.as<uint8_t>(), H.p, H.rhop.as<float>(), H.s);
        RT_HIP(hipStreamSynchronize(H.s));
        H.stencil_ready = false;
    })
}

int rt_get_density_box(const rt_hier* h, float* rho_box) {
    RT_TRY({ need(rho_box); Hier& H = H_(h); download_box<float>(H, H.lv[0].g, 1, H.rho.as<float>(), rho_box); })
}

int rt_set_density_active(rt_hier* h, const float* rho_active) {
    RT_TRY({
        need(rho_active);
        Hier& H = H_(h);
        std::vector<float> box(H.lv[0].ne_box());
        expand_active<float>(H.lv[0].eact, rho_active, box.data(), 1);
        upload_box<float>(H, H.lv[0].g, 1, box.data(), H.rho.as<float>());
        launch_rhop(H.lv[0].g, H.rho.as<float>(), H.eflag.as<uint8_t>(), H.p, H.rhop.as<float>(), H.s);
        RT_HIP(hipStreamSynchronize(H.s));
        H.stencil_ready = false;
    })
}

int rt_get_density_active(const rt_hier* h, float* rho_active) {
    RT_TRY({
        need(rho_active);
        Hier& H = H_(h);
        std::vector<float> box(H.lv[0].ne_box());
        download_box<float>(H, H.lv[0].g, 1, H.rho.as<float>(), box.data());
        compress_active<float>(H.lv[0].eact, box.data(), rho_active, 1);
    })
}

int rt_get_rhop_box(const rt_hier* h, float* out) {
    RT_TRY({ need(out); Hier& H = H_(h); download_box<float>(H, H.lv[0].g, 1, H.rhop.as<float>(), out); })
}

int rt_update_stencil(rt_hier* h) { RT_TRY({ update_stencil(H_(h)); }) }

int rt_get_stencil_box(const rt_hier* h, int level, double* st_box) {
    RT_TRY({
        need(st_box);
        Hier& H = H_(h);
        if (level < 1) throw std::invalid_argument("the finest level has no stored stencil");
        Level& L = LV(H, level);
        const size_t nb = L.nn_box();
        for (int k = 0; k < 243; k++)
            download_box<double>(H, L.g, 0, L.st64.as<double>() + (size_t)k * L.g.N, st_box + k * nb);
    })
}

int rt_set_graphs(rt_hier* h, int on) {
    RT_TRY({
        Hier& H = H_(h);
        H.drop_graph();
        H.use_graphs = on != 0;
    })
}

int rt_set_island_deflation(rt_hier* h, float theta, int min_size) {
    RT_TRY({
        Hier& H = H_(h);
        if (min_size < 1) throw std::invalid_argument("island min_size must be positive");
        H.drop_graph();
        H.isl_theta = theta;
        H.isl_min = min_size;
        if (H.stencil_ready) build_islands(H);
    })
}

int rt_num_islands(const rt_hier* h, int* n) {
    RT_TRY({ need(n); *n = H_(h).nisl; })
}

int rt_set_warm_start(rt_hier* h, int m) {
    RT_TRY({
        Hier& H = H_(h);
        if (m < 0 || m > 8) throw std::invalid_argument("warm-start subspace size must be 0..8");
        H.ws_m = m;
        H.ws_hist.clear();
        H.ws_count = H.ws_head = 0;
    })
}

static void need_x4_level(const Hier& H) {
    if (H.lv.size() < 2 || H.lv[1].scale != 4)
        throw std::invalid_argument("the learned warm start needs an x4 coarse level");
}

// number of eigenvalues of the symmetric tridiagonal (d, e) below x (Sturm count)
static int sturm_count(const std::vector<double>& d, const std::vector<double>& e, double x) {
    int c = 0;
    double q = 1.0;
    for (size_t i = 0; i < d.size(); i++) {
        q = d[i] - x - (i ? e[i - 1] * e[i - 1] / q : 0.0);
        if (q == 0.0) q = -1e-300;
        if (q < 0) c++;
    }
    return c;
}

int rt_cg_coefficients(const rt_hier* h, int nmax, double* alpha, double* beta, int* n_out) {
    RT_TRY({
        need(n_out);
        const Hier& H = H_(h);
        *n_out = (int)H.cg_alpha.size();
        for (int i = 0; i < nmax && i < (int)H.cg_alpha.size(); i++) {
            if (alpha) alpha[i] = H.cg_alpha[i];
            if (beta) beta[i] = i < (int)H.cg_beta.size() ? H.cg_beta[i] : 0.0;
        }
    })
}

int rt_cg_ritz_values(const rt_hier* h, int k, double* smallest, double* largest, int* n_out) {
    RT_TRY({
        need(n_out);
        const Hier& H = H_(h);
        const int n = (int)std::min(H.cg_alpha.size(), H.cg_beta.size() + 1);
        *n_out = n;
        if (n == 0 || k < 1) return RT_OK;
        // T_jj = 1/alpha_j + beta_{j-1}/alpha_{j-1}, T_{j,j+1} = sqrt(beta_j)/alpha_j
        std::vector<double> d(n), e(n > 1 ? n - 1 : 0);
        for (int j = 0; j < n; j++) {
            d[j] = 1.0 / H.cg_alpha[j] + (j ? H.cg_beta[j - 1] / H.cg_alpha[j - 1] : 0.0);
            if (j + 1 < n) e[j] = std::sqrt(std::max(0.0, H.cg_beta[j])) / H.cg_alpha[j];
        }
        double lo = 0, hi = 0;
        for (int j = 0; j < n; j++) {
            const double r = (j ? std::fabs(e[j - 1]) : 0.0) + (j + 1 < n ? std::fabs(e[j]) : 0.0);
            lo = std::min(lo, d[j] - r);
            hi = std::max(hi, d[j] + r);
        }
        // i-th smallest eigenvalue by bisection on the Sturm count (Gershgorin bracket)
        auto kth = [&](int i) {
            double a = lo, b = hi;
            for (int it = 0; it < 200 && b - a > 1e-14 * std::max(1.0, std::fabs(b)); it++) {
                const double m = 0.5 * (a + b);
                if (sturm_count(d, e, m) > i) b = m; else a = m;
            }
            return 0.5 * (a + b);
        };
        for (int i = 0; i < k; i++) {
            if (smallest) smallest[i] = i < n ? kth(i) : 0.0;
            if (largest) largest[i] = i < n ? kth(n - 1 - i) : 0.0;
        }
    })
}

int rt_set_learned_warm_start(rt_hier* h, rt_net* net) {
    RT_TRY({
        Hier& H = H_(h);
        if (net && !net->n) throw std::invalid_argument("null rt_net handle");
        if (net) need_x4_level(H);
        H.net = net ? net->n : nullptr;
        if (!H.net) H.nn_u.b.release();
        for (double& v : H.nn_info) v = 0;
    })
}

int rt_learned_warm_start_info(const rt_hier* h, double* info) {
    RT_TRY({
        need(info);
        const Hier& H = H_(h);
        for (int k = 0; k < 4; k++) info[k] = H.nn_info[k];
    })
}

static double net_sample(rt_net* net, rt_hier* h, bool store) {
    Hier& H = H_(h);
    if (!net || !net->n) throw std::invalid_argument("null rt_net handle");
    need_x4_level(H);
    Level& L0 = H.lv[0];
    Level& C = H.lv[1];
    LevelVecs<double>& cv = vecs<double>(C);
    restrict_to<double>(H, 1, H.f64.cv(), cv.F.v());
    if (store) {
        net_add_sample(*net->n, L0.g, L0.e, H.rho.as<float>(), C.g, cv.F.cv(), H.x64.cv(), H.s);
        return 0.0;
    }
    return net_evaluate(*net->n, L0.g, L0.e, H.rho.as<float>(), C.g, cv.F.cv(), H.x64.cv(), H.s);
}

int rt_net_add_sample(rt_net* net, rt_hier* h) { RT_TRY({ net_sample(net, h, true); }) }

int rt_net_predict_coarse(rt_net* net, rt_hier* h) {
    RT_TRY({
        Hier& H = H_(h);
        if (!net || !net->n) throw std::invalid_argument("null rt_net handle");
        need_x4_level(H);
        if (H.vprec != 64)
            throw std::logic_error("rt_net_predict_coarse works on the FP64 level vectors (rt_set_vcycle_precision 64)");
        Level& L0 = H.lv[0];
        Level& C = H.lv[1];
        LevelVecs<double>& cv = vecs<double>(C);
        launch_fill<double>(cv.U.data(), 0.0, (size_t)3 * C.g.N, H.s);
        net_predict(*net->n, L0.g, L0.e, H.rho.as<float>(), C.g, cv.F.cv(), C.vflag.as<uint8_t>(), cv.U.v(), H.s);
        RT_HIP(hipStreamSynchronize(H.s));
    })
}

int rt_net_evaluate(rt_net* net, rt_hier* h, double* loss) {
    RT_TRY({
        need(loss);
        *loss = net_sample(net, h, false);
    })
}

int rt_set_fine_smoother(rt_hier* h, int kind) {
    RT_TRY({
        Hier& H = H_(h);
        if (kind != 0 && kind != 1) throw std::invalid_argument("fine smoother kind must be 0 or 1");
        H.drop_graph();
        H.fine_smoother = kind;
    })
}

int rt_set_coarse_stencil_bits(rt_hier* h, int bits) {
    RT_TRY({
        Hier& H = H_(h);
        if (bits != 16 && bits != 32) throw std::invalid_argument("coarse stencil bits must be 16 or 32");
        if (bits == H.coarse_bits) return RT_OK;
        H.drop_graph();
        H.coarse_bits = bits;
        if (bits == 16 && H.stencil_ready)
            for (size_t li = 1; li < H.lv.size(); li++) {
                Level& L = H.lv[li];
                L.st16.ensure(sizeof(uint16_t) * 243 * (size_t)L.g.N);
                launch_to_bf16(L.st64.as<double>(), L.st16.as<bf16>(), (size_t)243 * L.g.N, H.s);
            }
        RT_HIP(hipStreamSynchronize(H.s));
    })
}

int rt_set_preconditioner(rt_hier* h, int coarse_sweep_mult, int gamma) {
    RT_TRY({
        if (coarse_sweep_mult < 1 || coarse_sweep_mult > 16 || gamma < 1 || gamma > 2)
            throw std::invalid_argument("need 1 <= coarse_sweep_mult <= 16 and gamma in {1, 2}");
        Hier& H = H_(h);
        H.pc_mult = coarse_sweep_mult;
        H.pc_gamma = gamma;
    })
}

int rt_set_level_sweeps(rt_hier* h, int level, int mult) {
    RT_TRY({
        Hier& H = H_(h);
        if (level < 1 || level >= (int)H.lv.size()) throw std::out_of_range("level must be a coarse level");
        if (mult < 0 || mult > 16) throw std::invalid_argument("mult must be 0..16");
        if ((int)H.pc_lv_mult.size() < (int)H.lv.size()) H.pc_lv_mult.resize(H.lv.size(), 0);
        H.pc_lv_mult[level] = mult;
        H.drop_graph();
    })
}

int rt_set_coarse_direct_dofs(rt_hier* h, int max_dofs) {
    RT_TRY({
        if (max_dofs < 0 || max_dofs > 30000) throw std::invalid_argument("max_dofs must be 0..30000");
        Hier& H = H_(h);
        H.pc_direct_dofs = max_dofs;
        H.drop_graph();
        if (H.stencil_ready) factor_coarsest(H);
    })
}

int rt_preconditioner_levels(const rt_hier* h, int* levels) {
    RT_TRY({
        if (!levels) throw std::invalid_argument("null output");
        *levels = H_(h).stencil_ready ? H_(h).pc_levels : 0;
    })
}

int rt_set_vcycle_precision(rt_hier* h, int bits) {
    RT_TRY({
        if (bits != 32 && bits != 64) throw std::invalid_argument("precision must be 32 or 64");
        for (const Level& L : H_(h).lv) check_load_range(L, bits / 8);
        H_(h).vprec = bits;
    })
}

}  // extern "C"
template <typename T>
static V3<T>& level_vec(Hier& H, int level, int which) {
    if (which < 0 || which > 2) throw std::invalid_argument("which must be RT_VEC_U/F/R");
    LevelVecs<T>& v = vecs<T>(LV(H, level));
    return which == 0 ? v.U : which == 1 ? v.F : v.R;
}
extern "C" {

int rt_set_level_vec(rt_hier* h, int level, int which, const double* v_box) {
    RT_TRY({
        need(v_box);
        Hier& H = H_(h);
        if (H.vprec == 64) upload_vec<double>(H, LV(H, level).g, v_box, level_vec<double>(H, level, which));
        else upload_vec<float>(H, LV(H, level).g, v_box, level_vec<float>(H, level, which));
    })
}

int rt_get_level_vec(const rt_hier* h, int level, int which, double* v_box) {
    RT_TRY({
        need(v_box);
        Hier& H = H_(h);
        if (H.vprec == 64) download_vec<double>(H, LV(H, level).g, level_vec<double>(H, level, which), v_box);
        else download_vec<float>(H, LV(H, level).g, level_vec<float>(H, level, which), v_box);
    })
}

}  // extern "C"
template <typename T>
static void do_gs(Hier& H, int level, int n, int reverse) {
    LevelVecs<T>& v = vecs<T>(LV(H, level));
    if (level > 0 && !H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    gs<T>(H, level, v.U.v(), v.F.cv(), n, reverse != 0);
    RT_HIP(hipStreamSynchronize(H.s));
}
extern "C" {

int rt_gs_relax(rt_hier* h, int level, int nsweeps, int reverse) {
    RT_TRY({
        Hier& H = H_(h);
        if (H.vprec == 64) do_gs<double>(H, level, nsweeps, reverse); else do_gs<float>(H, level, nsweeps, reverse);
    })
}

}  // extern "C"
template <typename T>
static void do_res(Hier& H, int level) {
    LevelVecs<T>& v = vecs<T>(LV(H, level));
    if (level > 0 && !H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    residual<T>(H, level, v.U.cv(), v.F.cv(), v.R.v());
    RT_HIP(hipStreamSynchronize(H.s));
}
extern "C" {

int rt_update_residual(rt_hier* h, int level) {
    RT_TRY({ Hier& H = H_(h); if (H.vprec == 64) do_res<double>(H, level); else do_res<float>(H, level); })
}

}  // extern "C"
template <typename T>
static void do_restrict(Hier& H, int level) {
    if (level < 1) throw std::invalid_argument("restrict_residual targets a coarse level");
    LevelVecs<T>& f = vecs<T>(LV(H, level - 1));
    LevelVecs<T>& c = vecs<T>(LV(H, level));
    restrict_to<T>(H, level, f.R.cv(), c.F.v());
    RT_HIP(hipStreamSynchronize(H.s));
}
extern "C" {

int rt_restrict_residual(rt_hier* h, int level) {
    RT_TRY({ Hier& H = H_(h); if (H.vprec == 64) do_restrict<double>(H, level); else do_restrict<float>(H, level); })
}

}  // extern "C"
template <typename T>
static void do_prolong(Hier& H, int level) {
    if (level < 0 || level + 1 >= (int)H.lv.size()) throw std::invalid_argument("prolongate_correction needs a coarser level");
    LevelVecs<T>& f = vecs<T>(LV(H, level));
    LevelVecs<T>& c = vecs<T>(LV(H, level + 1));
    prolong_from<T>(H, level, f.U.v(), c.U.cv());
    RT_HIP(hipStreamSynchronize(H.s));
}
extern "C" {

int rt_prolongate_correction(rt_hier* h, int level) {
    RT_TRY({ Hier& H = H_(h); if (H.vprec == 64) do_prolong<double>(H, level); else do_prolong<float>(H, level); })
}

int rt_reset_displacement(rt_hier* h, int level) {
    RT_TRY({
        Hier& H = H_(h);
        Level& L = LV(H, level);
        if (H.vprec == 64) launch_fill<double>(vecs<double>(L).U.data(), 0.0, (size_t)3 * L.g.N, H.s);
        else launch_fill<float>(vecs<float>(L).U.data(), 0.f, (size_t)3 * L.g.N, H.s);
        RT_HIP(hipStreamSynchronize(H.s));
    })
}

int rt_solve_coarsest(rt_hier* h) {
    RT_TRY({
        Hier& H = H_(h);
        Level& L = H.lv.back();
        if (H.vprec == 64) coarsest<double>(H, vecs<double>(L).F.cv(), vecs<double>(L).U.v());
        else coarsest<float>(H, vecs<float>(L).F.cv(), vecs<float>(L).U.v());
        RT_HIP(hipStreamSynchronize(H.s));
    })
}

int rt_relative_residual(rt_hier* h, double* rel) {
    RT_TRY({
        need(rel);
        Hier& H = H_(h);
        Level& L = H.lv[0];
        const size_t n = (size_t)3 * L.g.N;
        if (H.vprec == 64) {
            LevelVecs<double>& v = vecs<double>(L);
            *rel = std::sqrt(dot<double, double>(H, v.R.data(), v.R.data(), n) / dot<double, double>(H, v.F.data(), v.F.data(), n));
        } else {
            LevelVecs<float>& v = vecs<float>(L);
            *rel = std::sqrt(dot<float, float>(H, v.R.data(), v.R.data(), n) / dot<float, float>(H, v.F.data(), v.F.data(), n));
        }
    })
}

int rt_vcycle(rt_hier* h, int pre, int post, double* rel_res) {
    RT_TRY({
        Hier& H = H_(h);
        double r = H.vprec == 64 ? vcycle_on_solution<double>(H, pre, post) : vcycle_on_solution<float>(H, pre, post);
        if (rel_res) *rel_res = r;
    })
}

}  // extern "C"
static void solve_mode(Hier& H, double tol, int maxit, int mode, int pre, int post, rt_solve_stats* stats) {
    if (pre < 0 || post < 0 || maxit < 0) throw std::invalid_argument("negative sweep or iteration count");
    if (mode == RT_SOLVE_FP64) pcg<double, double>(H, tol, maxit, pre, post, stats);
    else if (mode == RT_SOLVE_FP32) pcg<float, float>(H, tol, maxit, pre, post, stats);
    else if (mode == RT_SOLVE_MIXED) pcg<double, float>(H, tol, maxit, pre, post, stats);
    else if (mode == RT_SOLVE_FP32_IR) pcg_ir(H, tol, maxit, pre, post, 1e-4, stats);
    else if (mode == RT_SOLVE_RELIABLE) pcg_reliable(H, tol, maxit, pre, post, stats);
    else throw std::invalid_argument("unknown solver mode");
}
extern "C" {

int rt_solve(rt_hier* h, double tol, int maxit, int mode, int pre, int post, rt_solve_stats* stats) {
    RT_TRY({ solve_mode(H_(h), tol, maxit, mode, pre, post, stats); })
}

int rt_set_displacement_box(rt_hier* h, const double* u_box) {
    RT_TRY({
        need(u_box);
        Hier& H = H_(h);
        Level& L0 = H.lv[0];
        const size_t nb = L0.nn_box();
        std::vector<double> box(u_box, u_box + 3 * nb);
        for (size_t v = 0; v < nb; v++)
            if (!L0.vact[v])
                for (int c = 0; c < 3; c++) box[c * nb + v] = 0;
        upload_vec<double>(H, L0.g, box.data(), H.x64);
    })
}

int rt_get_displacement_box(const rt_hier* h, double* u_box) {
    RT_TRY({ need(u_box); Hier& H = H_(h); download_vec<double>(H, H.lv[0].g, H.x64, u_box); })
}

int rt_get_displacement_active(const rt_hier* h, double* u_active) {
    RT_TRY({
        need(u_active);
        Hier& H = H_(h);
        std::vector<double> box(3 * H.lv[0].nn_box());
        download_vec<double>(H, H.lv[0].g, H.x64, box.data());
        compress_active<double>(H.lv[0].vact, box.data(), u_active, 3);
    })
}

// Grid::applyK (Grid.cu:2519-2532, kernel :2439-2517): f = K(rho) u with support
// neighbours taken as zero; a support node outputs its displacement (the
// reference writes the grid's U there, which is u at every call site,
// optimization.cpp:390-531)
int rt_apply_k(rt_hier* h, const double* u_box, double* f_box) {
    RT_TRY({
        need(u_box);
        need(f_box);
        Hier& H = H_(h);
        Level& L0 = H.lv[0];
        const size_t nb = L0.nn_box();
        std::vector<double> box(u_box, u_box + 3 * nb);
        for (size_t v = 0; v < nb; v++)
            if (!L0.vact[v])
                for (int c = 0; c < 3; c++) box[c * nb + v] = 0;
        V3<double> X, Y;
        X.alloc(L0.g.N);
        Y.alloc(L0.g.N);
        upload_vec<double>(H, L0.g, box.data(), X);
        launch_fine_apply_modal<double>(L0.g, X.cv(), Y.v(), H.rhop.as<float>(), L0.vflag.as<uint8_t>(), nullptr,
                                        nullptr, H.s);
        download_vec<double>(H, L0.g, Y, f_box);
        for (size_t v = 0; v < nb; v++)
            if (L0.vact[v] && H.fixed_box[v])
                for (int c = 0; c < 3; c++) f_box[c * nb + v] = box[c * nb + v];
    })
}

int rt_compliance(rt_hier* h, double* c) {
    RT_TRY({
        need(c);
        Hier& H = H_(h);
        *c = dot<double, double>(H, H.f64.data(), H.x64.data(), (size_t)3 * H.lv[0].g.N);
    })
}

}  // extern "C"

// ---------------------------------------------------------------------------
// worst-case load search (modified power method)
// ---------------------------------------------------------------------------
// rigid motions of the load nodes at p in the compact layout [k][c * nl + i]:
// with normals (nrm[c][i]) each node's 3-vector reduced to its normal part,
// then modified Gram-Schmidt with dropping as rigid_basis
static void rigid_basis_load(const std::vector<double>& p, const std::vector<double>* nrm, std::vector<double>& out,
                             int& kept) {
    std::vector<double> R;
    const size_t nl = p.size() / 3, m = 3 * nl;
    // interleaved basis first (rigid_basis orthonormalises), so build the raw
    // columns here in the compact layout and orthonormalise the same way
    R.assign(m * 6, 0.0);
    for (size_t i = 0; i < nl; i++) {
        const double x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
        const double ph[3][3] = {{0, -z, y}, {z, 0, -x}, {-y, x, 0}};
        double col[6][3];
        for (int c = 0; c < 3; c++) {
            for (int k = 0; k < 6; k++) col[k][c] = 0;
            col[c][c] = 1.0;
            for (int k = 0; k < 3; k++) col[3 + k][c] = ph[c][k];
        }
        if (nrm)
            for (int k = 0; k < 6; k++) {
                const double d = col[k][0] * nrm[0][i] + col[k][1] * nrm[1][i] + col[k][2] * nrm[2][i];
                for (int c = 0; c < 3; c++) col[k][c] = d * nrm[c][i];
            }
        for (int k = 0; k < 6; k++)
            for (int c = 0; c < 3; c++) R[k * m + c * nl + i] = col[k][c];
    }
    bool keep[6];
    for (int i = 0; i < 6; i++) {
        double* ci = R.data() + i * m;
        double old = 0;
        for (size_t r = 0; r < m; r++) old += ci[r] * ci[r];
        old = std::sqrt(old);
        for (int j = 0; j < i; j++) {
            if (!keep[j]) continue;
            const double* cj = R.data() + j * m;
            double w = 0;
            for (size_t r = 0; r < m; r++) w += ci[r] * cj[r];
            for (size_t r = 0; r < m; r++) ci[r] -= w * cj[r];
        }
        double res = 0;
        for (size_t r = 0; r < m; r++) res += ci[r] * ci[r];
        res = std::sqrt(res);
        keep[i] = !(res < 1e-6 * old || old < 1e-6);
        if (keep[i])
            for (size_t r = 0; r < m; r++) ci[r] /= res;
    }
    out.clear();
    kept = 0;
    for (int i = 0; i < 6; i++)
        if (keep[i]) {
            out.insert(out.end(), R.begin() + i * m, R.begin() + (i + 1) * m);
            kept++;
        }
}

// setLoadNodes (projection.cpp:129-269), with-support modes: load nodes are
// existing, non-support nodes; normals (if given) are normalised per node
static void set_load_nodes(Hier& H, const uint8_t* load_box, const double* nrm_box) {
    Level& L0 = H.lv[0];
    const size_t nb = L0.nn_box();
    const int nx = L0.g.n[0], ny = L0.g.n[1];
    std::vector<int> slots;
    std::vector<double> nrm[3];
    for (size_t v = 0; v < nb; v++) {
        if (!load_box[v] || !L0.vact[v] || H.fixed_box[v]) continue;
        const int x = (int)(v % nx), y = (int)(v / nx % ny), z = (int)(v / ((size_t)nx * ny));
        if (nrm_box) {
            const double a = nrm_box[v], b = nrm_box[nb + v], c = nrm_box[2 * nb + v];
            const double l = std::sqrt(a * a + b * b + c * c);
            if (!(l > 0)) throw std::invalid_argument("zero normal at a load node");
            nrm[0].push_back(a / l); nrm[1].push_back(b / l); nrm[2].push_back(c / l);
        }
        slots.push_back(oct_p(L0.g, x, y, z));
    }
    H.nload = (int)slots.size();
    H.load_normals = nrm_box != nullptr;
    H.load_slots.release(); H.load_nrm.release(); H.pm_fs.release(); H.pm_old.release();
    if (!H.nload) return;
    const size_t nl = (size_t)H.nload;
    H.load_slots.alloc(sizeof(int) * nl);
    RT_HIP(hipMemcpy(H.load_slots.p, slots.data(), sizeof(int) * nl, hipMemcpyHostToDevice));
    if (H.load_normals) {
        H.load_nrm.alloc(sizeof(double) * 3 * nl);
        for (int c = 0; c < 3; c++)
            RT_HIP(hipMemcpy(H.load_nrm.as<double>() + c * nl, nrm[c].data(), sizeof(double) * nl,
                             hipMemcpyHostToDevice));
    }
    H.pm_fs.alloc(sizeof(double) * 3 * nl);
    H.pm_old.alloc(sizeof(double) * 3 * nl);
    // the admissible rigid motions of the load nodes for the no-support
    // modes (projection.cpp:208-257): with normals only their normal part
    // (Rd -= Nd^T Nd Rd), orthonormalised; stored compact [k][c * nl + i]
    std::vector<double> pos;
    pos.reserve(3 * nl);
    for (size_t v = 0; v < nb; v++) {
        if (!load_box[v] || !L0.vact[v] || H.fixed_box[v]) continue;
        pos.push_back((double)(v % nx)); pos.push_back((double)(v / nx % ny)); pos.push_back((double)(v / ((size_t)nx * ny)));
    }
    std::vector<double> R;
    rigid_basis_load(pos, H.load_normals ? nrm : nullptr, R, H.nrd);
    H.load_rd.release();
    if (H.nrd) {
        H.load_rd.alloc(sizeof(double) * R.size());
        RT_HIP(hipMemcpy(H.load_rd.p, R.data(), sizeof(double) * R.size(), hipMemcpyHostToDevice));
    }
}

static const int kPmParts = 256;

// fs = P x (compact); returns |fs|^2
static double pm_gather(Hier& H, const V3<double>& X) {
    launch_pm_gather(H.nload, H.load_slots.as<int>(), X.cv(), H.load_normals ? H.load_nrm.as<double>() : nullptr,
                     H.pm_fs.as<double>(), H.partials.as<double>(), kPmParts, H.s);
    return sum_partials_host(H, kPmParts);
}
// f64 <- scale * fs on the load nodes; returns (|new - old|, |old|)
static std::pair<double, double> pm_commit(Hier& H, double scale) {
    launch_pm_commit(H.nload, H.load_slots.as<int>(), H.pm_fs.as<double>(), scale, H.f64.v(), H.pm_old.as<double>(),
                     H.partials.as<double>(), kPmParts, H.s);
    const double d = sum_partials_at(H, H.partials.as<double>(), kPmParts);
    const double o = sum_partials_at(H, H.partials.as<double>() + kPmParts, kPmParts);
    return {std::sqrt(d), std::sqrt(o)};
}

// no-support modes (projection.cpp:322-335): the gathered load fs also loses
// its rigid part, fs -= Rd Rd^T fs; returns |fs|^2 (n2 unchanged with supports)
static double pm_balance(Hier& H, double n2) {
    if (has_supports(H) || !H.nrd) return n2;
    const int m = 3 * H.nload;
    double* w = H.scal.as<double>() + 8;   // the weights Rd^T fs, on the device
    const double one = 1.0, zero = 0.0, mone = -1.0;
    if (rocblas_dgemv(H.blas, rocblas_operation_transpose, m, H.nrd, &one, H.load_rd.as<double>(), m,
                      H.pm_fs.as<double>(), 1, &zero, w, 1) != rocblas_status_success ||
        rocblas_dgemv(H.blas, rocblas_operation_none, m, H.nrd, &mone, H.load_rd.as<double>(), m, w, 1, &one,
                      H.pm_fs.as<double>(), 1) != rocblas_status_success)
        throw std::runtime_error("rocblas_dgemv failed");
    return dot<double, double>(H, H.pm_fs.as<double>(), H.pm_fs.as<double>(), (size_t)m);
}

// forceProject (projection.cu:165-231) of the finest force, in place
static void force_project(Hier& H, bool scale_to_unit) {
    if (!H.nload) throw std::logic_error("no load nodes set");
    ensure_blas(H);
    const double n2 = pm_balance(H, pm_gather(H, H.f64));
    launch_fill<double>(H.f64.data(), 0.0, (size_t)3 * H.lv[0].g.N, H.s);
    double sc = 1.0;
    if (scale_to_unit) {
        if (!(n2 > 0)) throw std::domain_error("projected force is zero");
        sc = 1.0 / std::sqrt(n2);
    }
    pm_commit(H, sc);
}

// displacementProject (projection.cpp:374-394): x64 -= (x64 . R_k) R_k for
// the orthonormal rigid motions R_k of the finest nodes, one after the other
static void displacement_project(Hier& H) {
    Level& L0 = H.lv[0];
    const size_t n = (size_t)3 * L0.g.N;
    if (H.rigid_f.empty()) {
        const int nx = L0.g.n[0], ny = L0.g.n[1];
        const size_t nb = L0.nn_box();
        std::vector<double> pos;
        std::vector<size_t> ids;
        for (size_t v = 0; v < nb; v++)
            if (L0.vact[v]) {
                pos.push_back((double)(v % nx)); pos.push_back((double)(v / nx % ny));
                pos.push_back((double)(v / ((size_t)nx * ny)));
                ids.push_back(v);
            }
        std::vector<double> R;
        if (rigid_basis(pos, R) != 6) throw std::domain_error("fewer than 6 rigid motions");
        const size_t m = 3 * ids.size();
        std::vector<double> box(3 * nb);
        H.rigid_f.resize(6);
        for (int k = 0; k < 6; k++) {
            std::fill(box.begin(), box.end(), 0.0);
            for (size_t i = 0; i < ids.size(); i++)
                for (int c = 0; c < 3; c++) box[c * nb + ids[i]] = R[k * m + 3 * i + c];
            H.rigid_f[k].alloc(L0.g.N);
            upload_vec<double>(H, L0.g, box.data(), H.rigid_f[k]);
        }
    }
    for (int k = 0; k < 6; k++) {
        const double w = dot<double, double>(H, H.x64.data(), H.rigid_f[k].data(), n);
        const double* hv[2] = {H.x64.data(), H.rigid_f[k].data()};
        const double c[2] = {1.0, -w};
        launch_lincomb(H.x64.data(), hv, c, 2, n, H.s);
    }
}

// modifiedPM (optimization.cpp:123-261), with-support branch, from the current
// force: project + unitize, u = 0, then one V(1,1) per step; the new force is
// the unit projection of u; stop when |fs_new - fs_old| / |fs_old| <= 1e-4 and
// the cycle's relative residual <= 1e-2, or after max_itn steps.
template <typename T>
static double modified_pm(Hier& H, int max_itn, int* steps) {
    if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    force_project(H, true);
    launch_fill<double>(H.x64.data(), 0.0, (size_t)3 * H.lv[0].g.N, H.s);
    double fch = 1, rel = 1;
    int itn = 0, nv = 0;
    const bool free_body = !has_supports(H);
    while (itn++ < max_itn && (fch > 1e-4 || rel > 1e-2)) {
        rel = vcycle_on_solution<T>(H, 1, 1);
        nv++;
        // no supports: remove the rigid motion (optimization.cpp:189-194)
        if (free_body) displacement_project(H);
        if (!(rel <= 1e4)) throw std::domain_error("modified power method diverged");
        const double n2 = pm_balance(H, pm_gather(H, H.x64));
        if (!(n2 > 0)) throw std::domain_error("projected displacement is zero");
        const auto dn = pm_commit(H, 1.0 / std::sqrt(n2));
        fch = dn.first / dn.second;
    }
    *steps = nv;
    return dot<double, double>(H, H.f64.data(), H.x64.data(), (size_t)3 * H.lv[0].g.N);
}

// ---------------------------------------------------------------------------
// worst-case load search by block Lanczos (no reference counterpart; the
// reference's modifiedPM above finds the same top eigenpair of C = P N^T K^-1 N P
// on the admissible load space, P the per-node projector of forceProject)
// ---------------------------------------------------------------------------
// Dense symmetric eigenproblem by cyclic Jacobi rotations (small block
// tridiagonal matrices only): A (n x n, row-major) is destroyed, w gets the
// eigenvalues and V (row-major, columns = eigenvectors).
static void jacobi_eigen(int n, std::vector<double>& A, std::vector<double>& w, std::vector<double>& V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                const double a = A[(size_t)i * n + j] * A[(size_t)i * n + j];
                tot += a;
                if (i != j) off += a;
            }
        if (off <= 1e-30 * tot || off == 0) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0) continue;
                const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                const double th = (aqq - app) / (2 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
                const double c = 1 / std::sqrt(t * t + 1), sn = t * c;
                for (int k = 0; k < n; k++) {   // A <- J^T A J
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - sn * akq;
                    A[(size_t)k * n + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - sn * aqk;
                    A[(size_t)q * n + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - sn * vkq;
                    V[(size_t)k * n + q] = sn * vkp + c * vkq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; i++) w[i] = A[(size_t)i * n + i];
}

// ---------------------------------------------------------------------------
// block Lanczos on the GPU: the Krylov basis is an m x (max_blocks + 1) k
// column-major FP64 matrix in device memory (m = 3 x load nodes); projections
// against it are rocBLAS GEMMs on the handle's stream; the block tridiagonal T
// and its eigenproblem ((j + 1) k square) stay on the host.
// ---------------------------------------------------------------------------
static double ddot(Hier& H, int n, const double* x, const double* y) {
    double r = 0;
    if (rocblas_ddot(H.blas, n, x, 1, y, 1, &r) != rocblas_status_success) throw std::runtime_error("rocblas_ddot failed");
    return r;
}
static void daxpy(Hier& H, int n, double a, const double* x, double* y) {
    if (rocblas_daxpy(H.blas, n, &a, x, 1, y, 1) != rocblas_status_success) throw std::runtime_error("rocblas_daxpy failed");
}
static void dscal(Hier& H, int n, double a, double* x) {
    if (rocblas_dscal(H.blas, n, &a, x, 1) != rocblas_status_success) throw std::runtime_error("rocblas_dscal failed");
}

// Orthonormalises the k columns of W (m x k, device) against the first nq
// columns of Q (two classical Gram-Schmidt passes as GEMMs), then among
// themselves (modified Gram-Schmidt, two passes). W_in = Q C + W_out R with R
// upper triangular (k x k, row-major, host); a column that vanishes (norm
// below 1e-12 of its norm on entry) is zeroed with R_jj = 0. Returns the rank.
static int block_orth_dev(Hier& H, int m, int k, double* W, const double* Q, int nq, double* C,
                          std::vector<double>& R, std::vector<char>& zero) {
    std::vector<double> n0(k);
    for (int j = 0; j < k; j++) n0[j] = std::sqrt(std::max(0.0, ddot(H, m, W + (size_t)j * m, W + (size_t)j * m)));
    if (nq > 0)
        for (int pass = 0; pass < 2; pass++) {
            gemm(H, rocblas_operation_transpose, rocblas_operation_none, nq, k, m, 1.0, Q, m, W, m, 0.0, C, nq);
            gemm(H, rocblas_operation_none, rocblas_operation_none, m, k, nq, -1.0, Q, m, C, nq, 1.0, W, m);
        }
    R.assign((size_t)k * k, 0.0);
    zero.assign(k, 0);
    int rank = 0;
    for (int j = 0; j < k; j++) {
        double* wj = W + (size_t)j * m;
        for (int pass = 0; pass < 2; pass++)
            for (int i = 0; i < j; i++) {
                if (zero[i]) continue;
                const double c = ddot(H, m, W + (size_t)i * m, wj);
                R[(size_t)i * k + j] += c;
                daxpy(H, m, -c, W + (size_t)i * m, wj);
            }
        const double nj = std::sqrt(std::max(0.0, ddot(H, m, wj, wj)));
        if (nj <= 1e-12 * std::max(n0[j], 1e-300)) {
            dscal(H, m, 0.0, wj);
            zero[j] = 1;
            continue;
        }
        R[(size_t)j * k + j] = nj;
        dscal(H, m, 1.0 / nj, wj);
        rank++;
    }
    return rank;
}

// y = C q = P N^T K^-1 N q for a compact load vector q ([3][nl] SoA, device,
// in the range of P): force = q at the load nodes, a PCG solve from zero to
// solve_tol, then the projected gather of the displacement into y (device).
static void worst_apply(Hier& H, const double* q, double* y, double solve_tol, int mode, int pre, int post,
                        int* iters) {
    const size_t nl3 = (size_t)3 * H.nload, n = (size_t)3 * H.lv[0].g.N;
    RT_HIP(hipMemcpyAsync(H.pm_fs.p, q, sizeof(double) * nl3, hipMemcpyDeviceToDevice, H.s));
    launch_fill<double>(H.f64.data(), 0.0, n, H.s);
    pm_commit(H, 1.0);
    launch_fill<double>(H.x64.data(), 0.0, n, H.s);
    rt_solve_stats st{};
    solve_mode(H, solve_tol, 2000, mode, pre, post, &st);
    if (!(st.rel_res <= 10 * solve_tol)) throw std::runtime_error("worst-case solve did not converge");
    *iters += st.iterations;
    pm_gather(H, H.x64);
    RT_HIP(hipMemcpyAsync(y, H.pm_fs.p, sizeof(double) * nl3, hipMemcpyDeviceToDevice, H.s));
}

// ---------------------------------------------------------------------------
// Batched solves for the worst-case search: nb <= kBatchMax right-hand sides
// through one preconditioner cycle. The finest level (whose traffic is the
// vectors themselves) runs the single-vector kernels once per vector; every
// coarser level runs batched kernels that read each stencil coefficient once
// for all nb vectors, and the coarsest applies its inverse to all nb at once.
// The whole batched cycle is captured into one HIP graph. No counterpart in
// the reference, which solves one load at a time (optimization.cpp:123-261).
// ---------------------------------------------------------------------------
static bool batch_supported(const Hier& H) { return H.pc_gamma == 1 && H.coarse_bits == 32; }

// room per vector for the partial sums of any kernel the batched solves use
static int batch_part_stride(const Hier& H) {
    return std::max(std::max(1024, kStreamBlocks), fine_apply_partials_modal(H.lv[0].g));
}

static void ensure_batch(Hier& H, int nb) {
    if (nb < 1 || nb > kBatchMax) throw std::invalid_argument("batch of right-hand sides must be 1..8");
    bool grown = false;
    for (Level& L : H.lv)
        if (!L.bU.fits(L.g.N, nb)) {
            L.bU.alloc(L.g.N, nb);
            L.bF.alloc(L.g.N, nb);
            L.bR.alloc(L.g.N, nb);
            grown = true;
        }
    const int N = H.lv[0].g.N;
    for (BVec<double>* v : {&H.bx, &H.bb, &H.br, &H.bp, &H.bq})
        if (!v->fits(N, nb)) { v->alloc(N, nb); grown = true; }
    H.bfbuf.ensure(sizeof(double) * (size_t)nb * std::max(std::max(H.n3, H.pcn3), 1));
    for (BVec<float>* v : {&H.bp32, &H.bq32, &H.bdx32})
        if (!v->fits(N, nb)) { v->alloc(N, nb); grown = true; }
    H.bpart.ensure(sizeof(double) * (size_t)nb * batch_part_stride(H));
    H.bscal.ensure(sizeof(double) * kBatchMax);
    if (grown && H.bgraph) {
        (void)hipGraphExecDestroy(H.bgraph);
        H.bgraph = nullptr;
        H.bkey = -1;
    }
}

// V(pre, post) on the nb vectors L0.bF -> L0.bU (L0.bU zero on entry), the
// preconditioner's cycle (vcycle_core with zero_support and islands)
static void vcycle_batch(Hier& H, int nb, int pre, int post) {
    const int n = H.pc_levels > 0 ? H.pc_levels : (int)H.lv.size();
    const bool early = n < (int)H.lv.size();
    // zero_start: the fine vectors hold no valid values yet (first pre-smoothing)
    auto smooth = [&](int i, int nsw, bool rev, bool zero_start) {
        Level& L = H.lv[i];
        if (i == 0) {
            for (int b = 0; b < nb; b++) gs<float>(H, 0, L.bU.v(b), L.bF.cv(b), nsw, rev, true, zero_start);
            return;
        }
        for (int it = 0; it < nsw; it++)
            for (int cc = 0; cc < 8; cc++)
                launch_coarse_gs_batch<float, float>(L.g, nb, L.bU.data(), L.bF.data(), L.st32.as<float>(),
                                                     L.vflag.as<uint8_t>(), rev ? 7 - cc : cc, rev, H.s);
    };
    auto cyc = [&](auto&& self, int i) -> void {
        Level& L = H.lv[i];
        if (i == n - 1) {
            launch_coarsest_batch<float>(early ? H.pcAinv.as<double>() : H.Ainv.as<double>(), early ? H.pcn3 : H.n3,
                                         early ? H.pcrowslot.as<int>() : H.rowslot.as<int>(), nb, L.g.N,
                                         L.bF.data(), L.bU.data(), H.bfbuf.as<double>(), H.s);
            return;
        }
        const int m = i == 0 ? 1 : level_mult(H, i);
        smooth(i, pre * m, false, i == 0);
        if (i == 0)
            for (int b = 0; b < nb; b++) residual<float>(H, 0, L.bU.cv(b), L.bF.cv(b), L.bR.v(b));
        else
            launch_coarse_residual_batch<float, float>(L.g, nb, L.bU.data(), L.bF.data(), L.bR.data(),
                                                       L.st32.as<float>(), L.vflag.as<uint8_t>(), H.s);
        Level& C = H.lv[i + 1];
        for (int b = 0; b < nb; b++) restrict_to<float>(H, i + 1, L.bR.cv(b), C.bF.v(b));
        launch_fill<float>(C.bU.data(), 0.f, (size_t)3 * C.g.N * nb, H.s);
        self(self, i + 1);
        for (int b = 0; b < nb; b++) prolong_from<float>(H, i, L.bU.v(b), C.bU.cv(b), i == 0);
        if (i == 0)
            for (int b = 0; b < nb; b++)
                island_correct<float>(L.g, H.nisl, H.isl_node.as<int>(), H.isl_cent.as<double>(),
                                      H.isl_Einv.as<double>(), H.isl_zr.as<double>(), H.isl_coef.as<double>(),
                                      L.bR.cv(b), L.bU.v(b), H.s);
        smooth(i, post * m, true, false);
    };
    cyc(cyc, 0);
}

// L0.bU = M^-1 L0.bF for nb vectors (from L0.bU = 0: the first fine sweep starts
// from zero itself); replayed from a HIP graph after the first (eager)
// application of a shape
static void precondition_batch(Hier& H, int nb, int pre, int post) {
    const long long key = (((long long)nb * 64 + pre) * 64 + post) * 64 + H.pc_mult;
    if (H.use_graphs && H.bgraph && H.bkey == key) {
        RT_HIP(hipGraphLaunch(H.bgraph, H.s));
        return;
    }
    vcycle_batch(H, nb, pre, post);
    if (!H.use_graphs) return;
    if (H.bgraph) (void)hipGraphExecDestroy(H.bgraph);
    H.bgraph = nullptr;
    H.bkey = -1;
    RT_HIP(hipStreamBeginCapture(H.s, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
        vcycle_batch(H, nb, pre, post);
    } catch (...) {
        (void)hipStreamEndCapture(H.s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    RT_HIP(hipStreamEndCapture(H.s, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    RT_HIP(e);
    H.bgraph = ex;
    H.bkey = key;
}

// out[b] = sum of the np partials of vector b (bpart + b * stride), one sync
static void batch_sums(Hier& H, int nb, int np, int stride, double* out) {
    for (int b = 0; b < nb; b++)
        launch_sum_partials(H.bpart.as<double>() + (size_t)b * stride, np, H.bscal.as<double>() + b, H.s);
    RT_HIP(hipMemcpyAsync(out, H.bscal.p, sizeof(double) * nb, hipMemcpyDeviceToHost, H.s));
    RT_HIP(hipStreamSynchronize(H.s));
}

// nb preconditioned CG solves K x_b = f_b (H.bb -> H.bx, from zero) in
// lockstep: FP64 Krylov vectors and dots per vector, the batched FP32 cycle as
// preconditioner; a vector stops updating once its recurrence residual is
// below tol |f_b|. iters[b] = its iterations, rel[b] = its true residual.
static void pcg_batch(Hier& H, int nb, double tol, int maxit, int pre, int post, int* iters, double* rel) {
    if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    ensure_batch(H, nb);
    Level& L0 = H.lv[0];
    const size_t n = (size_t)3 * L0.g.N;
    const int np = 1024, npa = batch_part_stride(H);
    double nf[kBatchMax], rr[kBatchMax], rz[kBatchMax], pq[kBatchMax];
    bool act[kBatchMax];
    for (int b = 0; b < nb; b++) {
        launch_fill<double>(H.bx.at(b), 0.0, n, H.s);
        launch_convert<double, double>(H.bb.at(b), H.br.at(b), n, H.s);
        launch_convert<double, float>(H.bb.at(b), L0.bF.at(b), n, H.s);
        launch_dot_partials_t<double, double>(H.bb.at(b), H.bb.at(b), n, H.bpart.as<double>() + (size_t)b * npa, np,
                                              H.s);
        iters[b] = 0;
    }
    batch_sums(H, nb, np, npa, nf);
    int live = 0;
    for (int b = 0; b < nb; b++) {
        nf[b] = std::sqrt(nf[b]);
        act[b] = nf[b] > 0;
        live += act[b];
    }
    if (live) {
        precondition_batch(H, nb, pre, post);
        for (int b = 0; b < nb; b++) {
            launch_dot_partials_t<double, float>(H.br.at(b), L0.bU.at(b), n, H.bpart.as<double>() + (size_t)b * npa,
                                                 np, H.s);
            launch_convert<float, double>(L0.bU.at(b), H.bp.at(b), n, H.s);
        }
        batch_sums(H, nb, np, npa, rz);
    }
    for (int it = 1; it <= maxit && live; it++) {
        int npq = 0;
        for (int b = 0; b < nb; b++) {
            if (!act[b]) continue;
            launch_fine_apply_modal<double>(L0.g, H.bp.cv(b), H.bq.v(b), H.rhop.as<float>(), L0.vflag.as<uint8_t>(),
                                            H.bpart.as<double>() + (size_t)b * npa, &npq, H.s);
        }
        batch_sums(H, nb, npq, npa, pq);
        for (int b = 0; b < nb; b++) {
            if (!act[b]) continue;
            if (!(pq[b] > 0)) throw std::domain_error("PCG breakdown: p.Ap <= 0");
            launch_pcg_update_x_r<double, float>(H.bx.at(b), H.br.at(b), H.bp.at(b), H.bq.at(b), rz[b] / pq[b], n,
                                                 H.bpart.as<double>() + (size_t)b * npa, np, L0.bF.at(b), H.s);
        }
        batch_sums(H, nb, np, npa, rr);
        live = 0;
        for (int b = 0; b < nb; b++) {
            if (!act[b]) continue;
            iters[b] = it;
            if (std::sqrt(rr[b]) <= tol * nf[b]) act[b] = false;
            else live++;
        }
        if (!live) break;
        precondition_batch(H, nb, pre, post);
        for (int b = 0; b < nb; b++)
            if (act[b])
                launch_dot_partials_t<double, float>(H.br.at(b), L0.bU.at(b), n,
                                                     H.bpart.as<double>() + (size_t)b * npa, np, H.s);
        double rzn[kBatchMax];
        batch_sums(H, nb, np, npa, rzn);
        for (int b = 0; b < nb; b++) {
            if (!act[b]) continue;
            launch_pcg_update_p<double, float>(H.bp.at(b), L0.bU.at(b), rzn[b] / rz[b], n, H.s);
            rz[b] = rzn[b];
        }
    }
    // true residuals |f_b - K x_b| / |f_b|
    for (int b = 0; b < nb; b++) {
        launch_fine_apply_modal<double>(L0.g, H.bx.cv(b), H.bq.v(b), H.rhop.as<float>(), L0.vflag.as<uint8_t>(),
                                        nullptr, nullptr, H.s);
        launch_sub_conv<double, double>(H.bb.at(b), H.bq.at(b), H.br.at(b), (double*)nullptr, n,
                                        H.bpart.as<double>() + (size_t)b * npa, np, H.s);
    }
    batch_sums(H, nb, np, npa, rr);
    for (int b = 0; b < nb; b++) rel[b] = nf[b] > 0 ? std::sqrt(rr[b]) / nf[b] : 0.0;
}

// The same nb solves with reliable updates (pcg_reliable's scheme per
// vector): FP32 r, p, q and FP32 operator, the solution increment in FP32,
// flushed into the FP64 solution with an FP64 true residual whenever a
// vector's FP32 residual has dropped 1/RT_REPLACE_DROP since its last
// replacement, and before it stops. A vector whose replaced residual comes
// out 10x above its best one (the FP32 recurrence lost it) is solved again
// from zero with FP64 Krylov vectors (pcg_batch).
static void pcg_batch_reliable(Hier& H, int nb, double tol, int maxit, int pre, int post, int* iters, double* rel) {
    if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    ensure_batch(H, nb);
    Level& L0 = H.lv[0];
    const size_t n = (size_t)3 * L0.g.N;
    const int np = 1024, ns = kStreamBlocks, npa = batch_part_stride(H);
    double* part = H.bpart.as<double>();
    double nf[kBatchMax], rmax[kBatchMax], best[kBatchMax], rz[kBatchMax], pq[kBatchMax], rn[kBatchMax],
        rt[kBatchMax];
    bool act[kBatchMax], bad[kBatchMax], dirty[kBatchMax];
    // apend[b]: the step whose alpha p is not yet on dx (added by the next
    // direction update or by the flush, as in pcg_reliable)
    double apend[kBatchMax] = {};
    for (int b = 0; b < nb; b++) {
        launch_fill<double>(H.bx.at(b), 0.0, n, H.s);
        launch_fill<float>(H.bdx32.at(b), 0.f, n, H.s);
        launch_convert<double, float>(H.bb.at(b), L0.bF.at(b), n, H.s);
        launch_dot_partials_t<double, double>(H.bb.at(b), H.bb.at(b), n, part + (size_t)b * npa, np, H.s);
        iters[b] = 0;
        bad[b] = dirty[b] = false;
    }
    batch_sums(H, nb, np, npa, nf);
    int live = 0;
    for (int b = 0; b < nb; b++) {
        nf[b] = std::sqrt(nf[b]);
        rmax[b] = nf[b];
        best[b] = 1.0;
        rel[b] = 0.0;
        act[b] = nf[b] > 0;
        live += act[b];
    }
    // true residuals of the vectors in `which`: r64 = f - K x64, rt[b] = |r64| / |f|
    auto true_res = [&](const bool* which) {
        for (int b = 0; b < nb; b++) {
            if (!which[b]) continue;
            launch_fine_apply_modal<double>(L0.g, H.bx.cv(b), H.bq.v(b), H.rhop.as<float>(), L0.vflag.as<uint8_t>(),
                                            nullptr, nullptr, H.s);
            launch_sub_conv<double, double>(H.bb.at(b), H.bq.at(b), H.br.at(b), (double*)nullptr, n,
                                            part + (size_t)b * npa, np, H.s);
        }
        double ss[kBatchMax];
        batch_sums(H, nb, np, npa, ss);
        for (int b = 0; b < nb; b++)
            if (which[b]) rt[b] = std::sqrt(ss[b]) / nf[b];
    };
    if (live) {
        precondition_batch(H, nb, pre, post);
        for (int b = 0; b < nb; b++) {
            launch_dot_partials_t<float, float>(L0.bF.at(b), L0.bU.at(b), n, part + (size_t)b * npa, ns, H.s,
                                                vec_chunks(H));
            launch_convert<float, float>(L0.bU.at(b), H.bp32.at(b), n, H.s);
        }
        batch_sums(H, nb, ns, npa, rz);
    }
    for (int it = 1; it <= maxit && live; it++) {
        int npq = 0;
        for (int b = 0; b < nb; b++)
            if (act[b])
                launch_fine_apply_modal<float>(L0.g, H.bp32.cv(b), H.bq32.v(b), H.rhop.as<float>(),
                                               L0.vflag.as<uint8_t>(), part + (size_t)b * npa, &npq, H.s);
        batch_sums(H, nb, npq, npa, pq);
        for (int b = 0; b < nb; b++) {
            if (!act[b]) continue;
            if (!(pq[b] > 0)) throw std::domain_error("PCG breakdown: p.Ap <= 0");
            apend[b] = rz[b] / pq[b];
            launch_pcg_ru_dxr(L0.bF.at(b), H.bq32.at(b), nullptr, nullptr, apend[b], n, part + (size_t)b * npa, ns,
                              H.s, vec_chunks(H));
            dirty[b] = true;
        }
        batch_sums(H, nb, ns, npa, rn);
        bool rep[kBatchMax] = {};
        int nrep = 0;
        for (int b = 0; b < nb; b++) {
            if (!act[b]) continue;
            iters[b] = it;
            rn[b] = std::sqrt(rn[b]);
            if (rn[b] / nf[b] <= tol || rn[b] < RT_REPLACE_DROP * rmax[b] || it == maxit) {
                launch_pcg_flush_x(H.bx.at(b), H.bdx32.at(b), n, H.s, H.bp32.at(b), apend[b]);
                apend[b] = 0.0;
                dirty[b] = false;
                rep[b] = true;
                nrep++;
            }
        }
        if (nrep) {
            true_res(rep);
            for (int b = 0; b < nb; b++) {
                if (!rep[b]) continue;
                rel[b] = rt[b];
                if (rt[b] <= tol || it == maxit) { act[b] = false; continue; }
                if (!(rt[b] <= 10 * best[b])) { act[b] = false; bad[b] = true; continue; }
                best[b] = std::min(best[b], rt[b]);
                launch_convert<double, float>(H.br.at(b), L0.bF.at(b), n, H.s);
                rmax[b] = rt[b] * nf[b];
            }
        }
        live = 0;
        for (int b = 0; b < nb; b++) live += act[b];
        if (!live) break;
        precondition_batch(H, nb, pre, post);
        for (int b = 0; b < nb; b++)
            if (act[b])
                launch_dot_partials_t<float, float>(L0.bF.at(b), L0.bU.at(b), n, part + (size_t)b * npa, ns, H.s,
                                                    vec_chunks(H));
        double rzn[kBatchMax];
        batch_sums(H, nb, ns, npa, rzn);
        for (int b = 0; b < nb; b++) {
            if (!act[b]) continue;
            if (apend[b] != 0.0)
                launch_pcg_p_dx(H.bp32.at(b), L0.bU.at(b), rzn[b] / rz[b], H.bdx32.at(b), apend[b], n, H.s,
                                vec_chunks(H));
            else
                launch_pcg_update_p<float, float>(H.bp32.at(b), L0.bU.at(b), rzn[b] / rz[b], n, H.s, vec_chunks(H));
            apend[b] = 0.0;
            rz[b] = rzn[b];
        }
    }
    for (int b = 0; b < nb; b++)
        if (dirty[b]) launch_pcg_flush_x(H.bx.at(b), H.bdx32.at(b), n, H.s, H.bp32.at(b), apend[b]);
    // a vector whose FP32 recurrence failed: all of them again with FP64
    // Krylov vectors (pcg_batch zero-starts every vector; the others' results
    // are kept and restored)
    int nbad = 0;
    for (int b = 0; b < nb; b++) nbad += bad[b];
    if (nbad) {
        int it2[kBatchMax];
        double rel2[kBatchMax];
        std::vector<std::vector<double>> keep(nb);
        for (int b = 0; b < nb; b++)
            if (!bad[b]) {
                keep[b].resize(n);
                RT_HIP(hipMemcpyAsync(keep[b].data(), H.bx.at(b), sizeof(double) * n, hipMemcpyDeviceToHost, H.s));
            }
        RT_HIP(hipStreamSynchronize(H.s));
        pcg_batch(H, nb, tol, maxit, pre, post, it2, rel2);
        for (int b = 0; b < nb; b++) {
            if (bad[b]) {
                iters[b] += it2[b];
                rel[b] = rel2[b];
            } else {
                RT_HIP(hipMemcpyAsync(H.bx.at(b), keep[b].data(), sizeof(double) * n, hipMemcpyHostToDevice, H.s));
            }
        }
        RT_HIP(hipStreamSynchronize(H.s));
    }
}

// Y[:, c] = C Q[:, c] for the nb columns of a block (worst_apply for all of
// them in one batched solve); zero columns give zero
static void worst_apply_batch(Hier& H, int nb, const double* Q, double* Y, const std::vector<char>& zero,
                              double solve_tol, int mode, int pre, int post, int* iters) {
    const size_t nl3 = (size_t)3 * H.nload, n = (size_t)3 * H.lv[0].g.N;
    ensure_batch(H, nb);
    for (int b = 0; b < nb; b++) {
        launch_fill<double>(H.bb.at(b), 0.0, n, H.s);
        if (!zero[b])
            launch_pm_commit(H.nload, H.load_slots.as<int>(), Q + (size_t)b * nl3, 1.0, H.bb.v(b),
                             H.pm_old.as<double>(), H.partials.as<double>(), kPmParts, H.s);
    }
    int it[kBatchMax];
    double rel[kBatchMax];
    if (mode == RT_SOLVE_RELIABLE) pcg_batch_reliable(H, nb, solve_tol, 2000, pre, post, it, rel);
    else pcg_batch(H, nb, solve_tol, 2000, pre, post, it, rel);
    for (int b = 0; b < nb; b++) {
        if (zero[b]) {
            RT_HIP(hipMemsetAsync(Y + (size_t)b * nl3, 0, sizeof(double) * nl3, H.s));
            continue;
        }
        if (!(rel[b] <= 10 * solve_tol)) throw std::runtime_error("worst-case solve did not converge");
        *iters += it[b];
        launch_pm_gather(H.nload, H.load_slots.as<int>(), H.bx.cv(b), H.load_normals ? H.load_nrm.as<double>() : nullptr,
                         Y + (size_t)b * nl3, H.partials.as<double>(), kPmParts, H.s);
    }
}

// Block Lanczos with full reorthogonalisation on C restricted to the
// admissible load space, block size k, from a seeded random block. Stops when
// the residual of the top Ritz pair, |B_j y_last|, falls below tol * theta, or
// after max_blocks blocks. Leaves the unit Ritz load in the force and its
// displacement (solved to solve_tol) in the solution; returns theta.
static double block_lanczos(Hier& H, int k, int max_blocks, double tol, double solve_tol, int mode, int pre,
                            int post, uint64_t seed, int* blocks_out, int* solves_out, int* iters_out) {
    if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    if (!H.nload) throw std::logic_error("no load nodes set");
    if (k < 1 || k > 64 || max_blocks < 1) throw std::invalid_argument("block size must be 1..64, max_blocks >= 1");
    const size_t nl = (size_t)H.nload, m = 3 * nl;
    if (m > (size_t)INT32_MAX / 2) throw std::invalid_argument("too many load nodes");
    std::vector<double> nrm;
    if (H.load_normals) {
        nrm.resize(m);
        RT_HIP(hipMemcpy(nrm.data(), H.load_nrm.p, sizeof(double) * m, hipMemcpyDeviceToHost));
    }
    const size_t dim = nrm.empty() ? m : nl;   // dimension of the admissible space
    k = (int)std::min<size_t>((size_t)k, dim);
    const int mi = (int)m;
    // the k right-hand sides of a block are solved together (FP64 Krylov
    // vectors, batched FP32 cycle) when the cycle has a batched form
    const bool batched = H.batch_solves && k <= kBatchMax && batch_supported(H) &&
                         (mode == RT_SOLVE_MIXED || mode == RT_SOLVE_RELIABLE);
    // basis: block j in columns [j k, (j + 1) k); block j + 1 is the work block
    DBuf Qd, Cd, Md, fd;
    Qd.alloc(sizeof(double) * m * (size_t)(max_blocks + 1) * k);
    Cd.alloc(sizeof(double) * (size_t)(max_blocks + 1) * k * k);
    Md.alloc(sizeof(double) * (size_t)k * k);
    fd.alloc(sizeof(double) * m);
    double* Q = Qd.as<double>();
    auto col = [&](int c) { return Q + (size_t)c * m; };
    // seeded start block (splitmix64 uniforms in [-1, 1)), projected per load node onto n n^T
    {
        uint64_t st = seed;
        auto next = [&st]() {
            uint64_t z = (st += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
        };
        std::vector<double> v((size_t)k * m);
        for (int c = 0; c < k; c++) {
            double* x = v.data() + (size_t)c * m;
            for (size_t i = 0; i < m; i++) x[i] = next();
            if (!nrm.empty())
                for (size_t i = 0; i < nl; i++) {
                    const double d = x[i] * nrm[i] + x[nl + i] * nrm[nl + i] + x[2 * nl + i] * nrm[2 * nl + i];
                    x[i] = d * nrm[i]; x[nl + i] = d * nrm[nl + i]; x[2 * nl + i] = d * nrm[2 * nl + i];
                }
        }
        RT_HIP(hipMemcpyAsync(Q, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, H.s));
    }
    std::vector<double> R;
    std::vector<char> zero;
    int r = block_orth_dev(H, mi, k, Q, nullptr, 0, Cd.as<double>(), R, zero);
    if (r < 1) throw std::domain_error("block Lanczos start block is zero");
    std::vector<double> Bprev;                               // k x k row-major, W_{j-1} = Q_j B_{j-1}
    std::vector<std::vector<double>> Ablk, Bblk;             // diagonal and sub-diagonal blocks of T
    int solves = 0, iters = 0, j = 0;
    double theta = 0;
    std::vector<double> ritz;                                // top Ritz vector coefficients
    for (j = 0; j < max_blocks; j++) {
        double* blk = col(j * k);
        double* W = col((j + 1) * k);
        if (batched) {
            worst_apply_batch(H, k, blk, W, zero, solve_tol, mode, pre, post, &iters);
            for (int c = 0; c < k; c++) solves += !zero[c];
        } else {
            for (int c = 0; c < k; c++) {
                if (zero[c]) {
                    RT_HIP(hipMemsetAsync(W + (size_t)c * m, 0, sizeof(double) * m, H.s));
                    continue;
                }
                worst_apply(H, blk + (size_t)c * m, W + (size_t)c * m, solve_tol, mode, pre, post, &iters);
                solves++;
            }
        }
        // A_j = Q_j^T W (symmetrised); column-major k x k, A(a, b) = q_a . w_b
        std::vector<double> A((size_t)k * k);
        gemm(H, rocblas_operation_transpose, rocblas_operation_none, k, k, mi, 1.0, blk, mi, W, mi, 0.0, Md.as<double>(), k);
        RT_HIP(hipMemcpyAsync(A.data(), Md.p, sizeof(double) * k * k, hipMemcpyDeviceToHost, H.s));
        RT_HIP(hipStreamSynchronize(H.s));
        for (int a = 0; a < k; a++)
            for (int b = a + 1; b < k; b++) A[(size_t)a * k + b] = A[(size_t)b * k + a] =
                0.5 * (A[(size_t)a * k + b] + A[(size_t)b * k + a]);
        Ablk.push_back(A);
        // W -= Q_j A_j + Q_{j-1} B_{j-1}^T, then full reorthogonalisation against Q_0..Q_j
        RT_HIP(hipMemcpyAsync(Md.p, A.data(), sizeof(double) * k * k, hipMemcpyHostToDevice, H.s));
        gemm(H, rocblas_operation_none, rocblas_operation_none, mi, k, k, -1.0, blk, mi, Md.as<double>(), k, 1.0, W, mi);
        if (j > 0) {
            // (B^T)(a, b) = B[b][a]: the row-major B read as column-major is B^T
            RT_HIP(hipMemcpyAsync(Md.p, Bprev.data(), sizeof(double) * k * k, hipMemcpyHostToDevice, H.s));
            gemm(H, rocblas_operation_none, rocblas_operation_none, mi, k, k, -1.0, col((j - 1) * k), mi,
                 Md.as<double>(), k, 1.0, W, mi);
        }
        std::vector<double> B;
        r = block_orth_dev(H, mi, k, W, Q, (j + 1) * k, Cd.as<double>(), B, zero);   // W = Q_{j+1} B_j
        Bblk.push_back(B);
        // T: block tridiagonal, (j+1) k square
        const int nt = (j + 1) * k;
        std::vector<double> T((size_t)nt * nt, 0.0);
        for (int i = 0; i <= j; i++) {
            for (int a = 0; a < k; a++)
                for (int b = 0; b < k; b++) T[(size_t)(i * k + a) * nt + i * k + b] = Ablk[i][(size_t)a * k + b];
            if (i < j)
                for (int a = 0; a < k; a++)
                    for (int b = 0; b < k; b++) {
                        const double v = Bblk[i][(size_t)a * k + b];   // block (i+1, i) = B_i
                        T[(size_t)((i + 1) * k + a) * nt + i * k + b] = v;
                        T[(size_t)(i * k + b) * nt + (i + 1) * k + a] = v;
                    }
        }
        std::vector<double> w, V;
        jacobi_eigen(nt, T, w, V);
        int top = 0;
        for (int i = 1; i < nt; i++) if (w[i] > w[top]) top = i;
        theta = w[top];
        ritz.resize(nt);
        for (int i = 0; i < nt; i++) ritz[i] = V[(size_t)i * nt + top];
        // residual of the top Ritz pair: |B_j y_last|
        double res = 0;
        for (int a = 0; a < k; a++) {
            double s2 = 0;
            for (int b = 0; b < k; b++) s2 += B[(size_t)a * k + b] * ritz[(size_t)j * k + b];
            res += s2 * s2;
        }
        res = std::sqrt(res);
        Bprev = B;
        if (res <= tol * std::fabs(theta) || r == 0 || (size_t)(j + 2) * k > dim) { j++; break; }
    }
    // Ritz load f = Q y, unit length, then its displacement
    const int nt = (int)ritz.size();
    DBuf yd;
    yd.alloc(sizeof(double) * nt);
    RT_HIP(hipMemcpyAsync(yd.p, ritz.data(), sizeof(double) * nt, hipMemcpyHostToDevice, H.s));
    gemm(H, rocblas_operation_none, rocblas_operation_none, mi, 1, nt, 1.0, Q, mi, yd.as<double>(), nt, 0.0,
         fd.as<double>(), mi);
    const double fn = std::sqrt(ddot(H, mi, fd.as<double>(), fd.as<double>()));
    if (!(fn > 0)) throw std::domain_error("Ritz load is zero");
    dscal(H, mi, 1.0 / fn, fd.as<double>());
    worst_apply(H, fd.as<double>(), col(0), solve_tol, mode, pre, post, &iters);
    solves++;
    RT_HIP(hipStreamSynchronize(H.s));
    if (blocks_out) *blocks_out = j;
    if (solves_out) *solves_out = solves;
    if (iters_out) *iters_out = iters;
    return dot<double, double>(H, H.f64.data(), H.x64.data(), (size_t)3 * H.lv[0].g.N);
}

extern "C" {

int rt_set_load_nodes(rt_hier* h, const uint8_t* load_box, const double* normals_box) {
    RT_TRY({
        need(load_box);
        set_load_nodes(H_(h), load_box, normals_box);
    })
}

int rt_num_load_nodes(const rt_hier* h, int* n) {
    RT_TRY({ need(n); *n = H_(h).nload; })
}

int rt_force_project(rt_hier* h) {
    RT_TRY({ force_project(H_(h), false); })
}

int rt_rand_force(rt_hier* h, uint64_t seed) {
    RT_TRY({
        Hier& H = H_(h);
        Level& L0 = H.lv[0];
        launch_rand_force(L0.g, L0.vflag.as<uint8_t>(), seed, H.f64.v(), H.s);
        RT_HIP(hipStreamSynchronize(H.s));
    })
}

int rt_modified_pm(rt_hier* h, int max_itn, double* worst_compliance, int* vcycles) {
    RT_TRY({
        need(worst_compliance);
        Hier& H = H_(h);
        if (max_itn < 1) throw std::invalid_argument("max_itn must be positive");
        int nv = 0;
        *worst_compliance = H.vprec == 64 ? modified_pm<double>(H, max_itn, &nv) : modified_pm<float>(H, max_itn, &nv);
        if (vcycles) *vcycles = nv;
    })
}

int rt_set_block_solves(rt_hier* h, int batched) {
    RT_TRY({ H_(h).batch_solves = batched != 0; })
}

int rt_solve_batch(rt_hier* h, int nb, const double* f_box, double* u_box, double tol, int maxit, int pre, int post,
                   int* iterations, double* rel_res) {
    RT_TRY({
        need(f_box); need(u_box); need(iterations); need(rel_res);
        Hier& H = H_(h);
        if (nb < 1 || nb > kBatchMax) throw std::invalid_argument("nb must be 1..8");
        if (tol <= 0 || maxit < 0 || pre < 0 || post < 0) throw std::invalid_argument("bad tolerance or counts");
        if (!batch_supported(H))
            throw std::logic_error("batched solves need a V-cycle (gamma 1) with FP32 coarse stencils");
        if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
        ensure_batch(H, nb);
        Level& L0 = H.lv[0];
        const size_t nbx = L0.nn_box();
        std::vector<double> box(3 * nbx);
        for (int b = 0; b < nb; b++) {
            std::copy(f_box + (size_t)b * 3 * nbx, f_box + (size_t)(b + 1) * 3 * nbx, box.begin());
            for (size_t v = 0; v < nbx; v++)
                if (!L0.vact[v] || H.fixed_box[v])
                    for (int c = 0; c < 3; c++) box[c * nbx + v] = 0;
            const Vec3<double> fv = H.bb.v(b);
            for (int c = 0; c < 3; c++) {
                std::vector<double> comp(box.begin() + c * nbx, box.begin() + (c + 1) * nbx);
                upload_box<double>(H, L0.g, 0, comp.data(), fv.c[c]);
            }
        }
        pcg_batch(H, nb, tol, maxit, pre, post, iterations, rel_res);
        for (int b = 0; b < nb; b++) {
            const Vec3<double> xv = H.bx.v(b);
            for (int c = 0; c < 3; c++)
                download_box<double>(H, L0.g, 0, xv.c[c], u_box + (size_t)b * 3 * nbx + (size_t)c * nbx);
        }
    })
}

int rt_block_lanczos(rt_hier* h, int k, int max_blocks, double tol, double solve_tol, int mode, int pre, int post,
                     uint64_t seed, double* worst_compliance, rt_lanczos_stats* stats) {
    RT_TRY({
        need(worst_compliance);
        Hier& H = H_(h);
        if (!has_supports(H))
            throw std::logic_error("block Lanczos needs supports; the no-support modes use rt_modified_pm");
        int blocks = 0, solves = 0, iters = 0;
        *worst_compliance = block_lanczos(H, k, max_blocks, tol, solve_tol, mode, pre, post, seed, &blocks, &solves,
                                          &iters);
        if (stats) {
            stats->blocks = blocks;
            stats->solves = solves;
            stats->pcg_iterations = iters;
        }
    })
}

int rt_compute_sensitivity(rt_hier* h) {
    RT_TRY({
        Hier& H = H_(h);
        launch_fill<float>(H.sens.as<float>(), 0.f, (size_t)H.lv[0].g.N, H.s);
        launch_sensitivity<double>(H.lv[0].g, H.x64.cv(), H.rho.as<float>(), H.eflag.as<uint8_t>(), H.p,
                                   H.sens.as<float>(), H.s);
        RT_HIP(hipStreamSynchronize(H.s));
    })
}

int rt_filter_sensitivity(rt_hier* h, double radius) {
    RT_TRY({
        Hier& H = H_(h);
        if (!(radius > 0)) throw std::invalid_argument("filter radius must be positive");
        RT_HIP(hipMemcpyAsync(H.sens_tmp.p, H.sens.p, sizeof(float) * H.lv[0].g.N, hipMemcpyDeviceToDevice, H.s));
        launch_filter(H.lv[0].g, H.sens_tmp.as<float>(), H.eflag.as<uint8_t>(), (float)radius, H.sens.as<float>(), H.s);
        RT_HIP(hipStreamSynchronize(H.s));
    })
}

int rt_get_sensitivity_box(const rt_hier* h, float* g_box) {
    RT_TRY({ need(g_box); Hier& H = H_(h); download_box<float>(H, H.lv[0].g, 1, H.sens.as<float>(), g_box); })
}

int rt_set_sensitivity_box(rt_hier* h, const float* g_box) {
    RT_TRY({ need(g_box); Hier& H = H_(h); upload_box<float>(H, H.lv[0].g, 1, g_box, H.sens.as<float>()); })
}

int rt_update_densities(rt_hier* h, float vgoal, float step, float damp, float rho_min, float* multiplier, int* trials) {
    RT_TRY({
        Hier& H = H_(h);
        Level& L0 = H.lv[0];
        const int np = 1024;
        launch_maxabs(L0.g, H.sens.as<float>(), H.eflag.as<uint8_t>(), H.partials.as<double>(), np, H.s);
        std::vector<double> pm(np);
        RT_HIP(hipMemcpyAsync(pm.data(), H.partials.p, sizeof(double) * np, hipMemcpyDeviceToHost, H.s));
        RT_HIP(hipStreamSynchronize(H.s));
        float gmax = 0.f;
        for (double v : pm) gmax = std::max(gmax, (float)v);
        const double nact = (double)L0.n_eact;
        float lo = 0.f, up = gmax, thr = (lo + up) / 2;
        float V = 2.f;
        int itn = 0, ntr = 0;
        do {   // bisection of optimization.cu:186-219
            thr = (lo + up) / 2;
            launch_oc_trial(L0.g, H.rho.as<float>(), H.sens.as<float>(), H.eflag.as<uint8_t>(), thr, step, damp,
                            rho_min, nullptr, H.partials.as<double>(), np, H.s);
            V = (float)(sum_partials_host(H, np) / nact);
            ntr++;
            if (V > vgoal) lo = thr;
            else if (V < vgoal) up = thr;
        } while (std::fabs(V - vgoal) > 1e-4 && itn++ < 30);
        launch_oc_trial(L0.g, H.rho.as<float>(), H.sens.as<float>(), H.eflag.as<uint8_t>(), thr, step, damp, rho_min,
                        H.rho.as<float>(), H.partials.as<double>(), np, H.s);
        launch_rhop(L0.g, H.rho.as<float>(), H.eflag.as<uint8_t>(), H.p, H.rhop.as<float>(), H.s);
        RT_HIP(hipStreamSynchronize(H.s));
        H.stencil_ready = false;
        if (multiplier) *multiplier = thr;
        if (trials) *trials = ntr;
    })
}

int rt_volume(rt_hier* h, double* v) {
    RT_TRY({
        need(v);
        Hier& H = H_(h);
        launch_sum_active(H.lv[0].g, H.rho.as<float>(), H.eflag.as<uint8_t>(), H.partials.as<double>(), 1024, H.s);
        *v = sum_partials_host(H, 1024) / (double)H.lv[0].n_eact;
    })
}

int rt_voxelize_mesh(const float* coords, int nverts, const int* faces, int nfaces, int prefer_reso, uint32_t* bits,
                     size_t maxwords, int* out_reso, float* out_box) {
    RT_TRY({
        need(coords); need(faces); need(bits); need(out_reso); need(out_box);
        if (nverts <= 0 || nfaces <= 0 || prefer_reso <= 0) throw std::invalid_argument("empty mesh or bad resolution");
        float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
        for (int i = 0; i < nverts; i++)
            for (int j = 0; j < 3; j++) { lo[j] = std::min(lo[j], coords[3 * i + j]); hi[j] = std::max(hi[j], coords[3 * i + j]); }
        float bmin[3], bmax[3], unit[3];
        const int R = voxel_frame(lo, hi, prefer_reso, bmin, bmax, unit);
        const size_t nw = ((size_t)R * R * R + 31) / 32;
        if (maxwords < nw) throw std::invalid_argument("bit buffer too small for R^3 bits");
        std::vector<float> tris((size_t)9 * nfaces);
        for (int t = 0; t < nfaces; t++)
            for (int k = 0; k < 3; k++) {
                const int vi = faces[3 * t + k];
                if (vi < 0 || vi >= nverts) throw std::invalid_argument("face index out of range");
                for (int j = 0; j < 3; j++) tris[(size_t)9 * t + 3 * k + j] = coords[3 * vi + j];
            }
        hipStream_t s;
        RT_HIP(hipStreamCreate(&s));
        DBuf dt, db;
        dt.alloc(sizeof(float) * tris.size());
        db.alloc(sizeof(uint32_t) * nw);
        RT_HIP(hipMemcpyAsync(dt.p, tris.data(), sizeof(float) * tris.size(), hipMemcpyHostToDevice, s));
        launch_voxelize(dt.as<float>(), nfaces, bmin, unit, R, db.as<uint32_t>(), s);
        RT_HIP(hipMemcpyAsync(bits, db.p, sizeof(uint32_t) * nw, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        RT_HIP(hipStreamDestroy(s));
        for (int j = 0; j < 3; j++) { out_reso[j] = R; out_box[j] = bmin[j]; out_box[3 + j] = bmax[j]; }
    })
}

int rt_synchronize(rt_hier* h) { RT_TRY({ RT_HIP(hipStreamSynchronize(H_(h).s)); }) }

}  // extern "C"
template <typename T>
static double time_vcycle(Hier& H, int pre, int post, int n) {
    if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    LevelVecs<T>& pc = pcvecs<T>(H);
    const size_t nn = (size_t)3 * H.lv[0].g.N;
    launch_convert<double, T>(H.f64.data(), pc.F.data(), nn, H.s);
    precondition<T>(H, pc, pre, post);   // eager first application (and graph capture)
    hipEvent_t a, b;
    RT_HIP(hipEventCreate(&a));
    RT_HIP(hipEventCreate(&b));
    RT_HIP(hipEventRecord(a, H.s));
    for (int i = 0; i < n; i++) precondition<T>(H, pc, pre, post);
    RT_HIP(hipEventRecord(b, H.s));
    RT_HIP(hipEventSynchronize(b));
    float ms = 0;
    RT_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / std::max(1, n);
}
extern "C" {

int rt_time_vcycle(rt_hier* h, int bits, int pre, int post, int n, double* ms) {
    RT_TRY({
        need(ms);
        Hier& H = H_(h);
        *ms = bits == 64 ? time_vcycle<double>(H, pre, post, n) : time_vcycle<float>(H, pre, post, n);
    })
}

int rt_time_vcycle_batch(rt_hier* h, int nb, int pre, int post, int n, double* ms) {
    RT_TRY({
        need(ms);
        Hier& H = H_(h);
        if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
        if (!batch_supported(H)) throw std::logic_error("the batched cycle needs gamma 1 and FP32 coarse stencils");
        ensure_batch(H, nb);
        Level& L0 = H.lv[0];
        const size_t nn = (size_t)3 * L0.g.N;
        for (int b = 0; b < nb; b++) launch_convert<double, float>(H.f64.data(), L0.bF.at(b), nn, H.s);
        precondition_batch(H, nb, pre, post);   // eager first application (and graph capture)
        hipEvent_t a, e;
        RT_HIP(hipEventCreate(&a));
        RT_HIP(hipEventCreate(&e));
        RT_HIP(hipEventRecord(a, H.s));
        for (int i = 0; i < n; i++) precondition_batch(H, nb, pre, post);
        RT_HIP(hipEventRecord(e, H.s));
        RT_HIP(hipEventSynchronize(e));
        float t = 0;
        RT_HIP(hipEventElapsedTime(&t, a, e));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(e);
        *ms = t / std::max(1, n);
    })
}

}  // extern "C"
template <typename T>
static double time_sweep(Hier& H, int n) {
    LevelVecs<T>& pc = pcvecs<T>(H);
    const size_t nn = (size_t)3 * H.lv[0].g.N;
    launch_convert<double, T>(H.f64.data(), pc.F.data(), nn, H.s);
    hipEvent_t a, b;
    RT_HIP(hipEventCreate(&a));
    RT_HIP(hipEventCreate(&b));
    RT_HIP(hipEventRecord(a, H.s));
    gs<T>(H, 0, pc.U.v(), pc.F.cv(), n, false, true);   // the preconditioner's smoother (f = 0 at supports)
    RT_HIP(hipEventRecord(b, H.s));
    RT_HIP(hipEventSynchronize(b));
    float ms = 0;
    RT_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / std::max(1, n);
}
extern "C" {

int rt_time_fine_sweep(rt_hier* h, int bits, int n, double* ms) {
    RT_TRY({
        need(ms);
        Hier& H = H_(h);
        *ms = bits == 64 ? time_sweep<double>(H, n) : time_sweep<float>(H, n);
    })
}

}  // extern "C"
template <typename T>
static double time_residual(Hier& H, int n) {
    LevelVecs<T>& pc = pcvecs<T>(H);
    const size_t nn = (size_t)3 * H.lv[0].g.N;
    launch_convert<double, T>(H.f64.data(), pc.F.data(), nn, H.s);
    launch_convert<double, T>(H.x64.data(), pc.U.data(), nn, H.s);
    hipEvent_t a, b;
    RT_HIP(hipEventCreate(&a));
    RT_HIP(hipEventCreate(&b));
    RT_HIP(hipEventRecord(a, H.s));
    for (int i = 0; i < n; i++) residual<T>(H, 0, pc.U.cv(), pc.F.cv(), pc.R.v());
    RT_HIP(hipEventRecord(b, H.s));
    RT_HIP(hipEventSynchronize(b));
    float ms = 0;
    RT_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / std::max(1, n);
}
extern "C" {

int rt_time_fine_residual(rt_hier* h, int bits, int n, double* ms) {
    RT_TRY({
        need(ms);
        Hier& H = H_(h);
        *ms = bits == 64 ? time_residual<double>(H, n) : time_residual<float>(H, n);
    })
}


}  // extern "C"

// ---------------------------------------------------------------------------
// autotuner: times the alternatives of each level's launch shape on this
// hierarchy's own sizes (FP32, the preconditioner's precision) and keeps the
// fastest
// ---------------------------------------------------------------------------
template <typename F>
static double time_ms(Hier& H, int reps, F&& body) {
    hipEvent_t a, b;
    RT_HIP(hipEventCreate(&a));
    RT_HIP(hipEventCreate(&b));
    body();   // warm-up
    RT_HIP(hipEventRecord(a, H.s));
    for (int i = 0; i < reps; i++) body();
    RT_HIP(hipEventRecord(b, H.s));
    RT_HIP(hipEventSynchronize(b));
    float ms = 0;
    RT_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / reps;
}

// info[0..4] = {z planes per fine residual workgroup, ms of residual + apply,
// fine smoother kind, ms per sweep, bit l set: coarse level l uses the sliced
// smoother}, info[5] bit l: sliced coarse residual
static void autotune(Hier& H, int reps, double* info) {
    if (!H.stencil_ready) throw std::logic_error("update_stencil has not been called");
    Level& L0 = H.lv[0];
    LevelVecs<float>& pc = pcvecs<float>(H);
    const size_t nn = (size_t)3 * L0.g.N;
    launch_convert<double, float>(H.f64.data(), pc.F.data(), nn, H.s);
    launch_convert<double, float>(H.x64.data(), pc.U.data(), nn, H.s);
    // finest residual / apply: z planes per workgroup
    double best = 1e30;
    int best_zc = L0.g.zc;
    for (int zc : {8, 12, 16, 24, 32, 48}) {
        L0.g.zc = zc;
        if (fine_apply_partials_modal(L0.g) > H.nparts_max) continue;
        build_tile_mask(L0);
        const double t = time_ms(H, reps, [&] {
            launch_fine_residual_modal<float>(L0.g, pc.U.cv(), pc.F.cv(), pc.R.v(), H.rhop.as<float>(),
                                              L0.vflag.as<uint8_t>(), H.s);
            int np = 0;
            launch_fine_apply_modal<float>(L0.g, pc.U.cv(), pc.R.v(), H.rhop.as<float>(), L0.vflag.as<uint8_t>(),
                                           H.partials.as<double>(), &np, H.s);
        });
        if (t < best) { best = t; best_zc = zc; }
    }
    L0.g.zc = best_zc;
    build_tile_mask(L0);
    info[0] = best_zc;
    info[1] = best;
    // finest smoother implementation
    best = 1e30;
    int best_kind = H.fine_smoother;
    for (int kind : {0, 1}) {
        H.fine_smoother = kind;
        const double t = time_ms(H, std::max(1, reps / 2), [&] { gs<float>(H, 0, pc.U.v(), pc.F.cv(), 1, false, true); });
        if (t < best) { best = t; best_kind = kind; }
    }
    H.fine_smoother = best_kind;
    info[2] = best_kind;
    info[3] = best;
    // coarse levels (not the coarsest): sliced or thread-per-node smoother and residual
    int gs_mask = 0, res_mask = 0;
    for (size_t li = 1; li + 1 < H.lv.size(); li++) {
        Level& L = H.lv[li];
        LevelVecs<float>& v = vecs<float>(L);
        double tb = 1e30;
        int kb = 1;
        for (int k : {1, 0}) {
            L.g.coarse_gs = k;
            const double t = time_ms(H, reps, [&] { gs<float>(H, (int)li, v.U.v(), v.F.cv(), 1, false); });
            if (t < tb) { tb = t; kb = k; }
        }
        L.g.coarse_gs = kb;
        gs_mask |= kb << li;
        tb = 1e30;
        kb = 1;
        for (int k : {1, 0}) {
            L.g.coarse_res = k;
            const double t = time_ms(H, reps, [&] { residual<float>(H, (int)li, v.U.cv(), v.F.cv(), v.R.v()); });
            if (t < tb) { tb = t; kb = k; }
        }
        L.g.coarse_res = kb;
        res_mask |= kb << li;
    }
    info[4] = gs_mask;
    info[5] = res_mask;
    H.drop_graph();
    RT_HIP(hipStreamSynchronize(H.s));
}

extern "C" {

int rt_autotune(rt_hier* h, int reps, double* info) {
    RT_TRY({
        need(info);
        if (reps < 1) throw std::invalid_argument("reps must be positive");
        autotune(H_(h), reps, info);
    })
}
}  // extern "C"
