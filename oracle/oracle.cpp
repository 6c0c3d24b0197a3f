// ORACLE — test infrastructure only.
//
// An FP64 CPU restatement of robtop's multigrid FEM / SIMP hot path, used as
// the checker for the HIP library. Only the tests, the smoke check and the
// CPU-baseline leg of the benchmark load it; the product never does.
//
// Each function cites the reference file:line it restates. The arithmetic
// follows the reference kernels (FP64 accumulation, rho^p through FP32 powf,
// FP32 filter weights, FP32 OC update). Data are held on a dense box that
// covers every active element of the finest level, x fastest; absent nodes and
// elements are masked out exactly as the reference's "-1" neighbour ids are.
//
// Pinning (independent checks, see tests/test_oracle_*.py):
//   KE            exact 3x3x3 Gauss integration, symmetry, 6 rigid modes
//   levels/masks  numpy OR-pooling, node counts of the reference's rule
//   BitSAT rank   the reference selfTest known answer rank(i) = i/2
//   GS/residual   scipy-assembled global K with the WS rules, point GS
//   transfers     explicit sparse trilinear P
//   Galerkin      dense P^T A P with the Dirichlet-penalised element matrices
//   coarsest      numpy dense solve
//   V-cycle       composition of the above; converged vs scipy spsolve
//   sensitivity   per-element u^T KE u
//   filter / OC   brute-force numpy
//   voxelizer     axis-aligned box meshes (exact answer), numpy restatement
#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>
#include <stdexcept>
#include <string>
#include <omp.h>

namespace {

thread_local std::string g_err;

struct Level {
    int layer = 0;      // reference layer id (0, 2, 3, ...; layer 1 is the dummy)
    int scale = 1;      // fine elements per element edge
    int o[3] = {0, 0, 0};   // element-box origin in this level's absolute element coordinates
    int e[3] = {0, 0, 0};   // element box dims
    int n[3] = {0, 0, 0};   // node box dims = e + 1
    std::vector<uint8_t> eact, vact, vfix;
    std::vector<double> U[3], F[3], R[3];
    std::vector<double> st;     // stencil [27][9][nnodes]
    size_t nn() const { return (size_t)n[0] * n[1] * n[2]; }
    size_t ne() const { return (size_t)e[0] * e[1] * e[2]; }
    size_t vid(int x, int y, int z) const { return (size_t)x + (size_t)n[0] * ((size_t)y + (size_t)n[1] * z); }
    size_t eid(int x, int y, int z) const { return (size_t)x + (size_t)e[0] * ((size_t)y + (size_t)e[1] * z); }
    bool vin(int x, int y, int z) const { return x >= 0 && y >= 0 && z >= 0 && x < n[0] && y < n[1] && z < n[2]; }
    bool ein(int x, int y, int z) const { return x >= 0 && y >= 0 && z >= 0 && x < e[0] && y < e[1] && z < e[2]; }
    bool vexist(int x, int y, int z) const { return vin(x, y, z) && vact[vid(x, y, z)]; }
    bool eexist(int x, int y, int z) const { return ein(x, y, z) && eact[eid(x, y, z)]; }
};

struct Hier {
    int R = 0;
    int nlayers = 0;            // including the dummy layer 1
    std::vector<Level> lv;      // non-dummy levels, finest first
    std::vector<int> active_per_layer;
    double KE[24][24];
    double E = 1e6, nu = 0.4;
    float p = 3.f;
    std::vector<float> rho;     // finest-level element densities (box layout)
    std::vector<float> rhop;    // rho^p per element (FP32 powf, Grid.cu:834), or set directly
    std::vector<uint8_t> shell; // finest-level shell flags (box layout)
    std::vector<uint8_t> load;  // finest-level load-node flags (box layout)
    std::vector<double> nrm[3]; // unit outward normals at load nodes (constrained-direction mode)
    bool normals = false;
    // coarsest dense system
    std::vector<int> lastrow;   // node -> row block, -1 if absent
    int nlastrows = 0;
    std::vector<double> Lfac;   // Cholesky factor (row-major, lower)
    // no supports: orthonormal rigid motions of the coarsest rows (N x 6,
    // row-major) and of the finest nodes (6 box vectors), and of the load
    // nodes (the admissible ones, nrd of them; 6 box vectors)
    std::vector<double> Rc;
    std::vector<double> Rf[6][3];
    std::vector<double> Rd[6][3];
    int nrd = 0;
};

// whether any finest node is a support
static bool has_support(const Hier* h) {
    for (uint8_t f : h->lv[0].vfix) if (f) return true;
    return false;
}

// rigid motions at points p (n x 3): columns e_x, e_y, e_z and the rotations
// (projection.cpp:46-112: [I | phat(p)] per node, phat(p) x = p x (rows
// (0, -p2, p1), (p2, 0, -p0), (-p1, p0, 0))), dof 3 i + c; out: 3n x 6 row-major
static void rigid_columns(const std::vector<double>& p, std::vector<double>& R) {
    const size_t n = p.size() / 3;
    R.assign(3 * n * 6, 0.0);
    for (size_t i = 0; i < n; i++) {
        const double x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
        const double ph[3][3] = {{0, -z, y}, {z, 0, -x}, {-y, x, 0}};
        for (int c = 0; c < 3; c++) {
            R[(3 * i + c) * 6 + c] = 1.0;
            for (int k = 0; k < 3; k++) R[(3 * i + c) * 6 + 3 + k] = ph[c][k];
        }
    }
}

// modified Gram-Schmidt on the 6 columns in order (projection.cpp:96-101, and
// :232-257 dropping a column whose norm falls below 1e-6 of its own or 1e-6);
// returns the number of columns kept, compacted to the front
static int orthonormalize(std::vector<double>& R, size_t rows) {
    bool keep[6];
    for (int i = 0; i < 6; i++) {
        double old = 0;
        for (size_t r = 0; r < rows; r++) old += R[r * 6 + i] * R[r * 6 + i];
        old = std::sqrt(old);
        for (int j = 0; j < i; j++) {
            if (!keep[j]) continue;
            double w = 0;
            for (size_t r = 0; r < rows; r++) w += R[r * 6 + i] * R[r * 6 + j];
            for (size_t r = 0; r < rows; r++) R[r * 6 + i] -= w * R[r * 6 + j];
        }
        double res = 0;
        for (size_t r = 0; r < rows; r++) res += R[r * 6 + i] * R[r * 6 + i];
        res = std::sqrt(res);
        keep[i] = !(res < 1e-6 * old || old < 1e-6);
        if (keep[i])
            for (size_t r = 0; r < rows; r++) R[r * 6 + i] /= res;
    }
    int m = 0;
    for (int i = 0; i < 6; i++) {
        if (!keep[i]) continue;
        if (m != i)
            for (size_t r = 0; r < rows; r++) R[r * 6 + m] = R[r * 6 + i];
        m++;
    }
    for (size_t r = 0; r < rows; r++)
        for (int i = m; i < 6; i++) R[r * 6 + i] = 0.0;
    return m;
}

// ---------------------------------------------------------------------------
// KE: element/templateMatrix.cpp:15-29 (dN) and :31-124 (2x2x2 Gauss, D(E,nu))
// ---------------------------------------------------------------------------
void dN(int i, double elen, const double prm[3], double out[3]) {
    int id[3] = {i % 2, i / 2 % 2, i / 4};
    double r[3];
    for (int k = 0; k < 3; k++) r[k] = id[k] ? prm[k] : (1 - prm[k]);
    for (int k = 0; k < 3; k++) out[k] = (id[k] ? 1.0 / elen : -1.0 / elen) * r[(k + 1) % 3] * r[(k + 2) % 3];
}

void build_ke(double elen, double E, double nu, double KE[24][24]) {
    double D[6][6] = {{0}};
    const double mu = nu;
    double Dv[6][6] = {{1 - mu, mu, mu, 0, 0, 0}, {mu, 1 - mu, mu, 0, 0, 0}, {mu, mu, 1 - mu, 0, 0, 0},
                       {0, 0, 0, (1 - 2 * mu) / 2, 0, 0}, {0, 0, 0, 0, (1 - 2 * mu) / 2, 0}, {0, 0, 0, 0, 0, (1 - 2 * mu) / 2}};
    for (int i = 0; i < 6; i++) for (int j = 0; j < 6; j++) D[i][j] = Dv[i][j] * (E / ((1 + mu) * (1 - 2 * mu)));
    for (int i = 0; i < 24; i++) for (int j = 0; j < 24; j++) KE[i][j] = 0;
    const double pp = std::sqrt(3.0) / 3;
    for (int g = 0; g < 8; g++) {
        int sx = 2 * (g % 2) - 1, sy = 2 * (g / 2 % 2) - 1, sz = 2 * (g / 4) - 1;
        double prm[3] = {(sx * pp + 1) / 2, (sy * pp + 1) / 2, (sz * pp + 1) / 2};
        double gN[8][3];
        for (int k = 0; k < 8; k++) dN(k, elen, prm, gN[k]);
        double B[6][24] = {{0}};
        for (int a = 0; a < 3; a++) for (int b = 0; b < 8; b++) B[a][3 * b + a] = gN[b][a];
        for (int b = 0; b < 8; b++) {
            int off = 3 * b;
            B[3][off + 1] = gN[b][2]; B[3][off + 2] = gN[b][1];   // tau_yz
            B[4][off] = gN[b][2];     B[4][off + 2] = gN[b][0];   // tau_xz
            B[5][off] = gN[b][1];     B[5][off + 1] = gN[b][0];   // tau_xy
        }
        double DB[6][24];
        for (int i = 0; i < 6; i++) for (int j = 0; j < 24; j++) {
            double s = 0; for (int k = 0; k < 6; k++) s += D[i][k] * B[k][j]; DB[i][j] = s;
        }
        for (int i = 0; i < 24; i++) for (int j = 0; j < 24; j++) {
            double s = 0; for (int k = 0; k < 6; k++) s += B[k][i] * DB[k][j]; KE[i][j] += s;
        }
    }
    double sc = std::pow(elen / 2, 3);
    for (int i = 0; i < 24; i++) for (int j = 0; j < 24; j++) KE[i][j] *= sc;
}


inline int colour_of(int x, int y, int z) { return (x & 1) | ((y & 1) << 1) | ((z & 1) << 2); }  // Grid.cu:579

// ---------------------------------------------------------------------------
// Hierarchy from solid bits: Grid.cpp:455-520 (coarsening while > 400 active,
// :229), Grid.cu:2555-2583 (vertex active iff an incident element is),
// Grid.cu:2612-2632 (coarse element active iff a child is). The dense box is
// this oracle's own layout; the active sets are the reference's.
// ---------------------------------------------------------------------------
Hier* create(int R, const uint32_t* bits) {
    if (R < 4 || (R & (R - 1))) throw std::runtime_error("R must be a power of two >= 4");
    std::vector<std::vector<uint8_t>> act;
    std::vector<uint8_t> a0((size_t)R * R * R);
    for (size_t i = 0; i < a0.size(); i++) a0[i] = (bits[i >> 5] >> (i & 31)) & 1u;
    act.push_back(std::move(a0));
    std::vector<int> counts;
    auto count = [](const std::vector<uint8_t>& a) { size_t c = 0; for (auto v : a) c += v; return (int)c; };
    counts.push_back(count(act[0]));
    int reso = R;
    while (counts.back() > 400) {
        int fr = reso; reso >>= 1;
        if (reso < 1) break;
        std::vector<uint8_t> c((size_t)reso * reso * reso, 0);
        for (int z = 0; z < fr; z++) for (int y = 0; y < fr; y++) for (int x = 0; x < fr; x++)
            if (act.back()[x + (size_t)fr * (y + (size_t)fr * z)])
                c[(x >> 1) + (size_t)reso * ((y >> 1) + (size_t)reso * (z >> 1))] = 1;
        act.push_back(std::move(c));
        counts.push_back(count(act.back()));
    }
    int nl = (int)act.size();
    if (nl < 3) throw std::runtime_error("fewer than three layers: the x4 first coarsening needs layers 0 and 2");
    int K = nl - 1;
    int lo[3] = {R, R, R}, hi[3] = {0, 0, 0};
    for (int z = 0; z < R; z++) for (int y = 0; y < R; y++) for (int x = 0; x < R; x++)
        if (act[0][x + (size_t)R * (y + (size_t)R * z)]) {
            int p[3] = {x, y, z};
            for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k] + 1); }
        }
    if (counts[0] == 0) throw std::runtime_error("no active element");
    int A = 1 << K;
    Hier* h = new Hier();
    h->R = R; h->nlayers = nl; h->active_per_layer = counts;
    for (int layer = 0; layer < nl; layer++) {
        if (layer == 1) continue;   // dummy layer (Grid.cpp:390-393)
        Level L;
        L.layer = layer; L.scale = 1 << layer;
        int r = R >> layer;
        for (int k = 0; k < 3; k++) {
            int l0 = (lo[k] / A) * A, h0 = std::min(R, ((hi[k] + A - 1) / A) * A);
            L.o[k] = l0 >> layer; L.e[k] = (h0 - l0) >> layer; L.n[k] = L.e[k] + 1;
        }
        L.eact.assign(L.ne(), 0);
        for (int z = 0; z < L.e[2]; z++) for (int y = 0; y < L.e[1]; y++) for (int x = 0; x < L.e[0]; x++) {
            size_t g = (size_t)(x + L.o[0]) + (size_t)r * ((size_t)(y + L.o[1]) + (size_t)r * (z + L.o[2]));
            L.eact[L.eid(x, y, z)] = act[layer][g];
        }
        L.vact.assign(L.nn(), 0);
        L.vfix.assign(L.nn(), 0);
        for (int z = 0; z < L.n[2]; z++) for (int y = 0; y < L.n[1]; y++) for (int x = 0; x < L.n[0]; x++) {
            bool any = false;
            for (int i = 0; i < 8 && !any; i++) any = L.eexist(x + i % 2 - 1, y + i / 2 % 2 - 1, z + i / 4 - 1);
            L.vact[L.vid(x, y, z)] = any;
        }
        for (int c = 0; c < 3; c++) { L.U[c].assign(L.nn(), 0); L.F[c].assign(L.nn(), 0); L.R[c].assign(L.nn(), 0); }
        if (layer >= 2) L.st.assign(27 * 9 * L.nn(), 0.0);
        h->lv.push_back(std::move(L));
    }
    h->rho.assign(h->lv[0].ne(), 0.f);
    h->rhop.assign(h->lv[0].ne(), 0.f);
    h->shell.assign(h->lv[0].ne(), 0);
    build_ke(1.0, h->E, h->nu, h->KE);
    return h;
}

// ---------------------------------------------------------------------------
// Fine-level nodal Gauss-Seidel, WS mode: Grid.cu:922-1066, driven colour by
// colour as in Grid.cu:1068-1092. `reverse` runs colours 7..0 and components
// z,y,x (the symmetric counterpart used by the PCG preconditioner; the
// reference only has the forward order).
// ---------------------------------------------------------------------------
void fine_gs(Hier* h, int nsweep, bool reverse) {
    Level& L = h->lv[0];
    for (int s = 0; s < nsweep; s++) {
        for (int cc = 0; cc < 8; cc++) {
            int c = reverse ? 7 - cc : cc;
            int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
            // absolute parity: the box origin is a multiple of the coarsest spacing (even)
#pragma omp parallel for collapse(2) schedule(static)
            for (int z = cz; z < L.n[2]; z += 2) for (int y = cy; y < L.n[1]; y += 2) for (int x = cx; x < L.n[0]; x += 2) {
                size_t v = L.vid(x, y, z);
                if (!L.vact[v]) continue;
                bool vfix = L.vfix[v];
                double KeU[3] = {0, 0, 0}, S[9] = {0};
                for (int w = 0; w < 8; w++) {
                    int ex = x + w % 2 - 1, ey = y + w / 2 % 2 - 1, ez = z + w / 4 - 1;
                    if (!L.eexist(ex, ey, ez)) continue;
                    double pen = (double)h->rhop[L.eid(ex, ey, ez)];
                    int vi = 7 - w;
                    double Se[9] = {0};
                    for (int vj = 0; vj < 8; vj++) {
                        int px = vj % 2 + w % 2, py = vj / 2 % 2 + w / 2 % 2, pz = vj / 4 + w / 4;
                        int lid = px + py * 3 + pz * 9;
                        int jx = x + px - 1, jy = y + py - 1, jz = z + pz - 1;
                        size_t j = L.vid(jx, jy, jz);
                        bool jfix = L.vfix[j];
                        if (lid != 13 && !jfix) {
                            double u[3] = {L.U[0][j], L.U[1][j], L.U[2][j]};
                            for (int k = 0; k < 3; k++) for (int q = 0; q < 3; q++) KeU[k] += pen * h->KE[k + vi * 3][q + vj * 3] * u[q];
                        }
                        if (lid == 13) {
                            if (!jfix) for (int i = 0; i < 9; i++) Se[i] = pen * h->KE[vi * 3 + i / 3][vi * 3 + i % 3];
                            else { Se[0] = 1; Se[4] = 1; Se[8] = 1; }
                        }
                    }
                    for (int i = 0; i < 9; i++) S[i] += Se[i];
                }
                if (vfix) { KeU[0] = KeU[1] = KeU[2] = 0; }
                double u[3] = {L.U[0][v], L.U[1][v], L.U[2][v]};
                double f[3] = {L.F[0][v], L.F[1][v], L.F[2][v]};
                if (!reverse) {
                    u[0] = (f[0] - S[1] * u[1] - S[2] * u[2] - KeU[0]) / S[0];
                    u[1] = (f[1] - S[3] * u[0] - S[5] * u[2] - KeU[1]) / S[4];
                    u[2] = (f[2] - S[6] * u[0] - S[7] * u[1] - KeU[2]) / S[8];
                } else {
                    u[2] = (f[2] - S[6] * u[0] - S[7] * u[1] - KeU[2]) / S[8];
                    u[1] = (f[1] - S[3] * u[0] - S[5] * u[2] - KeU[1]) / S[4];
                    u[0] = (f[0] - S[1] * u[1] - S[2] * u[2] - KeU[0]) / S[0];
                }
                L.U[0][v] = u[0]; L.U[1][v] = u[1]; L.U[2][v] = u[2];
            }
        }
    }
}

// Fine residual r = f - K u, WS mode: Grid.cu:1501-1557.
// With apply_only the result is K u (used as the PCG operator check).
void fine_residual(Hier* h, const std::vector<double>* Uin, const std::vector<double>* Fin,
                   std::vector<double>* Rout, bool apply_only) {
    Level& L = h->lv[0];
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < L.n[2]; z++) for (int y = 0; y < L.n[1]; y++) for (int x = 0; x < L.n[0]; x++) {
        size_t v = L.vid(x, y, z);
        if (!L.vact[v]) { for (int c = 0; c < 3; c++) Rout[c][v] = 0; continue; }
        double KU[3] = {0, 0, 0};
        for (int w = 0; w < 8; w++) {
            int ex = x + w % 2 - 1, ey = y + w / 2 % 2 - 1, ez = z + w / 4 - 1;
            if (!L.eexist(ex, ey, ez)) continue;
            double pen = (double)h->rhop[L.eid(ex, ey, ez)];
            int vi = 7 - w;
            for (int vj = 0; vj < 8; vj++) {
                int px = vj % 2 + w % 2, py = vj / 2 % 2 + w / 2 % 2, pz = vj / 4 + w / 4;
                size_t j = L.vid(x + px - 1, y + py - 1, z + pz - 1);
                double u[3] = {Uin[0][j], Uin[1][j], Uin[2][j]};
                if (L.vfix[j]) { u[0] = u[1] = u[2] = 0; }
                for (int r = 0; r < 3; r++) for (int q = 0; q < 3; q++) KU[r] += pen * h->KE[r + vi * 3][q + vj * 3] * u[q];
            }
        }
        if (L.vfix[v]) { KU[0] = KU[1] = KU[2] = 0; }
        for (int c = 0; c < 3; c++) Rout[c][v] = apply_only ? KU[c] : Fin[c][v] - KU[c];
    }
}

// Coarse-level GS with stored stencils: Grid.cu:678-790 (forward order);
// `reverse` as for fine_gs.
void coarse_gs(Hier* h, int li, int nsweep, bool reverse) {
    Level& L = h->lv[li];
    size_t nn = L.nn();
    for (int s = 0; s < nsweep; s++) {
        for (int cc = 0; cc < 8; cc++) {
            int c = reverse ? 7 - cc : cc;
            int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
#pragma omp parallel for collapse(2) schedule(static)
            for (int z = cz; z < L.n[2]; z += 2) for (int y = cy; y < L.n[1]; y += 2) for (int x = cx; x < L.n[0]; x += 2) {
                size_t v = L.vid(x, y, z);
                if (!L.vact[v]) continue;
                double Au[3] = {0, 0, 0};
                for (int nb = 0; nb < 27; nb++) {
                    if (nb == 13) continue;
                    int jx = x + nb % 3 - 1, jy = y + nb / 3 % 3 - 1, jz = z + nb / 9 - 1;
                    if (!L.vexist(jx, jy, jz)) continue;
                    size_t j = L.vid(jx, jy, jz);
                    for (int r = 0; r < 3; r++) for (int q = 0; q < 3; q++)
                        Au[r] += L.st[((size_t)nb * 9 + r * 3 + q) * nn + v] * L.U[q][j];
                }
                auto S = [&](int r, int q) { return L.st[((size_t)13 * 9 + r * 3 + q) * nn + v]; };
                double u[3] = {L.U[0][v], L.U[1][v], L.U[2][v]};
                double f[3] = {L.F[0][v], L.F[1][v], L.F[2][v]};
                if (!reverse) {
                    u[0] = (f[0] - Au[0] - (S(0, 1) * u[1] + S(0, 2) * u[2])) / S(0, 0);
                    u[1] = (f[1] - Au[1] - (S(1, 0) * u[0] + S(1, 2) * u[2])) / S(1, 1);
                    u[2] = (f[2] - Au[2] - (S(2, 0) * u[0] + S(2, 1) * u[1])) / S(2, 2);
                } else {
                    u[2] = (f[2] - Au[2] - (S(2, 0) * u[0] + S(2, 1) * u[1])) / S(2, 2);
                    u[1] = (f[1] - Au[1] - (S(1, 0) * u[0] + S(1, 2) * u[2])) / S(1, 1);
                    u[0] = (f[0] - Au[0] - (S(0, 1) * u[1] + S(0, 2) * u[2])) / S(0, 0);
                }
                L.U[0][v] = u[0]; L.U[1][v] = u[1]; L.U[2][v] = u[2];
            }
        }
    }
}

// Coarse residual: Grid.cu:1349-1448
void coarse_residual(Hier* h, int li) {
    Level& L = h->lv[li];
    size_t nn = L.nn();
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < L.n[2]; z++) for (int y = 0; y < L.n[1]; y++) for (int x = 0; x < L.n[0]; x++) {
        size_t v = L.vid(x, y, z);
        if (!L.vact[v]) { for (int c = 0; c < 3; c++) L.R[c][v] = 0; continue; }
        double Au[3] = {0, 0, 0};
        for (int nb = 0; nb < 27; nb++) {
            int jx = x + nb % 3 - 1, jy = y + nb / 3 % 3 - 1, jz = z + nb / 9 - 1;
            if (!L.vexist(jx, jy, jz)) continue;
            size_t j = L.vid(jx, jy, jz);
            for (int r = 0; r < 3; r++) for (int q = 0; q < 3; q++)
                Au[r] += L.st[((size_t)nb * 9 + r * 3 + q) * nn + v] * L.U[q][j];
        }
        for (int c = 0; c < 3; c++) L.R[c][v] = L.F[c][v] - Au[c];
    }
}

// absolute position of a node of level `L` in units of level `Lf` nodes
inline void node_abs(const Level& L, int x, int y, int z, int out[3]) {
    out[0] = L.o[0] + x; out[1] = L.o[1] + y; out[2] = L.o[2] + z;
}

// Restriction F_c = P^T r_f. x4: Grid.cu:1733-1794 (7^3 window, weight
// (4-|dx|)(4-|dy|)(4-|dz|)/64); x2: Grid.cu:1680-1731 (1, 1/2, 1/4, 1/8).
void restrict_residual(Hier* h, int li) {
    Level& C = h->lv[li];
    Level& Fl = h->lv[li - 1];
    int ratio = C.scale / Fl.scale;   // 4 or 2
    int win = ratio - 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < C.n[2]; z++) for (int y = 0; y < C.n[1]; y++) for (int x = 0; x < C.n[0]; x++) {
        size_t v = C.vid(x, y, z);
        double sum[3] = {0, 0, 0};
        if (C.vact[v]) {
            // coarse node position in fine-box coordinates
            int fx = (C.o[0] + x) * ratio - Fl.o[0], fy = (C.o[1] + y) * ratio - Fl.o[1], fz = (C.o[2] + z) * ratio - Fl.o[2];
            for (int dz = -win; dz <= win; dz++) for (int dy = -win; dy <= win; dy++) for (int dx = -win; dx <= win; dx++) {
                if (!Fl.vexist(fx + dx, fy + dy, fz + dz)) continue;
                double w = (double)((ratio - std::abs(dx)) * (ratio - std::abs(dy)) * (ratio - std::abs(dz))) / (double)(ratio * ratio * ratio);
                size_t j = Fl.vid(fx + dx, fy + dy, fz + dz);
                for (int c = 0; c < 3; c++) sum[c] += w * Fl.R[c][j];
            }
        }
        for (int c = 0; c < 3; c++) C.F[c][v] = sum[c];
    }
}

// Prolongation u_f += P u_c. x4: Grid.cu:1859-1895; x2: Grid.cu:1818-1856.
void prolongate(Hier* h, int li) {
    Level& Fl = h->lv[li];
    Level& C = h->lv[li + 1];
    int ratio = C.scale / Fl.scale;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < Fl.n[2]; z++) for (int y = 0; y < Fl.n[1]; y++) for (int x = 0; x < Fl.n[0]; x++) {
        size_t v = Fl.vid(x, y, z);
        if (!Fl.vact[v]) continue;
        int a[3] = {Fl.o[0] + x, Fl.o[1] + y, Fl.o[2] + z};
        int pos[3], cb[3];
        for (int k = 0; k < 3; k++) { pos[k] = a[k] % ratio; cb[k] = a[k] / ratio - C.o[k]; }
        double c3[3] = {0, 0, 0};
        for (int i = 0; i < 8; i++) {
            int cp[3] = {i % 2 * ratio, i / 2 % 2 * ratio, i / 4 * ratio};
            int wp[3] = {std::abs(cp[0] - pos[0]), std::abs(cp[1] - pos[1]), std::abs(cp[2] - pos[2])};
            if (wp[0] >= ratio || wp[1] >= ratio || wp[2] >= ratio) continue;
            double w = (double)((ratio - wp[0]) * (ratio - wp[1]) * (ratio - wp[2])) / (double)(ratio * ratio * ratio);
            int qx = cb[0] + i % 2, qy = cb[1] + i / 2 % 2, qz = cb[2] + i / 4;
            if (!C.vexist(qx, qy, qz)) continue;
            size_t q = C.vid(qx, qy, qz);
            for (int c = 0; c < 3; c++) c3[c] += w * C.U[c][q];
        }
        for (int c = 0; c < 3; c++) Fl.U[c][v] += c3[c];
    }
}

// Galerkin x4 stencil from rho and KE with the Dirichlet penalty, restating
// restrict_stencil_nondyadic_OTFA_WS_kernel (Grid.cu:392-506) loop by loop:
// 64 odd "centre" vertices x 8 elements x 8 rows x 8 cols x 27 coarse targets.
void galerkin4_loop(Hier* h) {
    Level& C = h->lv[1];
    Level& Fl = h->lv[0];
    size_t nn = C.nn();
    std::fill(C.st.begin(), C.st.end(), 0.0);
    float W[4][4][4];
    for (int t = 0; t < 64; t++) { int i = t % 4, j = t % 16 / 4, k = t / 16; W[k][j][i] = (4 - i) * (4 - j) * (4 - k) / 64.f; }
    const double DIR = 1e6f;   // DIRICHLET_DIAGONAL_WEIGHT, Grid.cu:14
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int z = 0; z < C.n[2]; z++) for (int y = 0; y < C.n[1]; y++) for (int x = 0; x < C.n[0]; x++) {
        size_t vid = C.vid(x, y, z);
        if (!C.vact[vid]) continue;
        int base[3] = {(C.o[0] + x) * 4 - 4 - Fl.o[0], (C.o[1] + y) * 4 - 4 - Fl.o[1], (C.o[2] + z) * 4 - 4 - Fl.o[2]};
        for (int ke = 0; ke < 9; ke++) {
            int row = ke / 3, col = ke % 3;
            double cs[27] = {0};
            for (int i = 0; i < 64; i++) {
                int i2[3] = {(i % 4) * 2 + 1, (i % 16 / 4) * 2 + 1, (i / 16) * 2 + 1};
                if (!Fl.vexist(base[0] + i2[0], base[1] + i2[1], base[2] + i2[2])) continue;
                for (int j = 0; j < 8; j++) {
                    int ep[3] = {i2[0] + j % 2 - 1, i2[1] + j % 4 / 2 - 1, i2[2] + j / 4 - 1};
                    int gx = base[0] + ep[0], gy = base[1] + ep[1], gz = base[2] + ep[2];
                    if (!Fl.eexist(gx, gy, gz)) continue;
                    float rho_p = h->rhop[Fl.eid(gx, gy, gz)];
                    bool vfix[8];
                    for (int k = 0; k < 8; k++) vfix[k] = Fl.vfix[Fl.vid(gx + k % 2, gy + k / 2 % 2, gz + k / 4)];
                    for (int ki = 0; ki < 8; ki++) {
                        int vp[3] = {ep[0] + ki % 2, ep[1] + ki % 4 / 2, ep[2] + ki / 4};
                        int wp[3] = {std::abs(vp[0] - 4), std::abs(vp[1] - 4), std::abs(vp[2] - 4)};
                        if (wp[0] >= 4 || wp[1] >= 4 || wp[2] >= 4) continue;
                        double wi = W[wp[0]][wp[1]][wp[2]];
                        double w_ki = wi * rho_p;
                        for (int kj = 0; kj < 8; kj++) {
                            int kp[3] = {ep[0] + kj % 2, ep[1] + kj % 4 / 2, ep[2] + kj / 4};
                            double wk = w_ki * h->KE[ki * 3 + row][kj * 3 + col];
                            if (vfix[kj] || vfix[ki]) {
                                wk = 0;
                                if (ki == kj && row == col) wk = wi * DIR;
                            }
                            for (int vs = 0; vs < 27; vs++) {
                                int sp[3] = {vs % 3 * 4, vs % 9 / 3 * 4, vs / 9 * 4};
                                int wj[3] = {std::abs(sp[0] - kp[0]), std::abs(sp[1] - kp[1]), std::abs(sp[2] - kp[2])};
                                if (wj[0] >= 4 || wj[1] >= 4 || wj[2] >= 4) continue;
                                cs[vs] += wk * W[wj[0]][wj[1]][wj[2]];
                            }
                        }
                    }
                }
            }
            for (int vs = 0; vs < 27; vs++) C.st[((size_t)vs * 9 + ke) * nn + vid] = cs[vs];
        }
    }
}

// Galerkin x2 from the finer stored stencils: restrict_stencil_dyadic_kernel
// (Grid.cu:120-177).
void galerkin2_loop(Hier* h, int li) {
    Level& C = h->lv[li];
    Level& Fl = h->lv[li - 1];
    size_t nnc = C.nn(), nnf = Fl.nn();
    std::fill(C.st.begin(), C.st.end(), 0.0);
    const double w[4] = {1.0, 1.0 / 2, 1.0 / 4, 1.0 / 8};
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int z = 0; z < C.n[2]; z++) for (int y = 0; y < C.n[1]; y++) for (int x = 0; x < C.n[0]; x++) {
        size_t vid = C.vid(x, y, z);
        if (!C.vact[vid]) continue;
        int fc[3] = {(C.o[0] + x) * 2 - Fl.o[0], (C.o[1] + y) * 2 - Fl.o[1], (C.o[2] + z) * 2 - Fl.o[2]};
        for (int ke = 0; ke < 9; ke++) {
            double cs[27] = {0};
            for (int i = 0; i < 27; i++) {
                int np[3] = {i % 3 + 1, i % 9 / 3 + 1, i / 9 + 1};
                int wn[3] = {std::abs(np[0] - 2), std::abs(np[1] - 2), std::abs(np[2] - 2)};
                double weight = w[wn[0] + wn[1] + wn[2]];
                int vx = fc[0] + np[0] - 2, vy = fc[1] + np[1] - 2, vz = fc[2] + np[2] - 2;
                if (!Fl.vexist(vx, vy, vz)) continue;
                size_t vn = Fl.vid(vx, vy, vz);
                for (int j = 0; j < 27; j++) {
                    if (!Fl.vexist(vx + j % 3 - 1, vy + j % 9 / 3 - 1, vz + j / 9 - 1)) continue;
                    double kij = Fl.st[((size_t)j * 9 + ke) * nnf + vn] * weight;
                    int vjp[3] = {np[0] + j % 3 - 1, np[1] + j % 9 / 3 - 1, np[2] + j / 9 - 1};
                    for (int vs = 0; vs < 27; vs++) {
                        int sp[3] = {vs % 3 * 2, vs % 9 / 3 * 2, vs / 9 * 2};
                        int ws[3] = {std::abs(sp[0] - vjp[0]), std::abs(sp[1] - vjp[1]), std::abs(sp[2] - vjp[2])};
                        if (ws[0] >= 2 || ws[1] >= 2 || ws[2] >= 2) continue;
                        cs[vs] += w[ws[0] + ws[1] + ws[2]] * kij;
                    }
                }
            }
            for (int vs = 0; vs < 27; vs++) C.st[((size_t)vs * 9 + ke) * nnc + vid] = cs[vs];
        }
    }
}

// Coarsest dense system: Grid.cpp:1884-1939 assembles K over existing nodes;
// the reference solves it by BDCSVD least squares (Grid.cpp:1981-2043). With
// supports the system is SPD and any accurate dense solve agrees to rounding;
// this oracle factors it by Cholesky.
void build_coarsest(Hier* h) {
    Level& L = h->lv.back();
    size_t nn = L.nn();
    h->lastrow.assign(nn, -1);
    int rows = 0;
    for (size_t v = 0; v < nn; v++) if (L.vact[v]) h->lastrow[v] = rows++;
    h->nlastrows = rows;
    int N = rows * 3;
    std::vector<double> K((size_t)N * N, 0.0);
    for (int z = 0; z < L.n[2]; z++) for (int y = 0; y < L.n[1]; y++) for (int x = 0; x < L.n[0]; x++) {
        size_t v = L.vid(x, y, z);
        if (!L.vact[v]) continue;
        for (int nb = 0; nb < 27; nb++) {
            int jx = x + nb % 3 - 1, jy = y + nb / 3 % 3 - 1, jz = z + nb / 9 - 1;
            if (!L.vexist(jx, jy, jz)) continue;
            size_t j = L.vid(jx, jy, jz);
            for (int r = 0; r < 3; r++) for (int q = 0; q < 3; q++)
                K[(size_t)(h->lastrow[v] * 3 + r) * N + h->lastrow[j] * 3 + q] = L.st[((size_t)nb * 9 + r * 3 + q) * nn + v];
        }
    }
    // No supports (the no-support modes, projection.cpp:374-411): K has the
    // rigid motions R of the level's nodes as its null space. The reference
    // takes BDCSVD's minimum-norm solution (Grid.cpp:1981-2043); the same
    // solution is P (K + a R R^T)^-1 P b with P = I - R R^T and a > 0 (the mean
    // diagonal), which stays a Cholesky solve.
    h->Rc.clear();
    if (!has_support(h)) {
        std::vector<double> pos;
        for (int z = 0; z < L.n[2]; z++) for (int y = 0; y < L.n[1]; y++) for (int x = 0; x < L.n[0]; x++)
            if (L.vact[L.vid(x, y, z)]) { pos.push_back(x); pos.push_back(y); pos.push_back(z); }
        rigid_columns(pos, h->Rc);
        if (orthonormalize(h->Rc, (size_t)N) != 6) throw std::runtime_error("coarsest level has fewer than 6 rigid motions");
        double a = 0;
        for (int i = 0; i < N; i++) a += K[(size_t)i * N + i];
        a /= N;
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) {
                double s = 0;
                for (int k = 0; k < 6; k++) s += h->Rc[(size_t)i * 6 + k] * h->Rc[(size_t)j * 6 + k];
                K[(size_t)i * N + j] += a * s;
            }
    }
    // Cholesky (lower), in place
    for (int j = 0; j < N; j++) {
        double d = K[(size_t)j * N + j];
        for (int k = 0; k < j; k++) d -= K[(size_t)j * N + k] * K[(size_t)j * N + k];
        if (!(d > 0)) throw std::runtime_error("coarsest system is not positive definite (no supports?)");
        d = std::sqrt(d);
        K[(size_t)j * N + j] = d;
#pragma omp parallel for schedule(static)
        for (int i = j + 1; i < N; i++) {
            double s = K[(size_t)i * N + j];
            for (int k = 0; k < j; k++) s -= K[(size_t)i * N + k] * K[(size_t)j * N + k];
            K[(size_t)i * N + j] = s / d;
        }
    }
    h->Lfac.swap(K);
}

void solve_coarsest(Hier* h) {
    Level& L = h->lv.back();
    int N = h->nlastrows * 3;
    std::vector<double> b(N);
    for (size_t v = 0; v < L.nn(); v++) if (h->lastrow[v] >= 0) for (int c = 0; c < 3; c++) b[h->lastrow[v] * 3 + c] = L.F[c][v];
    const std::vector<double>& K = h->Lfac;
    // b <- P b, x <- P x with P = I - Rc Rc^T (no supports)
    auto project = [&](std::vector<double>& v) {
        if (h->Rc.empty()) return;
        double w[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < N; i++) for (int k = 0; k < 6; k++) w[k] += h->Rc[(size_t)i * 6 + k] * v[i];
        for (int i = 0; i < N; i++) for (int k = 0; k < 6; k++) v[i] -= h->Rc[(size_t)i * 6 + k] * w[k];
    };
    project(b);
    for (int i = 0; i < N; i++) { double s = b[i]; for (int k = 0; k < i; k++) s -= K[(size_t)i * N + k] * b[k]; b[i] = s / K[(size_t)i * N + i]; }
    for (int i = N - 1; i >= 0; i--) { double s = b[i]; for (int k = i + 1; k < N; k++) s -= K[(size_t)k * N + i] * b[k]; b[i] = s / K[(size_t)i * N + i]; }
    project(b);
    for (size_t v = 0; v < L.nn(); v++) for (int c = 0; c < 3; c++) L.U[c][v] = h->lastrow[v] >= 0 ? b[h->lastrow[v] * 3 + c] : 0.0;
}

void update_stencil(Hier* h) {
    galerkin4_loop(h);
    for (size_t li = 2; li < h->lv.size(); li++) galerkin2_loop(h, (int)li);
    build_coarsest(h);
}

void update_residual(Hier* h, int li) {
    if (li == 0) fine_residual(h, h->lv[0].U, h->lv[0].F, h->lv[0].R, false);
    else coarse_residual(h, li);
}

void gs_relax(Hier* h, int li, int n, bool reverse) {
    if (li == 0) fine_gs(h, n, reverse); else coarse_gs(h, li, n, reverse);
}

double v3norm(const std::vector<double>* v) {
    double s = 0;
    for (int c = 0; c < 3; c++) for (double x : v[c]) s += x * x;
    return std::sqrt(s);
}

// V-cycle: Grid.cpp:1068-1113 (dummy layer 1 skipped; coarsest solved directly)
double vcycle(Hier* h, int pre, int post, bool symmetric) {
    int n = (int)h->lv.size();
    for (int i = 0; i < n; i++) {
        if (i > 0) {
            update_residual(h, i - 1);
            restrict_residual(h, i);
            for (int c = 0; c < 3; c++) std::fill(h->lv[i].U[c].begin(), h->lv[i].U[c].end(), 0.0);
        }
        if (i < n - 1) gs_relax(h, i, pre, false);
        else solve_coarsest(h);
    }
    for (int i = n - 2; i >= 0; i--) {
        prolongate(h, i);
        gs_relax(h, i, post, symmetric);
    }
    update_residual(h, 0);
    return v3norm(h->lv[0].R) / v3norm(h->lv[0].F);
}

// forceProject, with-support modes (projection.cu:165-231, projection.cpp:289-360):
// the force is cleared off the load nodes; with normals (wscf) the tangential
// part f - Nd^T Nd f is removed, Nd holding the orthonormal tangent pair of
// each node (projection.cpp:176-197), which leaves (f.n) n
void force_project(Hier* h, std::vector<double>* f) {
    Level& L = h->lv[0];
    for (size_t v = 0; v < L.nn(); v++) {
        if (!h->load[v]) { for (int c = 0; c < 3; c++) f[c][v] = 0; continue; }
        if (h->normals) {
            const double d = f[0][v] * h->nrm[0][v] + f[1][v] * h->nrm[1][v] + f[2][v] * h->nrm[2][v];
            for (int c = 0; c < 3; c++) f[c][v] = d * h->nrm[c][v];
        }
    }
    // no supports (nsff, nscf; projection.cpp:322-335): the load must also be
    // balanced, f -= Rd Rd^T f with Rd the admissible rigid motions of the
    // load nodes
    if (!has_support(h)) {
        double w[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < h->nrd; k++)
            for (int c = 0; c < 3; c++)
                for (size_t v = 0; v < L.nn(); v++) w[k] += h->Rd[k][c][v] * f[c][v];
        for (int k = 0; k < h->nrd; k++)
            for (int c = 0; c < 3; c++)
                for (size_t v = 0; v < L.nn(); v++) f[c][v] -= w[k] * h->Rd[k][c][v];
    }
}

// displacementProject (projection.cpp:374-394): u -= (u . R_k) R_k for the
// orthonormal rigid motions R_k of the finest nodes, one after the other
static void displacement_project(Hier* h, std::vector<double>* u) {
    Level& L = h->lv[0];
    const size_t nn = L.nn();
    if (h->Rf[0][0].empty()) {
        std::vector<double> pos;
        std::vector<size_t> ids;
        for (int z = 0; z < L.n[2]; z++) for (int y = 0; y < L.n[1]; y++) for (int x = 0; x < L.n[0]; x++)
            if (L.vact[L.vid(x, y, z)]) { pos.push_back(x); pos.push_back(y); pos.push_back(z); ids.push_back(L.vid(x, y, z)); }
        std::vector<double> R;
        rigid_columns(pos, R);
        if (orthonormalize(R, 3 * ids.size()) != 6) throw std::runtime_error("fewer than 6 rigid motions");
        for (int k = 0; k < 6; k++)
            for (int c = 0; c < 3; c++) {
                h->Rf[k][c].assign(nn, 0.0);
                for (size_t i = 0; i < ids.size(); i++) h->Rf[k][c][ids[i]] = R[(3 * i + c) * 6 + k];
            }
    }
    for (int k = 0; k < 6; k++) {
        double w = 0;
        for (int c = 0; c < 3; c++) for (size_t v = 0; v < nn; v++) w += u[c][v] * h->Rf[k][c][v];
        for (int c = 0; c < 3; c++) for (size_t v = 0; v < nn; v++) u[c][v] -= w * h->Rf[k][c][v];
    }
}

// modifiedPM, with-support branch (optimization.cpp:123-261) from the force
// already in F: project, unitize, u = 0; then one V(1,1) per step, F <- unit
// projection of U, fch = |fs_new - fs_old| / |fs_old| over the load nodes
// (Grid.cu:2255-2286), until fch <= 1e-4 and rel_res <= 1e-2 or 500 steps.
// Returns the worst compliance F.U; *steps = number of V-cycles.
double modified_pm(Hier* h, int max_itn, int* steps) {
    Level& L = h->lv[0];
    const size_t nn = L.nn();
    auto unitize = [&]() {
        const double n = v3norm(L.F);
        for (int c = 0; c < 3; c++) for (size_t v = 0; v < nn; v++) L.F[c][v] /= n;
    };
    force_project(h, L.F);
    unitize();
    for (int c = 0; c < 3; c++) std::fill(L.U[c].begin(), L.U[c].end(), 0.0);
    std::vector<double> fs[3];
    for (int c = 0; c < 3; c++) fs[c] = L.F[c];
    double fch = 1, rel = 1;
    int itn = 0, nv = 0;
    while (itn++ < max_itn && (fch > 1e-4 || rel > 1e-2)) {
        rel = vcycle(h, 1, 1, false);
        nv++;
        // no supports: remove the rigid motion (optimization.cpp:189-194)
        if (!has_support(h)) displacement_project(h, L.U);
        if (rel > 1e4) throw std::runtime_error("modified power method diverged");
        for (int c = 0; c < 3; c++) L.F[c] = L.U[c];
        force_project(h, L.F);
        unitize();
        double d = 0, o = 0;
        for (int c = 0; c < 3; c++)
            for (size_t v = 0; v < nn; v++) {
                if (!h->load[v]) continue;
                d += (L.F[c][v] - fs[c][v]) * (L.F[c][v] - fs[c][v]);
                o += fs[c][v] * fs[c][v];
            }
        fch = std::sqrt(d) / std::sqrt(o);
        for (int c = 0; c < 3; c++) fs[c] = L.F[c];
    }
    *steps = nv;
    double cw = 0;
    for (int c = 0; c < 3; c++) for (size_t v = 0; v < nn; v++) cw += L.F[c][v] * L.U[c][v];
    return cw;
}

// Sensitivity: optimization.cu:43-107, summed per element
// g_e = -sum_v u_v . (penal * KE[v] u_e), penal = p * powf(rho, p-1) (FP32)
void sensitivity(Hier* h, float* g) {
    Level& L = h->lv[0];
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < L.e[2]; z++) for (int y = 0; y < L.e[1]; y++) for (int x = 0; x < L.e[0]; x++) {
        size_t e = L.eid(x, y, z);
        if (!L.eact[e]) { g[e] = 0; continue; }
        double ue[24];
        for (int k = 0; k < 8; k++) {
            size_t j = L.vid(x + k % 2, y + k / 2 % 2, z + k / 4);
            for (int c = 0; c < 3; c++) ue[k * 3 + c] = L.U[c][j];
        }
        double penal = (double)(h->p * powf(h->rho[e], h->p - 1));
        float acc = 0.f;
        for (int vi = 0; vi < 8; vi++) {   // per-vertex partials, each rounded to FP32 (optimization.cu:104)
            double part = 0;
            for (int r = 0; r < 3; r++) {
                double ku = 0;
                for (int q = 0; q < 24; q++) ku += h->KE[vi * 3 + r][q] * ue[q];
                part -= ue[vi * 3 + r] * (ku * penal);
            }
            acc += (float)part;
        }
        g[e] = acc;
    }
}

// Sensitivity filter: Grid.cu:2323-2409 with the weight of Grid.cu:2420-2423.
void filter(Hier* h, float Rf, const float* gin, float* gout) {
    Level& L = h->lv[0];
    float R2 = Rf * Rf;
    int Ri = (int)(Rf + 0.5f);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < L.e[2]; z++) for (int y = 0; y < L.e[1]; y++) for (int x = 0; x < L.e[0]; x++) {
        size_t e = L.eid(x, y, z);
        if (!L.eact[e]) { gout[e] = 0; continue; }
        float wsum = 0; double gsum = 0;
        for (int dx = -Ri; dx <= Ri; dx++) for (int dy = -Ri; dy <= Ri; dy++) for (int dz = -Ri; dz <= Ri; dz++) {
            int nx = x + dx, ny = y + dy, nz = z + dz;
            float r2 = (float)(dx * dx + dy * dy + dz * dz);
            if (r2 > R2) continue;
            if (!L.eexist(nx, ny, nz)) continue;
            float r = sqrtf(r2 / R2);
            float rr = r * r;
            float w = 1 - 6 * rr + 8 * rr * r - 3 * rr * rr;
            gsum += w * gin[L.eid(nx, ny, nz)];
            wsum += w;
        }
        gout[e] = (float)(gsum / wsum);
    }
}

// OC bisection: optimization.cu:136-226. The volume is averaged over active
// elements (the reference averages over its padded array).
float oc_update(Hier* h, float Vgoal, float step, float damp, float rhomin, const float* g, int* iters_out) {
    Level& L = h->lv[0];
    size_t ne = L.ne();
    std::vector<float> nr(ne);
    size_t nact = 0;
    float gmax = 0;
    for (size_t e = 0; e < ne; e++) if (L.eact[e]) { nact++; gmax = std::max(gmax, std::fabs(g[e])); }
    auto trial = [&](float thr, float* out) {
#pragma omp parallel for schedule(static)
        for (size_t e = 0; e < ne; e++) {
            if (!L.eact[e]) continue;
            float gg = g[e]; if (gg > 0) gg = 0; gg = std::fabs(gg);
            float ro = h->rho[e];
            float rn = ro * powf(gg / thr, damp);
            rn = std::min(std::max(rn, ro - step), ro + step);
            rn = std::min(std::max(rn, rhomin), 1.f);
            if (h->shell[e]) rn = 1;
            out[e] = rn;
        }
    };
    float lo = 0, up = gmax, thr = (lo + up) / 2;
    float V = 2; int itn = 0, steps = 0;
    do {
        thr = (lo + up) / 2;
        trial(thr, nr.data());
        double s = 0;
        for (size_t e = 0; e < ne; e++) if (L.eact[e]) s += nr[e];
        V = (float)(s / (double)nact);
        steps++;
        if (V > Vgoal) lo = thr; else if (V < Vgoal) up = thr;
    } while (std::fabs(V - Vgoal) > 1e-4 && itn++ < 30);
    trial(thr, h->rho.data());
    for (size_t e = 0; e < ne; e++) h->rhop[e] = powf(h->rho[e], h->p);
    if (iters_out) *iters_out = steps;
    return thr;
}

} // namespace

// ---------------------------------------------------------------------------
// Solid voxelizer, restating voxelize_solid.cu:28-150 (Schwarz-Seidel parity
// fill) with FP32 geometry and no FMA contraction, and voxelize_main.cpp:
// 312-404 / util.h:57-76 for the padded cube. Differences, by design:
//  * a triangle with |n.x| < 1e-6 is skipped ("continue"); the reference
//    kernel "return"s, which also drops the rest of that thread's grid-stride
//    triangles (voxelize_solid.cu:98-99) and so depends on the launch shape;
//  * bits are written LSB-first, i.e. after the reference's wordReverse_g.
// ---------------------------------------------------------------------------
namespace vox {
struct v2 { float x, y; };
struct v3 { float x, y, z; };
inline v2 sub2(v2 a, v2 b) { return {a.x - b.x, a.y - b.y}; }
inline v3 sub3(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// check_point_triangle, voxelize_solid.cu:29-50
inline int check_point_triangle(v2 v0, v2 v1, v2 v2_, v2 p) {
    const float fe = 0.000001f;
    v2 PA = sub2(p, v0), PB = sub2(p, v1), PC = sub2(p, v2_);
    float t1 = PA.x * PB.y - PA.y * PB.x;
    if (std::fabs(t1) < fe && PA.x * PB.x <= 0 && PA.y * PB.y <= 0) return 1;
    float t2 = PB.x * PC.y - PB.y * PC.x;
    if (std::fabs(t2) < fe && PB.x * PC.x <= 0 && PB.y * PC.y <= 0) return 2;
    float t3 = PC.x * PA.y - PC.y * PA.x;
    if (std::fabs(t3) < fe && PC.x * PA.x <= 0 && PC.y * PA.y <= 0) return 3;
    if (t1 * t2 > 0 && t1 * t3 > 0) return 0;
    return -1;
}
// get_x_coordinate, voxelize_solid.cu:53-56
inline float get_x(v3 n, v3 v0, v2 p) { return (-(n.y * (p.x - v0.y) + n.z * (p.y - v0.z)) / n.x + v0.x); }
// checkCCW, voxelize_solid.cu:59-68
inline bool ccw(v2 a, v2 b, v2 c) { v2 e0 = sub2(b, a), e1 = sub2(c, a); return (e0.x * e1.y - e1.x * e0.y) > 0; }
// TopLeftEdge, voxelize_solid.cu:71-74
inline bool top_left(v2 a, v2 b) { return (b.y < a.y) || (b.y == a.y && a.x > b.x); }

// one triangle (voxelize_solid.cu:78-150), bits LSB-first
void triangle(const float* tri, const float bmin[3], const float unit[3], int R, uint32_t* bits) {
    v3 v0 = {tri[0] - bmin[0], tri[1] - bmin[1], tri[2] - bmin[2]};
    v3 v1 = {tri[3] - bmin[0], tri[4] - bmin[1], tri[5] - bmin[2]};
    v3 v2_ = {tri[6] - bmin[0], tri[7] - bmin[1], tri[8] - bmin[2]};
    v3 e0 = sub3(v1, v0), e1 = sub3(v2_, v1);
    v3 c = {e0.y * e1.z - e0.z * e1.y, e0.z * e1.x - e0.x * e1.z, e0.x * e1.y - e0.y * e1.x};
    float dd = c.x * c.x + c.y * c.y;
    dd = dd + c.z * c.z;
    float inv = 1.0f / std::sqrt(dd);   // glm::normalize = v * inversesqrt(dot(v, v))
    v3 n = {c.x * inv, c.y * inv, c.z * inv};
    if (std::fabs(n.x) < 0.000001f) return;
    v2 a = {v0.y, v0.z}, b = {v1.y, v1.z}, d = {v2_.y, v2_.z};
    if (!ccw(a, b, d)) { v2 t = b; b = d; d = t; }
    float bmaxx = std::max(a.x, std::max(b.x, d.x)), bmaxy = std::max(a.y, std::max(b.y, d.y));
    float bminx = std::min(a.x, std::min(b.x, d.x)), bminy = std::min(a.y, std::min(b.y, d.y));
    float gmaxx = (float)std::floor(bmaxx / unit[1] - 0.5), gmaxy = (float)std::floor(bmaxy / unit[2] - 0.5);
    float gminx = (float)std::ceil(bminx / unit[1] - 0.5), gminy = (float)std::ceil(bminy / unit[2] - 0.5);
    // clamps to the cube: never bind for a mesh inside its padded cube
    const int y0 = std::max(0, (int)gminx), y1 = std::min(R - 1, (int)gmaxx);
    const int z0 = std::max(0, (int)gminy), z1 = std::min(R - 1, (int)gmaxy);
    for (int y = y0; y <= y1; y++) {
        for (int z = z0; z <= z1; z++) {
            v2 p = {(float)((y + 0.5) * unit[1]), (float)((z + 0.5) * unit[2])};
            int k = check_point_triangle(a, b, d, p);
            if ((k == 1 && top_left(a, b)) || (k == 2 && top_left(b, d)) || (k == 3 && top_left(d, a)) || (k == 0)) {
                int xmax = (int)(get_x(n, v0, p) / unit[0] - 0.5);
                if (xmax > R - 1) xmax = R - 1;
                for (int x = 0; x <= xmax; x++) {
                    size_t id = (size_t)x + (size_t)y * (size_t)R + (size_t)z * (size_t)R * (size_t)R;
                    bits[id >> 5] ^= (1u << (id & 31));
                }
            }
        }
    }
}
} // namespace vox

extern "C" {

const char* or_last_error(void) { return g_err.c_str(); }

void* or_create(int R, const uint32_t* bits) {
    try { return create(R, bits); } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}
void or_destroy(void* h) { delete (Hier*)h; }
int or_num_levels(void* hp) { return (int)((Hier*)hp)->lv.size(); }
int or_num_layers(void* hp) { return ((Hier*)hp)->nlayers; }
int or_layer_active(void* hp, int layer) { return ((Hier*)hp)->active_per_layer[layer]; }
// info: layer, scale, o[3], e[3], n[3]
void or_level_info(void* hp, int li, int* info) {
    Level& L = ((Hier*)hp)->lv[li];
    info[0] = L.layer; info[1] = L.scale;
    for (int k = 0; k < 3; k++) { info[2 + k] = L.o[k]; info[5 + k] = L.e[k]; info[8 + k] = L.n[k]; }
}
void or_get_masks(void* hp, int li, uint8_t* eact, uint8_t* vact) {
    Level& L = ((Hier*)hp)->lv[li];
    std::memcpy(eact, L.eact.data(), L.ne()); std::memcpy(vact, L.vact.data(), L.nn());
}
void or_set_material(void* hp, double E, double nu, float p) {
    Hier* h = (Hier*)hp; h->E = E; h->nu = nu; h->p = p; build_ke(1.0, E, nu, h->KE);
}
void or_ke(double elen, double E, double nu, double* out) {
    double KE[24][24]; build_ke(elen, E, nu, KE); std::memcpy(out, KE, sizeof(KE));
}
void or_set_fixed(void* hp, const uint8_t* vfix) {
    Level& L = ((Hier*)hp)->lv[0];
    for (size_t v = 0; v < L.nn(); v++) L.vfix[v] = vfix[v] && L.vact[v];
}
void or_set_rho(void* hp, const float* rho) {
    Hier* h = (Hier*)hp; std::memcpy(h->rho.data(), rho, h->rho.size() * 4);
    for (size_t e = 0; e < h->rho.size(); e++) h->rhop[e] = powf(h->rho[e], h->p);
}
// Overrides rho^p (for bit-level comparisons that must not depend on powf's last ulp)
void or_set_rhop(void* hp, const float* rhop) { Hier* h = (Hier*)hp; std::memcpy(h->rhop.data(), rhop, h->rhop.size() * 4); }
void or_get_rhop(void* hp, float* rhop) { Hier* h = (Hier*)hp; std::memcpy(rhop, h->rhop.data(), h->rhop.size() * 4); }
void or_get_rho(void* hp, float* rho) { Hier* h = (Hier*)hp; std::memcpy(rho, h->rho.data(), h->rho.size() * 4); }
void or_set_shell(void* hp, const uint8_t* s) { Hier* h = (Hier*)hp; std::memcpy(h->shell.data(), s, h->shell.size()); }
// which: 0 = U, 1 = F, 2 = R; layout [3][nnodes]
void or_set_vec(void* hp, int li, int which, const double* x) {
    Level& L = ((Hier*)hp)->lv[li];
    std::vector<double>* V = which == 0 ? L.U : which == 1 ? L.F : L.R;
    for (int c = 0; c < 3; c++) std::memcpy(V[c].data(), x + c * L.nn(), L.nn() * 8);
}
void or_get_vec(void* hp, int li, int which, double* x) {
    Level& L = ((Hier*)hp)->lv[li];
    std::vector<double>* V = which == 0 ? L.U : which == 1 ? L.F : L.R;
    for (int c = 0; c < 3; c++) std::memcpy(x + c * L.nn(), V[c].data(), L.nn() * 8);
}
void or_get_stencil(void* hp, int li, double* out) {
    Level& L = ((Hier*)hp)->lv[li]; std::memcpy(out, L.st.data(), L.st.size() * 8);
}
void or_set_stencil(void* hp, int li, const double* in) {
    Level& L = ((Hier*)hp)->lv[li]; std::memcpy(L.st.data(), in, L.st.size() * 8);
}
void or_gs_relax(void* hp, int li, int n, int reverse) { gs_relax((Hier*)hp, li, n, reverse != 0); }
void or_update_residual(void* hp, int li) { update_residual((Hier*)hp, li); }
// y = K x on the finest level with the WS rules (fixed rows 0, fixed columns skipped)
void or_apply(void* hp, const double* x, double* y) {
    Hier* h = (Hier*)hp; Level& L = h->lv[0]; size_t nn = L.nn();
    std::vector<double> X[3], Y[3];
    for (int c = 0; c < 3; c++) { X[c].assign(x + c * nn, x + (c + 1) * nn); Y[c].assign(nn, 0); }
    fine_residual(h, X, X, Y, true);
    for (int c = 0; c < 3; c++) std::memcpy(y + c * nn, Y[c].data(), nn * 8);
}
void or_restrict_residual(void* hp, int li) { restrict_residual((Hier*)hp, li); }
void or_prolongate_correction(void* hp, int li) { prolongate((Hier*)hp, li); }
int or_update_stencil(void* hp) {
    try { update_stencil((Hier*)hp); return 0; } catch (std::exception& e) { g_err = e.what(); return -1; }
}
int or_galerkin4(void* hp) { galerkin4_loop((Hier*)hp); return 0; }
int or_galerkin2(void* hp, int li) { galerkin2_loop((Hier*)hp, li); return 0; }
int or_build_coarsest(void* hp) {
    try { build_coarsest((Hier*)hp); return 0; } catch (std::exception& e) { g_err = e.what(); return -1; }
}
void or_solve_coarsest(void* hp) { solve_coarsest((Hier*)hp); }
double or_vcycle(void* hp, int pre, int post, int symmetric) { return vcycle((Hier*)hp, pre, post, symmetric != 0); }
double or_relative_residual(void* hp) { Hier* h = (Hier*)hp; return v3norm(h->lv[0].R) / v3norm(h->lv[0].F); }
double or_compliance(void* hp) {   // Grid.cpp:1976-1979: F . U
    Level& L = ((Hier*)hp)->lv[0]; double s = 0;
    for (int c = 0; c < 3; c++) for (size_t v = 0; v < L.nn(); v++) s += L.F[c][v] * L.U[c][v];
    return s;
}
void or_sensitivity(void* hp, float* g) { sensitivity((Hier*)hp, g); }
// load nodes (flags on the finest node box; supports excluded) and, if given,
// their normals (3 x nn), normalised here
void or_set_load_nodes(void* hp, const uint8_t* load, const double* normals) {
    Hier* h = (Hier*)hp; Level& L = h->lv[0]; const size_t nn = L.nn();
    h->load.assign(nn, 0);
    for (size_t v = 0; v < nn; v++) h->load[v] = (load[v] && L.vact[v] && !L.vfix[v]) ? 1 : 0;
    h->normals = normals != nullptr;
    for (int c = 0; c < 3; c++) h->nrm[c].assign(nn, 0.0);
    if (normals)
        for (size_t v = 0; v < nn; v++) {
            const double a = normals[v], b = normals[nn + v], d = normals[2 * nn + v];
            const double l = std::sqrt(a * a + b * b + d * d);
            if (h->load[v] && l > 0) { h->nrm[0][v] = a / l; h->nrm[1][v] = b / l; h->nrm[2][v] = d / l; }
        }
    // admissible rigid motions of the load nodes (projection.cpp:208-257):
    // with normals only their normal part, then orthonormalised
    std::vector<double> pos;
    std::vector<size_t> ids;
    for (int z = 0; z < L.n[2]; z++) for (int y = 0; y < L.n[1]; y++) for (int x = 0; x < L.n[0]; x++)
        if (h->load[L.vid(x, y, z)]) { pos.push_back(x); pos.push_back(y); pos.push_back(z); ids.push_back(L.vid(x, y, z)); }
    std::vector<double> R;
    rigid_columns(pos, R);
    if (h->normals)
        for (size_t i = 0; i < ids.size(); i++)
            for (int k = 0; k < 6; k++) {
                double d = 0;
                for (int c = 0; c < 3; c++) d += R[(3 * i + c) * 6 + k] * h->nrm[c][ids[i]];
                for (int c = 0; c < 3; c++) R[(3 * i + c) * 6 + k] = d * h->nrm[c][ids[i]];
            }
    h->nrd = ids.empty() ? 0 : orthonormalize(R, 3 * ids.size());
    for (int k = 0; k < 6; k++)
        for (int c = 0; c < 3; c++) {
            h->Rd[k][c].assign(nn, 0.0);
            if (k < h->nrd)
                for (size_t i = 0; i < ids.size(); i++) h->Rd[k][c][ids[i]] = R[(3 * i + c) * 6 + k];
        }
}
void or_force_project(void* hp) { Hier* h = (Hier*)hp; force_project(h, h->lv[0].F); }
double or_modified_pm(void* hp, int max_itn, int* steps) {
    try { return modified_pm((Hier*)hp, max_itn, steps); } catch (std::exception& e) { g_err = e.what(); return NAN; }
}
void or_filter(void* hp, float r, const float* gin, float* gout) { filter((Hier*)hp, r, gin, gout); }
float or_oc_update(void* hp, float Vgoal, float step, float damp, float rhomin, const float* g, int* iters) {
    return oc_update((Hier*)hp, Vgoal, step, damp, rhomin, g, iters);
}
// BitSAT rank (Grid.h:93-135): number of set bits before bit `id`, -1 if unset
int or_bitsat_rank(const uint32_t* words, int nwords, long id) {
    (void)nwords;
    long ent = id >> 5; int mod = (int)(id & 31);
    int acc = 0;
    for (long i = 0; i < ent; i++) acc += __builtin_popcount(words[i]);
    uint32_t w = words[ent];
    if (!((w >> mod) & 1u)) return -1;
    return acc + __builtin_popcount(w & ((1u << mod) - 1u));
}

// voxelize_mesh (voxelize_main.cpp:344-404): padded cube of side R = 2^k >
// prefer_reso (padd2SatifyResolution, :312-340), unit = side / R (util.h:64-67),
// then the parity fill. bits must hold R^3/32 words; out_box = {min[3], max[3]}.
int or_voxelize_mesh(const float* coords, int nverts, const int* faces, int nfaces, int prefer_reso,
                     uint32_t* bits, int maxwords, int* out_reso, float* out_box) {
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = 0; i < nverts; i++) for (int j = 0; j < 3; j++) {
        lo[j] = std::min(lo[j], coords[3 * i + j]); hi[j] = std::max(hi[j], coords[3 * i + j]);
    }
    unsigned padd = 1;
    for (int i = 0; i < 32; i++) { if ((unsigned)prefer_reso < padd) break; padd <<= 1; }
    float len[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    float maxlen = std::max(std::max(len[0], len[1]), len[2]);
    float bmin[3], bmax[3], unit[3];
    for (int j = 0; j < 3; j++) {
        float pl = maxlen * (float)padd / (float)prefer_reso;
        bmin[j] = lo[j] - (pl - len[j]) / 2;
        bmax[j] = hi[j] + (pl - len[j]) / 2;
        unit[j] = (bmax[j] - bmin[j]) / float(padd);
    }
    int R = (int)padd;
    size_t nw = ((size_t)R * R * R + 31) / 32;
    if ((size_t)maxwords < nw) { g_err = "bit buffer too small"; return -1; }
    std::memset(bits, 0, nw * 4);
    for (int t = 0; t < nfaces; t++) {
        float tri[9];
        for (int k = 0; k < 3; k++) for (int j = 0; j < 3; j++) tri[3 * k + j] = coords[3 * faces[3 * t + k] + j];
        vox::triangle(tri, bmin, unit, R, bits);
    }
    for (int j = 0; j < 3; j++) { out_reso[j] = R; out_box[j] = bmin[j]; out_box[3 + j] = bmax[j]; }
    return 0;
}

// The parity fill alone for a given cube (bmin, unit, R), for unit tests.
void or_voxelize_triangles(const float* tris, int nfaces, const float* bmin, const float* unit, int R, uint32_t* bits) {
    size_t nw = ((size_t)R * R * R + 31) / 32;
    std::memset(bits, 0, nw * 4);
    for (int t = 0; t < nfaces; t++) vox::triangle(tris + 9 * t, bmin, unit, R, bits);
}

} // extern "C"
