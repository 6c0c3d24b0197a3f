// Floating-cluster deflation for the MG-PCG preconditioner (no counterpart in
// the reference).
//
// SIMP designs contain clusters of stiff elements (rho >= theta) that reach
// the supported structure only through near-void elements (rho^p ~ 1e-9).
// Their rigid-body motions are near-null modes of K that the geometric
// hierarchy cannot represent, and each costs PCG iterations. The clusters
// ("islands") are found as face-connected components of stiff elements that
// touch no support; each contributes its six rigid-body modes Z_i (three
// translations, three rotations about its centroid, on the nodes of its
// elements) to an additive coarse correction
//     z = M_MG r + sum_i Z_i E_i^+ Z_i^T r,   E_i = Z_i^T K Z_i,
// which keeps the preconditioner symmetric positive (semi)definite. E_i is
// assembled element by element: only elements with some but not all nodes on
// the island contribute (rigid motions are in the null space of KE).
//
// Labels: union-find over the face links of stiff elements (each element
// unites with its -x, -y and -z neighbours, linking the larger root under the
// smaller with atomicMin until the link holds), then every element takes its
// root. Each component ends labelled with its smallest slot, in three passes
// and no host round trips.
#include "rt_internal.h"

namespace rt {

__global__ void k_isl_init(int N, const float* __restrict__ rho, const uint8_t* __restrict__ eflag, float theta,
                           int* __restrict__ lab) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < N; e += gridDim.x * blockDim.x)
        lab[e] = ((eflag[e] & EF_ACTIVE) && rho[e] >= theta) ? e : -1;
}

// labels are read with device-scope atomic loads: other workgroups relink
// roots while this one walks, and a plain load could be served stale from L1
__device__ __forceinline__ int uf_load(int* lab, int x) {
    return __hip_atomic_load(&lab[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// root of x: labels only decrease and lab[x] <= x, so the walk ends
__device__ __forceinline__ int uf_find(int* lab, int x) {
    int y = uf_load(lab, x);
    while (y != x) {
        x = y;
        y = uf_load(lab, x);
    }
    return x;
}

// unite the components of stiff elements a and b: repeatedly find both roots
// and hang the larger under the smaller; atomicMin only lowers a root's label,
// and a failed link (the root was relinked meanwhile) retries from the label
// it returned
__device__ __forceinline__ void uf_unite(int* lab, int a, int b) {
    for (;;) {
        a = uf_find(lab, a);
        b = uf_find(lab, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicMin(&lab[b], a);
        if (old == b) return;
        b = old;
    }
}

__global__ void k_isl_union(Geom g, int* __restrict__ lab) {
    const int ex = g.n[0] - 1, ey = g.n[1] - 1, ez = g.n[2] - 1;
    const int tot = ex * ey * ez;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x) {
        const int x = t % ex, y = (t / ex) % ey, z = t / (ex * ey);
        const int e = oct_p(g, x, y, z);
        if (lab[e] < 0) continue;
        if (x > 0) { const int n = oct_p(g, x - 1, y, z); if (lab[n] >= 0) uf_unite(lab, e, n); }
        if (y > 0) { const int n = oct_p(g, x, y - 1, z); if (lab[n] >= 0) uf_unite(lab, e, n); }
        if (z > 0) { const int n = oct_p(g, x, y, z - 1); if (lab[n] >= 0) uf_unite(lab, e, n); }
    }
}

__global__ void k_isl_flatten(int N, int* __restrict__ lab) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < N; e += gridDim.x * blockDim.x)
        if (lab[e] >= 0) lab[e] = uf_find(lab, e);
}

// elements with a support node anchor their component. Element counts and
// anchor flags are aggregated per wave before the atomics: most lanes of a
// wave share one label (a large component), and one counter hit by every
// element of the box would serialise.
__global__ void k_isl_anchor(Geom g, const int* __restrict__ lab, const uint8_t* __restrict__ vflag,
                             uint8_t* __restrict__ anchored, int* __restrict__ count) {
    const int ex = g.n[0] - 1, ey = g.n[1] - 1, ez = g.n[2] - 1;
    const int tot = ex * ey * ez;
    const int lane = (int)__lane_id();
    for (int t0 = blockIdx.x * blockDim.x; t0 < tot; t0 += gridDim.x * blockDim.x) {
        const int t = t0 + (int)threadIdx.x;
        int l = -1;
        bool fixed = false;
        if (t < tot) {
            const int x = t % ex, y = (t / ex) % ey, z = t / (ex * ey);
            l = lab[oct_p(g, x, y, z)];
            if (l >= 0) {
#pragma unroll
                for (int k = 0; k < 8; k++)
                    fixed |= (vflag[oct_p(g, x + (k & 1), y + ((k >> 1) & 1), z + (k >> 2))] & NF_FIXED) != 0;
            }
        }
        bool todo = l >= 0;
        while (__any(todo)) {
            const int lead = __ffsll((unsigned long long)__ballot(todo)) - 1;
            const int l0 = __shfl(l, lead);
            const bool mine = todo && l == l0;
            const unsigned long long m = __ballot(mine);
            const unsigned long long mf = __ballot(mine && fixed);
            if (lane == lead) {
                atomicAdd(&count[l0], (int)__popcll(m));
                if (mf) anchored[l0] = 1;
            }
            todo = todo && !mine;
        }
    }
}

// roots of floating components with at least min_size elements get island ids
__global__ void k_isl_select(int N, const int* __restrict__ lab, const uint8_t* __restrict__ anchored,
                             const int* __restrict__ count, int min_size, int max_isl, int* __restrict__ root_isl,
                             int* __restrict__ nisl) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < N; e += gridDim.x * blockDim.x) {
        int id = -1;
        if (lab[e] == e && !anchored[e] && count[e] >= min_size) {
            id = atomicAdd(nisl, 1);
            if (id >= max_isl) id = -1;
        }
        root_isl[e] = id;
    }
}

// node -> island: the island of the smallest-numbered island element around the
// node; support nodes and nodes between two islands get none
__global__ void k_isl_nodes(Geom g, const int* __restrict__ lab, const int* __restrict__ root_isl,
                            const uint8_t* __restrict__ vflag, int* __restrict__ node_isl,
                            double* __restrict__ cent) {
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int tot = nx * ny * nz;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x) {
        const int x = t % nx, y = (t / nx) % ny, z = t / (nx * ny);
        const int v = oct_p(g, x, y, z);
        int isl = -1;
        bool clash = false;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int xe = x - 1 + (k & 1), ye = y - 1 + ((k >> 1) & 1), ze = z - 1 + (k >> 2);
            if (xe < 0 || ye < 0 || ze < 0 || xe >= nx - 1 || ye >= ny - 1 || ze >= nz - 1) continue;
            const int l = lab[oct_p(g, xe, ye, ze)];
            if (l < 0) continue;
            const int i = root_isl[l];
            if (i < 0) continue;
            if (isl < 0) isl = i;
            else if (i != isl) clash = true;
        }
        if (clash || (vflag[v] & NF_FIXED)) isl = -1;
        node_isl[v] = isl;
        if (isl >= 0) {
            atomicAdd(&cent[4 * isl + 0], (double)x);
            atomicAdd(&cent[4 * isl + 1], (double)y);
            atomicAdd(&cent[4 * isl + 2], (double)z);
            atomicAdd(&cent[4 * isl + 3], 1.0);
        }
    }
}

__global__ void k_isl_centroid(int nisl, double* __restrict__ cent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nisl) return;
    const double c = cent[4 * i + 3];
    if (c > 0) { cent[4 * i + 0] /= c; cent[4 * i + 1] /= c; cent[4 * i + 2] /= c; }
}

// rigid-body mode m (0-2 translation along m, 3-5 rotation about axis m - 3)
// of a point at offset (dx, dy, dz) from the centroid, component c
__device__ __forceinline__ double rbm(int m, int c, double dx, double dy, double dz) {
    if (m < 3) return m == c ? 1.0 : 0.0;
    const int a = m - 3;
    // e_a x d
    if (a == 0) return c == 1 ? -dz : (c == 2 ? dy : 0.0);
    if (a == 1) return c == 0 ? dz : (c == 2 ? -dx : 0.0);
    return c == 0 ? -dy : (c == 1 ? dx : 0.0);
}

// E_i += rho^p Z_i|e^T KE Z_i|e for every element with some, not all, nodes on
// island i (up to two islands per element)
__global__ void __launch_bounds__(256) k_isl_energy(Geom g, const int* __restrict__ node_isl,
                                                    const float* __restrict__ rhop, const double* __restrict__ KE,
                                                    const double* __restrict__ cent, double* __restrict__ E) {
    const int ex = g.n[0] - 1, ey = g.n[1] - 1, ez = g.n[2] - 1;
    const int tot = ex * ey * ez;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x) {
        const int x = t % ex, y = (t / ex) % ey, z = t / (ex * ey);
        int ni[8];
        int first = -1, second = -1;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            ni[k] = node_isl[oct_p(g, x + (k & 1), y + ((k >> 1) & 1), z + (k >> 2))];
            if (ni[k] >= 0) {
                if (first < 0) first = ni[k];
                else if (ni[k] != first) second = ni[k];
            }
        }
        if (first < 0) continue;
        const double rp = (double)rhop[oct_p(g, x, y, z)];
        if (rp == 0) continue;
        for (int pass = 0; pass < 2; pass++) {
            const int i = pass == 0 ? first : second;
            if (i < 0) continue;
            int on = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) on += ni[k] == i;
            if (on == 8) continue;                    // rigid motion of a whole element: no energy
            const double cx = cent[4 * i + 0], cy = cent[4 * i + 1], cz = cent[4 * i + 2];
            double Zl[24][6];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const double dx = x + (k & 1) - cx, dy = y + ((k >> 1) & 1) - cy, dz = z + (k >> 2) - cz;
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int m = 0; m < 6; m++) Zl[3 * k + c][m] = ni[k] == i ? rbm(m, c, dx, dy, dz) : 0.0;
            }
            // W = KE Zl (24 x 6), then Zl^T W
            double out[36];
#pragma unroll
            for (int q = 0; q < 36; q++) out[q] = 0;
            for (int a = 0; a < 24; a++) {
                double w[6] = {0, 0, 0, 0, 0, 0};
                for (int b = 0; b < 24; b++) {
                    const double kab = KE[a * 24 + b];
#pragma unroll
                    for (int m = 0; m < 6; m++) w[m] += kab * Zl[b][m];
                }
#pragma unroll
                for (int m = 0; m < 6; m++)
#pragma unroll
                    for (int l = 0; l < 6; l++) out[m * 6 + l] += Zl[a][m] * w[l];
            }
#pragma unroll
            for (int q = 0; q < 36; q++)
                if (out[q] != 0) atomicAdd(&E[36 * i + q], rp * out[q]);
        }
    }
}

// E_i -> pseudo-inverse (cyclic Jacobi on the 6 x 6 block, eigenvalues below
// 1e-10 of the largest dropped); one thread per island
__global__ void k_isl_pinv(int nisl, const double* __restrict__ E, double* __restrict__ Einv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nisl) return;
    double A[6][6], V[6][6];
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) {
            A[a][b] = 0.5 * (E[36 * i + 6 * a + b] + E[36 * i + 6 * b + a]);
            V[a][b] = a == b ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 30; sweep++) {
        double off = 0;
        for (int p = 0; p < 6; p++)
            for (int q = p + 1; q < 6; q++) off += A[p][q] * A[p][q];
        if (off < 1e-300) break;
        for (int p = 0; p < 6; p++)
            for (int q = p + 1; q < 6; q++) {
                if (A[p][q] == 0) continue;
                const double th = (A[q][q] - A[p][p]) / (2 * A[p][q]);
                const double tt = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                const double c = 1 / sqrt(tt * tt + 1), s = tt * c;
                for (int k = 0; k < 6; k++) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 6; k++) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 6; k++) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    double wmax = 0;
    for (int k = 0; k < 6; k++) wmax = fmax(wmax, A[k][k]);
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) {
            double s = 0;
            for (int k = 0; k < 6; k++)
                if (wmax > 0 && A[k][k] > 1e-10 * wmax) s += V[a][k] * V[b][k] / A[k][k];
            Einv[36 * i + 6 * a + b] = s;
        }
}

// zr_i += Z_i^T r over the island's nodes
template <typename T>
__global__ void k_isl_project(Geom g, const int* __restrict__ node_isl, const double* __restrict__ cent,
                              Vec3<const T> R, double* __restrict__ zr) {
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int tot = nx * ny * nz;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x) {
        const int x = t % nx, y = (t / nx) % ny, z = t / (nx * ny);
        const int v = oct_p(g, x, y, z);
        const int i = node_isl[v];
        if (i < 0) continue;
        const double r0 = (double)R.c[0][v], r1 = (double)R.c[1][v], r2 = (double)R.c[2][v];
        const double dx = x - cent[4 * i + 0], dy = y - cent[4 * i + 1], dz = z - cent[4 * i + 2];
        atomicAdd(&zr[6 * i + 0], r0);
        atomicAdd(&zr[6 * i + 1], r1);
        atomicAdd(&zr[6 * i + 2], r2);
        atomicAdd(&zr[6 * i + 3], dy * r2 - dz * r1);   // (e_x x d) . r
        atomicAdd(&zr[6 * i + 4], dz * r0 - dx * r2);
        atomicAdd(&zr[6 * i + 5], dx * r1 - dy * r0);
    }
}

// c_i = E_i^+ zr_i, then zr_i = 0 for the next application
__global__ void k_isl_solve(int nisl, const double* __restrict__ Einv, double* __restrict__ zr,
                            double* __restrict__ coef) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nisl) return;
    double b[6];
    for (int m = 0; m < 6; m++) { b[m] = zr[6 * i + m]; zr[6 * i + m] = 0.0; }
    for (int a = 0; a < 6; a++) {
        double s = 0;
        for (int m = 0; m < 6; m++) s += Einv[36 * i + 6 * a + m] * b[m];
        coef[6 * i + a] = s;
    }
}

// z += Z_i c_i on the island's nodes
template <typename T>
__global__ void k_isl_correct(Geom g, const int* __restrict__ node_isl, const double* __restrict__ cent,
                              const double* __restrict__ coef, Vec3<T> Zv) {
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int tot = nx * ny * nz;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += gridDim.x * blockDim.x) {
        const int x = t % nx, y = (t / nx) % ny, z = t / (nx * ny);
        const int v = oct_p(g, x, y, z);
        const int i = node_isl[v];
        if (i < 0) continue;
        const double dx = x - cent[4 * i + 0], dy = y - cent[4 * i + 1], dz = z - cent[4 * i + 2];
        const double* c = coef + 6 * i;
        // translation + rotation (omega x d), omega = (c3, c4, c5)
        Zv.c[0][v] += (T)(c[0] + c[4] * dz - c[5] * dy);
        Zv.c[1][v] += (T)(c[1] + c[5] * dx - c[3] * dz);
        Zv.c[2][v] += (T)(c[2] + c[3] * dy - c[4] * dx);
    }
}

static unsigned grid_for(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 16384); }

int island_build(const Geom& g, const float* rho, const float* rhop, const uint8_t* eflag, const uint8_t* vflag,
                 const double* KE_dev, float theta, int min_size, int max_isl, int* lab, uint8_t* anchored, int* count,
                 int* root_isl, int* node_isl, double* cent, double* E, double* Einv, int* flag, hipStream_t s) {
    const long long ne = (long long)(g.n[0] - 1) * (g.n[1] - 1) * (g.n[2] - 1);
    const long long nn = (long long)g.n[0] * g.n[1] * g.n[2];
    hipLaunchKernelGGL(k_isl_init, dim3(grid_for(g.N)), dim3(256), 0, s, g.N, rho, eflag, theta, lab);
    RT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_isl_union, dim3(grid_for(ne)), dim3(256), 0, s, g, lab);
    RT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_isl_flatten, dim3(grid_for(g.N)), dim3(256), 0, s, g.N, lab);
    RT_HIP(hipGetLastError());
    RT_HIP(hipMemsetAsync(anchored, 0, (size_t)g.N, s));
    RT_HIP(hipMemsetAsync(count, 0, sizeof(int) * (size_t)g.N, s));
    hipLaunchKernelGGL(k_isl_anchor, dim3(grid_for(ne)), dim3(256), 0, s, g, lab, vflag, anchored, count);
    RT_HIP(hipGetLastError());
    RT_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_isl_select, dim3(grid_for(g.N)), dim3(256), 0, s, g.N, lab, anchored, count, min_size,
                       max_isl, root_isl, flag);
    RT_HIP(hipGetLastError());
    int nisl = 0;
    RT_HIP(hipMemcpyAsync(&nisl, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    nisl = std::min(nisl, max_isl);
    RT_HIP(hipMemsetAsync(cent, 0, sizeof(double) * 4 * (size_t)std::max(nisl, 1), s));
    hipLaunchKernelGGL(k_isl_nodes, dim3(grid_for(nn)), dim3(256), 0, s, g, lab, root_isl, vflag, node_isl, cent);
    RT_HIP(hipGetLastError());
    if (nisl == 0) return 0;
    hipLaunchKernelGGL(k_isl_centroid, dim3((nisl + 255) / 256), dim3(256), 0, s, nisl, cent);
    RT_HIP(hipMemsetAsync(E, 0, sizeof(double) * 36 * (size_t)nisl, s));
    hipLaunchKernelGGL(k_isl_energy, dim3(grid_for(ne)), dim3(256), 0, s, g, node_isl, rhop, KE_dev, cent, E);
    RT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_isl_pinv, dim3((nisl + 63) / 64), dim3(64), 0, s, nisl, E, Einv);
    RT_HIP(hipGetLastError());
    RT_HIP(hipStreamSynchronize(s));
    return nisl;
}

template <typename T>
void island_correct(const Geom& g, int nisl, const int* node_isl, const double* cent, const double* Einv, double* zr,
                    double* coef, Vec3<const T> R, Vec3<T> Zv, hipStream_t s) {
    if (nisl <= 0) return;
    const long long nn = (long long)g.n[0] * g.n[1] * g.n[2];
    hipLaunchKernelGGL((k_isl_project<T>), dim3(grid_for(nn)), dim3(256), 0, s, g, node_isl, cent, R, zr);
    hipLaunchKernelGGL(k_isl_solve, dim3((nisl + 255) / 256), dim3(256), 0, s, nisl, Einv, zr, coef);
    hipLaunchKernelGGL((k_isl_correct<T>), dim3(grid_for(nn)), dim3(256), 0, s, g, node_isl, cent, coef, Zv);
    RT_HIP(hipGetLastError());
}

template void island_correct<float>(const Geom&, int, const int*, const double*, const double*, double*, double*,
                                    Vec3<const float>, Vec3<float>, hipStream_t);
template void island_correct<double>(const Geom&, int, const int*, const double*, const double*, double*, double*,
                                     Vec3<const double>, Vec3<double>, hipStream_t);

}  // namespace rt
