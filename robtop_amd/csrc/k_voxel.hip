// Solid (parity) voxelization of a closed triangle mesh, restating
// voxelize_triangle_solid (voxelize_solid.cu:78-150). Built with
// -ffp-contract=off so that every FP32 step rounds as in the CPU restatement:
// the bit pattern depends on the exact edge-function values at voxel centres.
// Bits are LSB-first (bit id x + y R + z R^2), i.e. what the reference holds
// after wordReverse_g (Grid.cu:2662-2687). One thread per triangle; column
// flips use atomicXor so overlapping triangles commute.
#include "rt_internal.h"
#include "robtop_amd.h"

#include <cstring>
#include <mutex>
#include <string>

namespace rt {

struct f2 { float x, y; };

__device__ __forceinline__ int check_point_triangle(f2 v0, f2 v1, f2 v2, f2 p) {
    const float fe = 0.000001f;
    const f2 PA = {p.x - v0.x, p.y - v0.y}, PB = {p.x - v1.x, p.y - v1.y}, PC = {p.x - v2.x, p.y - v2.y};
    const float t1 = PA.x * PB.y - PA.y * PB.x;
    if (fabsf(t1) < fe && PA.x * PB.x <= 0 && PA.y * PB.y <= 0) return 1;
    const float t2 = PB.x * PC.y - PB.y * PC.x;
    if (fabsf(t2) < fe && PB.x * PC.x <= 0 && PB.y * PC.y <= 0) return 2;
    const float t3 = PC.x * PA.y - PC.y * PA.x;
    if (fabsf(t3) < fe && PC.x * PA.x <= 0 && PC.y * PA.y <= 0) return 3;
    if (t1 * t2 > 0 && t1 * t3 > 0) return 0;
    return -1;
}

__device__ __forceinline__ bool top_left(f2 a, f2 b) { return (b.y < a.y) || (b.y == a.y && a.x > b.x); }

__global__ void __launch_bounds__(64) k_voxelize(const float* __restrict__ tris, int ntris, float bx, float by,
                                                 float bz, float ux, float uy, float uz, int R,
                                                 unsigned int* __restrict__ bits) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntris) return;
    const float* tr = tris + 9 * (size_t)t;
    const float v0x = tr[0] - bx, v0y = tr[1] - by, v0z = tr[2] - bz;
    const float v1x = tr[3] - bx, v1y = tr[4] - by, v1z = tr[5] - bz;
    const float v2x = tr[6] - bx, v2y = tr[7] - by, v2z = tr[8] - bz;
    const float e0x = v1x - v0x, e0y = v1y - v0y, e0z = v1z - v0z;
    const float e1x = v2x - v1x, e1y = v2y - v1y, e1z = v2z - v1z;
    const float cx = e0y * e1z - e0z * e1y, cy = e0z * e1x - e0x * e1z, cz = e0x * e1y - e0y * e1x;
    float dd = cx * cx + cy * cy;
    dd = dd + cz * cz;
    const float inv = 1.0f / sqrtf(dd);
    const float nx = cx * inv, ny = cy * inv, nz = cz * inv;
    if (fabsf(nx) < 0.000001f) return;
    f2 a = {v0y, v0z}, b = {v1y, v1z}, d = {v2y, v2z};
    {   // make the yz projection counter-clockwise (voxelize_solid.cu:106-112)
        const f2 e0 = {b.x - a.x, b.y - a.y}, e1 = {d.x - a.x, d.y - a.y};
        if (!((e0.x * e1.y - e1.x * e0.y) > 0)) { f2 tt = b; b = d; d = tt; }
    }
    const float bmaxx = fmaxf(a.x, fmaxf(b.x, d.x)), bmaxy = fmaxf(a.y, fmaxf(b.y, d.y));
    const float bminx = fminf(a.x, fminf(b.x, d.x)), bminy = fminf(a.y, fminf(b.y, d.y));
    const int gmaxx = (int)(float)floor((double)(bmaxx / uy) - 0.5), gmaxy = (int)(float)floor((double)(bmaxy / uz) - 0.5);
    const int gminx = (int)(float)ceil((double)(bminx / uy) - 0.5), gminy = (int)(float)ceil((double)(bminy / uz) - 0.5);
    const size_t RR = (size_t)R;
    // The padded cube contains the mesh, so these clamps never bind for the
    // reference's inputs; they keep out-of-cube geometry from writing outside
    // the bit table.
    const int y0 = gminx < 0 ? 0 : gminx, y1 = gmaxx > R - 1 ? R - 1 : gmaxx;
    const int z0 = gminy < 0 ? 0 : gminy, z1 = gmaxy > R - 1 ? R - 1 : gmaxy;
    for (int y = y0; y <= y1; y++) {
        for (int z = z0; z <= z1; z++) {
            const f2 p = {(float)((y + 0.5) * (double)uy), (float)((z + 0.5) * (double)uz)};
            const int k = check_point_triangle(a, b, d, p);
            if ((k == 1 && top_left(a, b)) || (k == 2 && top_left(b, d)) || (k == 3 && top_left(d, a)) || (k == 0)) {
                const float xh = (-(ny * (p.x - v0y) + nz * (p.y - v0z)) / nx + v0x);
                int xmax = (int)((double)(xh / ux) - 0.5);
                if (xmax > R - 1) xmax = R - 1;
                for (int x = 0; x <= xmax; x++) {
                    const size_t id = (size_t)x + (size_t)y * RR + (size_t)z * RR * RR;
                    atomicXor(bits + (id >> 5), 1u << (id & 31));
                }
            }
        }
    }
}

void launch_voxelize(const float* tris, int ntris, const float bmin[3], const float unit[3], int R, uint32_t* bits,
                     hipStream_t s) {
    if (ntris <= 0) return;
    hipLaunchKernelGGL(k_voxelize, dim3((ntris + 63) / 64), dim3(64), 0, s, tris, ntris, bmin[0], bmin[1], bmin[2],
                       unit[0], unit[1], unit[2], R, (unsigned int*)bits);
    RT_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// rt_voxelize_mesh_noalloc: rt_voxelize_mesh on device arrays for callers that
// cannot allocate device memory. Its scratch (the staged triangles, the
// bounding-box partials and an index-error flag) lives in a file-scope device
// array of fixed size; each device holds its own copy, found once per device
// with hipGetSymbolAddress and used only through that address.
// ---------------------------------------------------------------------------
constexpr int kVoxScratchTris = 1 << 19;                 // triangles the scratch stages
constexpr int kVoxBoxBlocks = 256;                       // bounding-box partials
constexpr int kVoxHeader = 2048;                         // floats before the triangles
constexpr size_t kVoxScratchFloats = (size_t)kVoxHeader + (size_t)9 * kVoxScratchTris;
static_assert(6 * kVoxBoxBlocks + 8 <= kVoxHeader, "scratch header too small");

static __device__ __align__(256) float scratch[kVoxScratchFloats];

// block b: min and max over its grid-stride share of the vertices
__global__ void __launch_bounds__(256) k_vox_box_partials(const float* __restrict__ coords, int nverts,
                                                          float* __restrict__ part) {
    __shared__ float red[6][256];
    float v[6] = {3.0e38f, 3.0e38f, 3.0e38f, -3.0e38f, -3.0e38f, -3.0e38f};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nverts; i += gridDim.x * blockDim.x)
        for (int j = 0; j < 3; j++) {
            const float c = coords[3 * (size_t)i + j];
            v[j] = fminf(v[j], c);
            v[3 + j] = fmaxf(v[3 + j], c);
        }
    for (int j = 0; j < 6; j++) red[j][threadIdx.x] = v[j];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int j = 0; j < 6; j++)
                red[j][threadIdx.x] = j < 3 ? fminf(red[j][threadIdx.x], red[j][threadIdx.x + w])
                                            : fmaxf(red[j][threadIdx.x], red[j][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// one block: the box from the partials; clears the index-error flag
__global__ void __launch_bounds__(256) k_vox_box_final(const float* __restrict__ part, float* __restrict__ box,
                                                       int* __restrict__ flag) {
    __shared__ float red[6][256];
    for (int j = 0; j < 6; j++) red[j][threadIdx.x] = part[6 * threadIdx.x + j];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int j = 0; j < 6; j++)
                red[j][threadIdx.x] = j < 3 ? fminf(red[j][threadIdx.x], red[j][threadIdx.x + w])
                                            : fmaxf(red[j][threadIdx.x], red[j][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 6) box[threadIdx.x] = red[threadIdx.x][0];
    if (threadIdx.x == 0) *flag = 0;
}

// tris[t] = the three vertices of face t; an index outside [0, nverts) sets
// the flag and stages zeros
__global__ void k_vox_gather(const float* __restrict__ coords, int nverts, const int* __restrict__ faces, int nfaces,
                             float* __restrict__ tris, int* __restrict__ flag) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nfaces; t += gridDim.x * blockDim.x)
        for (int k = 0; k < 3; k++) {
            const int vi = faces[3 * (size_t)t + k];
            const bool ok = vi >= 0 && vi < nverts;
            if (!ok) atomicOr(flag, 1);
            for (int j = 0; j < 3; j++) tris[(size_t)9 * t + 3 * k + j] = ok ? coords[3 * (size_t)vi + j] : 0.f;
        }
}

__global__ void k_vox_clear(uint32_t* __restrict__ bits, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        bits[i] = 0u;
}

// the scratch's device address on each device, looked up once per device
static float* vox_scratch_ptr() {
    static std::mutex mu;
    static float* ptr[256] = {};
    int dev = 0;
    RT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 256) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        RT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[dev] = (float*)p;
    }
    return ptr[dev];
}

}  // namespace rt

extern "C" int rt_voxelize_mesh_noalloc(const float* d_coords, int nverts, const int* d_faces, int nfaces,
                                        int prefer_reso, uint32_t* d_bits, size_t maxwords, int* out_reso,
                                        float* out_box) {
    using namespace rt;
    try {
        if (!d_coords || !d_faces || !d_bits || !out_reso || !out_box) throw std::invalid_argument("null pointer");
        if (nverts <= 0 || nfaces <= 0 || prefer_reso <= 0) throw std::invalid_argument("empty mesh or bad resolution");
        if (nfaces > kVoxScratchTris)
            throw std::invalid_argument("rt_voxelize_mesh_noalloc: " + std::to_string(nfaces) +
                                        " faces exceed the scratch limit of " + std::to_string(kVoxScratchTris));
        int R = 1;
        while (R <= prefer_reso && R < (1 << 30)) R <<= 1;
        const size_t nw = ((size_t)R * R * R + 31) / 32;
        if (maxwords < nw) throw std::invalid_argument("bit buffer too small for R^3 bits");
        float* sc = vox_scratch_ptr();
        float* part = sc;
        float* box = sc + 6 * kVoxBoxBlocks;
        int* flag = (int*)(box + 6);
        float* tris = sc + kVoxHeader;
        hipStream_t s = nullptr;
        hipLaunchKernelGGL(k_vox_box_partials, dim3(kVoxBoxBlocks), dim3(256), 0, s, d_coords, nverts, part);
        RT_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_vox_box_final, dim3(1), dim3(256), 0, s, part, box, flag);
        RT_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_vox_gather, dim3(std::min(4096, (nfaces + 255) / 256)), dim3(256), 0, s, d_coords, nverts,
                           d_faces, nfaces, tris, flag);
        RT_HIP(hipGetLastError());
        float hb[7];
        RT_HIP(hipMemcpy(hb, box, sizeof(float) * 7, hipMemcpyDeviceToHost));
        int bad = 0;
        std::memcpy(&bad, hb + 6, sizeof(int));
        if (bad) throw std::invalid_argument("face index out of range");
        float bmin[3], bmax[3], unit[3];
        const int Rf = voxel_frame(hb, hb + 3, prefer_reso, bmin, bmax, unit);
        if (Rf != R) throw std::logic_error("cube side mismatch");
        hipLaunchKernelGGL(k_vox_clear, dim3((unsigned)std::min<size_t>(4096, (nw + 255) / 256)), dim3(256), 0, s, d_bits,
                           nw);
        RT_HIP(hipGetLastError());
        launch_voxelize(tris, nfaces, bmin, unit, R, d_bits, s);
        RT_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < 3; j++) { out_reso[j] = R; out_box[j] = bmin[j]; out_box[3 + j] = bmax[j]; }
        return RT_OK;
    } catch (...) {
        return status_of_current_exception();
    }
}
