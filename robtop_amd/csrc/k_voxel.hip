// Solid (parity) voxelization of a closed triangle mesh, restating
// voxelize_triangle_solid (voxelize_solid.cu:78-150). Built with
// -ffp-contract=off so that every FP32 step rounds as in the CPU restatement:
// the bit pattern depends on the exact edge-function values at voxel centres.
// Bits are LSB-first (bit id x + y R + z R^2), i.e. what the reference holds
// after wordReverse_g (Grid.cu:2662-2687). One thread per triangle; column
// flips use atomicXor so overlapping triangles commute.
#include "rt_internal.h"

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

}  // namespace rt
