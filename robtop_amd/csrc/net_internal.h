// Learned warm start: kernels of the 3D convolutional encoder-decoder that
// predicts a displacement field on the first coarse level (x4) from the
// density and load there, and the host-side hooks the solver calls.
//
// Tensors are dense and channel-last ("NDHWC"): position p = x + nx (y + ny z)
// of an nx x ny x nz node box holds C consecutive channels. Activations and
// weights are bfloat16 (stored as uint16_t, round to nearest even); the
// products accumulate in FP32 on the matrix cores (v_mfma_f32_16x16x16_bf16).
// Channel counts are multiples of 16 (the MFMA K and N tile); the input's 4
// channels (density, load x, y, z) and the output's 3 (displacement) are
// padded with zero channels whose weights stay zero.
//
// A 3x3x3 convolution is a zero-padded cross-correlation,
//   Y[p][co] = b[co] + sum_t sum_ci W[t][co][ci] X[p + d_t][ci],
// tap t = (dx+1) + 3 (dy+1) + 9 (dz+1), d_t = (dx, dy, dz), x fastest.
#pragma once
#include "rt_internal.h"

namespace rt {

// epilogues of the convolution kernel
enum ConvEpi {
    EPI_RELU = 0,       // Yb = bf16(relu(acc + b))
    EPI_RELU_SKIP = 1,  // Yb = bf16(relu(acc + b + aux))         aux: bf16 [P][COUT]
    EPI_F32 = 2,        // Yf = acc + b                            (network output)
    EPI_MASK = 3,       // Yb = bf16(acc * (aux > 0))              (backward: ReLU derivative)
    EPI_RAW_F32 = 4,    // Yf = acc                                (backward into a pool / upsample)
};

// Y = conv(X) with W laid out [27][cout][cin]; flip = 1 uses the offsets -d_t
// (the backward-data pass, with W holding each tap's transposed matrix)
void launch_conv(int cin, int cout, int epi, const int dims[3], int flip, const uint16_t* X, const uint16_t* W,
                 const float* bias, const uint16_t* aux, void* Y, hipStream_t s);
// partial weight gradients: part[w][t][cout][cin] = sum over wave w's
// positions of G[p][co] X[p + d_t][ci]; returns the partial count
int launch_conv_wgrad(int cin, int cout, const int dims[3], const uint16_t* X, const uint16_t* G, float* part,
                      hipStream_t s);
size_t conv_wgrad_part_elems(int cin, int cout);
// out[i] (+)= sum_w part[w][i], i < n
void launch_sum_parts(const float* part, int nparts, size_t n, float* out, int accumulate, hipStream_t s);
// bias gradient db[c] += sum_p G[p][c]
void launch_bias_grad(int c, size_t P, const uint16_t* G, float* part, float* db, hipStream_t s);
// 2x2x2 average pool (windows clipped at the upper faces), nearest upsample,
// and their transposes fused with the ReLU derivative
void launch_pool(int c, const int dims[3], const uint16_t* X, uint16_t* Y, hipStream_t s);
void launch_upsample(int c, const int dims_fine[3], const uint16_t* X, uint16_t* Y, hipStream_t s);
// G[q] = bf16((sum over q's children of H) * (act[q] > 0)), q on the coarse box
void launch_upsample_t(int c, const int dims_fine[3], const float* H, const uint16_t* act, uint16_t* G,
                       hipStream_t s);
// G[p] = bf16((H[parent(p)] / children(parent) + skip[p]) * (act[p] > 0)), p on the fine box
void launch_pool_t(int c, const int dims_fine[3], const float* H, const uint16_t* skip, const uint16_t* act,
                   uint16_t* G, hipStream_t s);
// loss sum_p sum_{c<3} (Y[p][c] - tscale T[p][c])^2 (partials), G = its
// gradient times gscale as bf16 [P][16]
void launch_loss(size_t P, const float* Y, const float* T, float tscale, float gscale, uint16_t* G, double* partials,
                 int nparts, hipStream_t s);
// Adam on n FP32 parameters: g *= gscale; m, v moments; w -= lr mhat / (sqrt(vhat) + eps)
void launch_adam(size_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                 float c1, float c2, float gscale, hipStream_t s);
// bfloat16 copies of one convolution's weights: Wf[t][co][ci] and the
// transposed Wb[t][ci][co]
void launch_pack_weights(int cin, int cout, const float* W, uint16_t* Wf, uint16_t* Wb, hipStream_t s);
// host tensors <-> device: X [P][4] floats -> bf16 [P][16]; Y [P][16] -> [P][3]
void launch_pack_input(size_t P, const float* x4, uint16_t* X, hipStream_t s);
void launch_unpack_output(size_t P, const float* Y, float* y3, hipStream_t s);

// features on the x4 level's node box: channel 0 = mean fine density over the
// 8^3 fine elements around the node (absent ones count 0), channels 1-3 = the
// restricted load Fc times fscale
void launch_net_features(const Geom& g0, const int e0[3], const float* rho, const Geom& gc, Vec3<const double> Fc,
                         float fscale, uint16_t* X, hipStream_t s);
// T[p][c] = fine displacement at the fine node under coarse node p; partials of sum T^2
void launch_net_target(const Geom& g0, const Geom& gc, Vec3<const double> U, float* T, double* partials, int nparts,
                       hipStream_t s);
// Uc = prediction (channels 0-2) at free active coarse nodes, 0 elsewhere
void launch_net_to_level(const Geom& gc, const float* Y, const uint8_t* vflag_c, Vec3<double> Uc, hipStream_t s);
// max |Fc| over the coarse level (as the bits of a non-negative float)
void launch_absmax3(const Geom& gc, Vec3<const double> Fc, unsigned* out, hipStream_t s);

// ---------------------------------------------------------------------------
// host hooks (rt_net.cpp)
// ---------------------------------------------------------------------------
struct Net;
// predicts the coarse displacement for the hierarchy's current density and
// load (Fc: load restricted to the x4 level) into Uc
void net_predict(Net& net, const Geom& g0, const int e0[3], const float* rho, const Geom& gc,
                 Vec3<const double> Fc, const uint8_t* vflag_c, Vec3<double> Uc, hipStream_t s);
// stores (features, target = U under the coarse nodes) as a training sample
void net_add_sample(Net& net, const Geom& g0, const int e0[3], const float* rho, const Geom& gc,
                    Vec3<const double> Fc, Vec3<const double> U, hipStream_t s);
// the loss of the prediction for such a sample (not stored)
double net_evaluate(Net& net, const Geom& g0, const int e0[3], const float* rho, const Geom& gc,
                    Vec3<const double> Fc, Vec3<const double> U, hipStream_t s);

}  // namespace rt

namespace rt {

// ---------------------------------------------------------------------------
// the network: 3-level encoder-decoder (16 / 32 / 64 channels), 3x3x3 convs,
// ReLU, 2x2x2 average pooling down, nearest upsampling up, additive skips
// ---------------------------------------------------------------------------
enum NetLayer { E0A, E0B, E1A, E1B, E2A, E2B, D1A, D1B, D0A, D0B, HEAD, NCONV };

struct NetSample {
    DBuf X, T;           // features bf16 [P][16], raw target FP32 [P][3]
    int dims[3];
    float tscale;        // 1 / rms(target)
};

struct Net {
    hipStream_t s = nullptr;
    int cin[NCONV], cout[NCONV];
    size_t woff[NCONV], boff[NCONV], nparam = 0;
    DBuf W, G, M, V;                        // FP32 parameters, gradients, Adam moments
    DBuf Wf[NCONV], Wb[NCONV];              // bf16 forward / transposed copies
    int d[3][3] = {};                       // node boxes of the three resolutions
    size_t cap[3] = {0, 0, 0};              // positions the activations are sized for
    DBuf F, a0, s0, q1, a1, s1, q2, a2, b2, u1, c1, t1, u0, c0, t0, Y;
    DBuf gY, gt0, gc0, hu0, gt1, gc1, hu1, gb2, ga2, hq2, gs1, ga1, hq1, gs0, ga0;
    DBuf wpart, bpart, lpart, T, absmax;
    std::vector<NetSample*> samples;
    int adam_t = 0;
    ~Net();
};

}  // namespace rt

struct rt_net { rt::Net* n; };
