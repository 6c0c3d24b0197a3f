// Host side of the learned warm start: the network's parameters and
// activations, its forward and backward passes over the kernels of k_net.hip,
// Adam training over a store of (features, displacement) samples that solves
// add, and the rt_net_* entry points of include/robtop_amd.h. The hierarchy
// side (feature extraction from a solved hierarchy, and the warm start itself)
// is in rt_host.cpp.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <random>

#include "../../include/robtop_amd.h"
#include "net_internal.h"

namespace rt {

static const int kCinReal[NCONV] = {4, 16, 16, 32, 32, 64, 64, 32, 32, 16, 16};
static const int kCoutReal[NCONV] = {16, 16, 32, 32, 64, 64, 32, 32, 16, 16, 3};
static const int kCin[NCONV] = {16, 16, 16, 32, 32, 64, 64, 32, 32, 16, 16};
static const int kCout[NCONV] = {16, 16, 32, 32, 64, 64, 32, 32, 16, 16, 16};
constexpr int kLossParts = 256;

Net::~Net() {
    for (NetSample* p : samples) delete p;
    if (s) (void)hipStreamDestroy(s);
}

// He-normal weights over the real fan-in, zero biases, zero padding channels
static void init_params(Net& N, uint64_t seed, std::vector<float>& w) {
    std::mt19937_64 gen(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    w.assign(N.nparam, 0.f);
    for (int l = 0; l < NCONV; l++) {
        const double sd = std::sqrt((l == HEAD ? 1.0 : 2.0) / (27.0 * kCinReal[l]));
        for (int t = 0; t < 27; t++)
            for (int co = 0; co < N.cout[l]; co++)
                for (int ci = 0; ci < N.cin[l]; ci++) {
                    const double v = nd(gen);
                    if (co < kCoutReal[l] && ci < kCinReal[l])
                        w[N.woff[l] + ((size_t)t * N.cout[l] + co) * N.cin[l] + ci] = (float)(sd * v);
                }
    }
}

static void pack_all(Net& N, hipStream_t s) {
    for (int l = 0; l < NCONV; l++)
        launch_pack_weights(N.cin[l], N.cout[l], N.W.as<float>() + N.woff[l], N.Wf[l].as<uint16_t>(),
                            N.Wb[l].as<uint16_t>(), s);
}

static Net* net_new(uint64_t seed) {
    std::unique_ptr<Net> N(new Net());
    RT_HIP(hipStreamCreate(&N->s));
    size_t off = 0;
    for (int l = 0; l < NCONV; l++) {
        N->cin[l] = kCin[l];
        N->cout[l] = kCout[l];
        N->woff[l] = off;
        off += (size_t)27 * kCin[l] * kCout[l];
        N->boff[l] = off;
        off += kCout[l];
    }
    N->nparam = off;
    N->W.alloc(sizeof(float) * off);
    N->G.alloc(sizeof(float) * off);
    N->M.alloc(sizeof(float) * off);
    N->V.alloc(sizeof(float) * off);
    size_t wp = 0;
    for (int l = 0; l < NCONV; l++) {
        N->Wf[l].alloc(sizeof(uint16_t) * 27 * kCin[l] * kCout[l]);
        N->Wb[l].alloc(sizeof(uint16_t) * 27 * kCin[l] * kCout[l]);
        wp = std::max(wp, conv_wgrad_part_elems(kCin[l], kCout[l]));
    }
    N->wpart.alloc(sizeof(float) * wp);
    N->bpart.alloc(sizeof(float) * 64 * 64);
    N->lpart.alloc(sizeof(double) * kLossParts);
    N->absmax.alloc(sizeof(unsigned));
    std::vector<float> w;
    init_params(*N, seed, w);
    RT_HIP(hipMemcpy(N->W.p, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
    pack_all(*N, N->s);
    RT_HIP(hipStreamSynchronize(N->s));
    return N.release();
}

// sizes every activation for a level-0 box d0 (grow only)
static void shape(Net& N, const int d0[3]) {
    for (int k = 0; k < 3; k++) {
        if (d0[k] < 1) throw std::invalid_argument("network box must be at least 1 node per axis");
        N.d[0][k] = d0[k];
        N.d[1][k] = (d0[k] + 1) / 2;
        N.d[2][k] = (N.d[1][k] + 1) / 2;
    }
    const size_t P0 = (size_t)d0[0] * d0[1] * d0[2];
    const size_t P1 = (size_t)N.d[1][0] * N.d[1][1] * N.d[1][2];
    const size_t P2 = (size_t)N.d[2][0] * N.d[2][1] * N.d[2][2];
    if (P0 > N.cap[0] || P1 > N.cap[1] || P2 > N.cap[2]) {
        // buffers are about to be reallocated: no queued work may still use them
        RT_HIP(hipDeviceSynchronize());
        N.cap[0] = std::max(N.cap[0], P0);
        N.cap[1] = std::max(N.cap[1], P1);
        N.cap[2] = std::max(N.cap[2], P2);
    }
    auto bf = [](DBuf& b, size_t P, int c) { b.ensure(sizeof(uint16_t) * P * c); };
    auto f32 = [](DBuf& b, size_t P, int c) { b.ensure(sizeof(float) * P * c); };
    bf(N.F, P0, 16); bf(N.a0, P0, 16); bf(N.s0, P0, 16); bf(N.q1, P1, 16); bf(N.a1, P1, 32); bf(N.s1, P1, 32);
    bf(N.q2, P2, 32); bf(N.a2, P2, 64); bf(N.b2, P2, 64); bf(N.u1, P1, 64); bf(N.c1, P1, 32); bf(N.t1, P1, 32);
    bf(N.u0, P0, 32); bf(N.c0, P0, 16); bf(N.t0, P0, 16); f32(N.Y, P0, 16);
    bf(N.gY, P0, 16); bf(N.gt0, P0, 16); bf(N.gc0, P0, 16); f32(N.hu0, P0, 32); bf(N.gt1, P1, 32);
    bf(N.gc1, P1, 32); f32(N.hu1, P1, 64); bf(N.gb2, P2, 64); bf(N.ga2, P2, 64); f32(N.hq2, P2, 32);
    bf(N.gs1, P1, 32); bf(N.ga1, P1, 32); f32(N.hq1, P1, 16); bf(N.gs0, P0, 16); bf(N.ga0, P0, 16);
}

static size_t P_of(const int d[3]) { return (size_t)d[0] * d[1] * d[2]; }
template <class T> static T* B(const DBuf& b) { return b.as<T>(); }
static const uint16_t* cb(const DBuf& b) { return b.as<uint16_t>(); }
static uint16_t* mb(const DBuf& b) { return b.as<uint16_t>(); }

static void conv_f(Net& N, int l, int epi, int lev, const DBuf& X, const DBuf* aux, DBuf& Y, hipStream_t s) {
    launch_conv(N.cin[l], N.cout[l], epi, N.d[lev], 0, cb(X), cb(N.Wf[l]), B<float>(N.W) + N.boff[l],
                aux ? cb(*aux) : nullptr, Y.p, s);
}
// backward-data: input gradient of layer l from its output gradient G
static void conv_b(Net& N, int l, int epi, int lev, const DBuf& G, const DBuf* aux, DBuf& out, hipStream_t s) {
    launch_conv(N.cout[l], N.cin[l], epi, N.d[lev], 1, cb(G), cb(N.Wb[l]), nullptr, aux ? cb(*aux) : nullptr,
                out.p, s);
}
static void grads(Net& N, int l, int lev, const DBuf& X, const DBuf& G, hipStream_t s) {
    const int np = launch_conv_wgrad(N.cin[l], N.cout[l], N.d[lev], cb(X), cb(G), B<float>(N.wpart), s);
    launch_sum_parts(B<float>(N.wpart), np, (size_t)27 * N.cin[l] * N.cout[l], B<float>(N.G) + N.woff[l], 1, s);
    launch_bias_grad(N.cout[l], P_of(N.d[lev]), cb(G), B<float>(N.bpart), B<float>(N.G) + N.boff[l], s);
}

// forward pass over N.F (already shaped); output in N.Y (FP32 [P0][16])
static void forward(Net& N, hipStream_t s) {
    conv_f(N, E0A, EPI_RELU, 0, N.F, nullptr, N.a0, s);
    conv_f(N, E0B, EPI_RELU, 0, N.a0, nullptr, N.s0, s);
    launch_pool(16, N.d[0], cb(N.s0), mb(N.q1), s);
    conv_f(N, E1A, EPI_RELU, 1, N.q1, nullptr, N.a1, s);
    conv_f(N, E1B, EPI_RELU, 1, N.a1, nullptr, N.s1, s);
    launch_pool(32, N.d[1], cb(N.s1), mb(N.q2), s);
    conv_f(N, E2A, EPI_RELU, 2, N.q2, nullptr, N.a2, s);
    conv_f(N, E2B, EPI_RELU, 2, N.a2, nullptr, N.b2, s);
    launch_upsample(64, N.d[1], cb(N.b2), mb(N.u1), s);
    conv_f(N, D1A, EPI_RELU_SKIP, 1, N.u1, &N.s1, N.c1, s);
    conv_f(N, D1B, EPI_RELU, 1, N.c1, nullptr, N.t1, s);
    launch_upsample(32, N.d[0], cb(N.t1), mb(N.u0), s);
    conv_f(N, D0A, EPI_RELU_SKIP, 0, N.u0, &N.s0, N.c0, s);
    conv_f(N, D0B, EPI_RELU, 0, N.c0, nullptr, N.t0, s);
    conv_f(N, HEAD, EPI_F32, 0, N.t0, nullptr, N.Y, s);
}

// backward pass from N.gY (bf16 [P0][16]); adds the parameter gradients to N.G
static void backward(Net& N, hipStream_t s) {
    grads(N, HEAD, 0, N.t0, N.gY, s);
    conv_b(N, HEAD, EPI_MASK, 0, N.gY, &N.t0, N.gt0, s);
    grads(N, D0B, 0, N.c0, N.gt0, s);
    conv_b(N, D0B, EPI_MASK, 0, N.gt0, &N.c0, N.gc0, s);
    grads(N, D0A, 0, N.u0, N.gc0, s);
    conv_b(N, D0A, EPI_RAW_F32, 0, N.gc0, nullptr, N.hu0, s);
    launch_upsample_t(32, N.d[0], B<float>(N.hu0), cb(N.t1), mb(N.gt1), s);
    grads(N, D1B, 1, N.c1, N.gt1, s);
    conv_b(N, D1B, EPI_MASK, 1, N.gt1, &N.c1, N.gc1, s);
    grads(N, D1A, 1, N.u1, N.gc1, s);
    conv_b(N, D1A, EPI_RAW_F32, 1, N.gc1, nullptr, N.hu1, s);
    launch_upsample_t(64, N.d[1], B<float>(N.hu1), cb(N.b2), mb(N.gb2), s);
    grads(N, E2B, 2, N.a2, N.gb2, s);
    conv_b(N, E2B, EPI_MASK, 2, N.gb2, &N.a2, N.ga2, s);
    grads(N, E2A, 2, N.q2, N.ga2, s);
    conv_b(N, E2A, EPI_RAW_F32, 2, N.ga2, nullptr, N.hq2, s);
    launch_pool_t(32, N.d[1], B<float>(N.hq2), cb(N.gc1), cb(N.s1), mb(N.gs1), s);
    grads(N, E1B, 1, N.a1, N.gs1, s);
    conv_b(N, E1B, EPI_MASK, 1, N.gs1, &N.a1, N.ga1, s);
    grads(N, E1A, 1, N.q1, N.ga1, s);
    conv_b(N, E1A, EPI_RAW_F32, 1, N.ga1, nullptr, N.hq1, s);
    launch_pool_t(16, N.d[0], B<float>(N.hq1), cb(N.gc0), cb(N.s0), mb(N.gs0), s);
    grads(N, E0B, 0, N.a0, N.gs0, s);
    conv_b(N, E0B, EPI_MASK, 0, N.gs0, &N.a0, N.ga0, s);
    grads(N, E0A, 0, N.F, N.ga0, s);
}

// loss = mean over positions and the 3 output channels of (Y - tscale T)^2;
// leaves its gradient in N.gY
static double loss_grad(Net& N, const float* T, float tscale, hipStream_t s) {
    const size_t P = P_of(N.d[0]);
    const float inv = 1.f / (3.f * (float)P);
    launch_loss(P, B<float>(N.Y), T, tscale, inv, mb(N.gY), B<double>(N.lpart), kLossParts, s);
    std::vector<double> h(kLossParts);
    RT_HIP(hipMemcpyAsync(h.data(), N.lpart.p, sizeof(double) * kLossParts, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    double sum = 0;
    for (double v : h) sum += v;
    return sum * inv;
}

void net_predict(Net& N, const Geom& g0, const int e0[3], const float* rho, const Geom& gc, Vec3<const double> Fc,
                 const uint8_t* vflag_c, Vec3<double> Uc, hipStream_t s) {
    RT_HIP(hipStreamSynchronize(N.s));
    shape(N, gc.n);
    RT_HIP(hipMemsetAsync(N.absmax.p, 0, sizeof(unsigned), s));
    launch_absmax3(gc, Fc, N.absmax.as<unsigned>(), s);
    unsigned bits = 0;
    RT_HIP(hipMemcpyAsync(&bits, N.absmax.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    float fmax;
    std::memcpy(&fmax, &bits, sizeof(float));
    launch_net_features(g0, e0, rho, gc, Fc, fmax > 0 ? 1.f / fmax : 0.f, mb(N.F), s);
    forward(N, s);
    launch_net_to_level(gc, B<float>(N.Y), vflag_c, Uc, s);
}

static NetSample* make_sample(Net& N, const Geom& g0, const int e0[3], const float* rho, const Geom& gc,
                              Vec3<const double> Fc, Vec3<const double> U, hipStream_t s) {
    RT_HIP(hipStreamSynchronize(N.s));
    std::unique_ptr<NetSample> S(new NetSample());
    const size_t P = (size_t)gc.n[0] * gc.n[1] * gc.n[2];
    for (int k = 0; k < 3; k++) S->dims[k] = gc.n[k];
    S->X.alloc(sizeof(uint16_t) * P * 16);
    S->T.alloc(sizeof(float) * P * 3);
    RT_HIP(hipMemsetAsync(N.absmax.p, 0, sizeof(unsigned), s));
    launch_absmax3(gc, Fc, N.absmax.as<unsigned>(), s);
    launch_net_target(g0, gc, U, S->T.as<float>(), B<double>(N.lpart), kLossParts, s);
    unsigned bits = 0;
    std::vector<double> h(kLossParts);
    RT_HIP(hipMemcpyAsync(&bits, N.absmax.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(h.data(), N.lpart.p, sizeof(double) * kLossParts, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    float fmax;
    std::memcpy(&fmax, &bits, sizeof(float));
    launch_net_features(g0, e0, rho, gc, Fc, fmax > 0 ? 1.f / fmax : 0.f, S->X.as<uint16_t>(), s);
    double ss = 0;
    for (double v : h) ss += v;
    if (!(ss > 0)) throw std::domain_error("training sample has a zero displacement");
    S->tscale = (float)(1.0 / std::sqrt(ss / (3.0 * (double)P)));
    RT_HIP(hipStreamSynchronize(s));
    return S.release();
}

void net_add_sample(Net& N, const Geom& g0, const int e0[3], const float* rho, const Geom& gc,
                    Vec3<const double> Fc, Vec3<const double> U, hipStream_t s) {
    N.samples.push_back(make_sample(N, g0, e0, rho, gc, Fc, U, s));
}

double net_evaluate(Net& N, const Geom& g0, const int e0[3], const float* rho, const Geom& gc,
                    Vec3<const double> Fc, Vec3<const double> U, hipStream_t s) {
    std::unique_ptr<NetSample> S(make_sample(N, g0, e0, rho, gc, Fc, U, s));
    shape(N, S->dims);
    RT_HIP(hipMemcpyAsync(N.F.p, S->X.p, sizeof(uint16_t) * P_of(S->dims) * 16, hipMemcpyDeviceToDevice, N.s));
    forward(N, N.s);
    return loss_grad(N, S->T.as<float>(), S->tscale, N.s);
}

static double train(Net& N, int steps, int batch, float lr, uint64_t seed) {
    if (N.samples.empty()) throw std::logic_error("no training samples");
    if (steps < 1 || batch < 1) throw std::invalid_argument("steps and batch must be positive");
    std::mt19937_64 gen(seed);
    std::uniform_int_distribution<size_t> pick(0, N.samples.size() - 1);
    double last = 0;
    for (int it = 0; it < steps; it++) {
        RT_HIP(hipMemsetAsync(N.G.p, 0, sizeof(float) * N.nparam, N.s));
        double lsum = 0;
        for (int b = 0; b < batch; b++) {
            const NetSample& S = *N.samples[pick(gen)];
            shape(N, S.dims);
            RT_HIP(hipMemcpyAsync(N.F.p, S.X.p, sizeof(uint16_t) * P_of(S.dims) * 16, hipMemcpyDeviceToDevice, N.s));
            forward(N, N.s);
            lsum += loss_grad(N, S.T.as<float>(), S.tscale, N.s);
            backward(N, N.s);
        }
        N.adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        launch_adam(N.nparam, B<float>(N.W), B<float>(N.G), B<float>(N.M), B<float>(N.V), lr, b1, b2, 1e-8f,
                    1.f - std::pow(b1, (float)N.adam_t), 1.f - std::pow(b2, (float)N.adam_t), 1.f / batch, N.s);
        pack_all(N, N.s);
        last = lsum / batch;
    }
    RT_HIP(hipStreamSynchronize(N.s));
    return last;
}

}  // namespace rt

using namespace rt;

static Net& N_(rt_net* n) {
    if (!n || !n->n) throw std::invalid_argument("null rt_net handle");
    return *n->n;
}
static void need_ptr(const void* p) {
    if (!p) throw std::invalid_argument("null pointer argument");
}

#define RT_NET_TRY(...)                                        \
    try {                                                      \
        __VA_ARGS__;                                           \
        return RT_OK;                                          \
    } catch (...) {                                            \
        return status_of_current_exception();                  \
    }

extern "C" {

int rt_net_create(uint64_t seed, rt_net** out) {
    RT_NET_TRY({
        need_ptr(out);
        *out = nullptr;
        std::unique_ptr<Net> n(net_new(seed));
        *out = new rt_net{n.release()};
    })
}

int rt_net_destroy(rt_net* n) {
    RT_NET_TRY({
        if (n) { delete n->n; delete n; }
    })
}

int rt_net_num_params(const rt_net* n, size_t* count) {
    RT_NET_TRY({
        need_ptr(count);
        *count = N_(const_cast<rt_net*>(n)).nparam;
    })
}

int rt_net_get_params(const rt_net* n, float* w) {
    RT_NET_TRY({
        need_ptr(w);
        Net& N = N_(const_cast<rt_net*>(n));
        RT_HIP(hipStreamSynchronize(N.s));
        RT_HIP(hipMemcpy(w, N.W.p, sizeof(float) * N.nparam, hipMemcpyDeviceToHost));
    })
}

int rt_net_set_params(rt_net* n, const float* w) {
    RT_NET_TRY({
        need_ptr(w);
        Net& N = N_(n);
        RT_HIP(hipStreamSynchronize(N.s));
        RT_HIP(hipMemcpy(N.W.p, w, sizeof(float) * N.nparam, hipMemcpyHostToDevice));
        RT_HIP(hipMemset(N.M.p, 0, sizeof(float) * N.nparam));
        RT_HIP(hipMemset(N.V.p, 0, sizeof(float) * N.nparam));
        N.adam_t = 0;
        pack_all(N, N.s);
        RT_HIP(hipStreamSynchronize(N.s));
    })
}

int rt_net_forward(rt_net* n, const int* dims, const float* x, float* y) {
    RT_NET_TRY({
        need_ptr(dims); need_ptr(x); need_ptr(y);
        Net& N = N_(n);
        shape(N, dims);
        const size_t P = P_of(dims);
        DBuf xin, yout;
        xin.alloc(sizeof(float) * P * 4);
        yout.alloc(sizeof(float) * P * 3);
        RT_HIP(hipMemcpyAsync(xin.p, x, sizeof(float) * P * 4, hipMemcpyHostToDevice, N.s));
        launch_pack_input(P, xin.as<float>(), mb(N.F), N.s);
        forward(N, N.s);
        launch_unpack_output(P, B<float>(N.Y), yout.as<float>(), N.s);
        RT_HIP(hipMemcpyAsync(y, yout.p, sizeof(float) * P * 3, hipMemcpyDeviceToHost, N.s));
        RT_HIP(hipStreamSynchronize(N.s));
    })
}

int rt_net_gradient(rt_net* n, const int* dims, const float* x, const float* target, double* loss, float* grad) {
    RT_NET_TRY({
        need_ptr(dims); need_ptr(x); need_ptr(target); need_ptr(loss); need_ptr(grad);
        Net& N = N_(n);
        shape(N, dims);
        const size_t P = P_of(dims);
        DBuf xin;
        xin.alloc(sizeof(float) * P * 4);
        N.T.ensure(sizeof(float) * P * 3);
        RT_HIP(hipMemcpyAsync(xin.p, x, sizeof(float) * P * 4, hipMemcpyHostToDevice, N.s));
        RT_HIP(hipMemcpyAsync(N.T.p, target, sizeof(float) * P * 3, hipMemcpyHostToDevice, N.s));
        RT_HIP(hipMemsetAsync(N.G.p, 0, sizeof(float) * N.nparam, N.s));
        launch_pack_input(P, xin.as<float>(), mb(N.F), N.s);
        forward(N, N.s);
        *loss = loss_grad(N, N.T.as<float>(), 1.f, N.s);
        backward(N, N.s);
        RT_HIP(hipMemcpyAsync(grad, N.G.p, sizeof(float) * N.nparam, hipMemcpyDeviceToHost, N.s));
        RT_HIP(hipStreamSynchronize(N.s));
    })
}

int rt_net_time_forward(rt_net* n, const int* dims, int reps, double* ms) {
    RT_NET_TRY({
        need_ptr(dims); need_ptr(ms);
        if (reps < 1) throw std::invalid_argument("reps must be positive");
        Net& N = N_(n);
        shape(N, dims);
        forward(N, N.s);   // warm-up (inputs: whatever the activation buffers hold)
        hipEvent_t e0, e1;
        RT_HIP(hipEventCreate(&e0));
        RT_HIP(hipEventCreate(&e1));
        RT_HIP(hipEventRecord(e0, N.s));
        for (int r = 0; r < reps; r++) forward(N, N.s);
        RT_HIP(hipEventRecord(e1, N.s));
        RT_HIP(hipEventSynchronize(e1));
        float t = 0;
        RT_HIP(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = t / reps;
    })
}

int rt_net_num_samples(const rt_net* n, int* count) {
    RT_NET_TRY({
        need_ptr(count);
        *count = (int)N_(const_cast<rt_net*>(n)).samples.size();
    })
}

int rt_net_clear_samples(rt_net* n) {
    RT_NET_TRY({
        Net& N = N_(n);
        RT_HIP(hipStreamSynchronize(N.s));
        for (NetSample* p : N.samples) delete p;
        N.samples.clear();
    })
}

int rt_net_adam_step(rt_net* n, const float* grad, float lr) {
    RT_NET_TRY({
        need_ptr(grad);
        Net& N = N_(n);
        RT_HIP(hipStreamSynchronize(N.s));
        RT_HIP(hipMemcpy(N.G.p, grad, sizeof(float) * N.nparam, hipMemcpyHostToDevice));
        N.adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        launch_adam(N.nparam, B<float>(N.W), B<float>(N.G), B<float>(N.M), B<float>(N.V), lr, b1, b2, 1e-8f,
                    1.f - std::pow(b1, (float)N.adam_t), 1.f - std::pow(b2, (float)N.adam_t), 1.f, N.s);
        pack_all(N, N.s);
        RT_HIP(hipStreamSynchronize(N.s));
    })
}

int rt_net_train(rt_net* n, int steps, int batch, float lr, uint64_t seed, double* mean_loss) {
    RT_NET_TRY({
        const double l = train(N_(n), steps, batch, lr, seed);
        if (mean_loss) *mean_loss = l;
    })
}

}  // extern "C"
