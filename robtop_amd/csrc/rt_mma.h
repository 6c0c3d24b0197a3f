// Device-side pieces of the MMA optimiser shared by k_mma.hip and rt_mma.cpp.
#pragma once
#include "rt_internal.h"

#include <algorithm>

#define RT_MMA_MAX_M 8

namespace rt {

// the optimiser's length-n FP32 vectors (device pointers)
struct MmaVecs {
    float *x, *xmin, *xmax, *al, *au, *alpha, *beta, *lastdx, *xi, *eta, *curx, *dx, *nx, *nxi, *neta;
};
// m multipliers passed by value
struct MmaLam {
    float v[RT_MMA_MAX_M];
};

int mma_blocks(int n);
int mma_newton_slots();
int mma_trial_slots();
void launch_mma_init(const MmaVecs& V, int n, const float* lo, const float* hi, hipStream_t s);
void launch_mma_pq(const MmaVecs& V, int n, int m, const float* df, const float* dg, float* p, float* q,
                   double* part, hipStream_t s);
void launch_mma_reset(const MmaVecs& V, int n, hipStream_t s);
void launch_mma_newton(const MmaVecs& V, const float* p, const float* q, const MmaLam& L, int n, int m, float eps,
                       double* part, hipStream_t s);
void launch_mma_dx(const MmaVecs& V, const float* p, const float* q, const MmaLam& L, const MmaLam& dlam, int n,
                   int m, float eps, double* part, hipStream_t s);
void launch_mma_trial(const MmaVecs& V, const float* p, const float* q, const MmaLam& Lnew, int n, int m, float eps,
                      float step, double* part, hipStream_t s);
void launch_mma_adjust(const MmaVecs& V, int n, double* part, hipStream_t s);

}  // namespace rt
