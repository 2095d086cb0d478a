// Batched persistent Poisson solver (host side).
#pragma once

#include "common.hpp"

namespace sil {

struct BatchProblem {
    int64_t offset; // first node of the problem in the concatenated arrays
    int32_t side;   // n: the problem has n×n nodes, (n−2)² unknowns
    int32_t pad;
};

struct BatchResult {
    int32_t iterations;
    int32_t converged;
    double true_relative_residual;
};

struct BatchLaunch {
    const BatchProblem* problems;
    const int* order;
    int nprob;
    int* queue;
    const float *wx32, *wy32;
    const double *wx64, *wy64;
    const double* b;
    double* x;
    float *r, *p, *q;
    const float* dinv32;
    int jacobi;
    double tol;
    int max_it;
    int replace;
    BatchResult* results;
    int block; // workgroup size
    int grid;  // persistent workgroups
};

// Wide path: large problems, each spread over many workgroups.  One CG
// iteration is a sequence of kernels (per-block partial sums, a per-problem
// scalar kernel that reduces them in a fixed order); kernel boundaries are
// the only synchronisation between workgroups.  Same arithmetic as the
// persistent kernel: FP32 recurrence, FP64 x, dot products and scalars, FP64
// residual replacement, convergence only on the FP64 residual.
struct WideState {
    double bb, tol2bb, rz, rz_old, rr_ref, rr_true, alpha, beta, pq, rz_new, rr_new;
    int32_t it, since, converged, done, replace_now, pad;
};

struct WideLaunch {
    const BatchProblem* problems;
    const int* ids;      // problem of each wide slot
    int nwide;
    int nbx;             // workgroups per problem
    const float *wx32, *wy32;
    const double *wx64, *wy64;
    const double* b;
    double* x;
    float *r, *p, *q;
    const float* dinv32;
    int jacobi;
    double tol;
    int max_it;
    int replace;
    WideState* state;
    double* part;        // nwide · nbx · 2
    int* done;           // nwide flags (host-visible through a copy)
    BatchResult* results;
};

int wide_init(const WideLaunch& L, hipStream_t s);
int wide_iteration(const WideLaunch& L, hipStream_t s);
int wide_finish(const WideLaunch& L, hipStream_t s);

int batched_weights(const BatchProblem* problems_dev, int nprob, int max_side, const double* kappa, float* wx32,
                    float* wy32, double* wx64, double* wy64, float* dinv32, hipStream_t s);
int batched_enqueue(const BatchLaunch& L, hipStream_t s);
int batched_occupancy(int block, int* blocks_per_cu);

}  // namespace sil
