// extern "C" entry points of the batched persistent Poisson solver.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <new>
#include <numeric>
#include <vector>

#include "batched.hpp"

using namespace sil;

struct sil_batched_poisson_s {
    int nprob = 0;
    int max_side = 0;
    int64_t total = 0;
    std::vector<BatchProblem> problems;
    BatchProblem* problems_dev = nullptr;
    int* order_dev = nullptr;
    int* queue = nullptr;
    float *wx32 = nullptr, *wy32 = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *dinv32 = nullptr;
    double *wx64 = nullptr, *wy64 = nullptr, *b = nullptr, *x = nullptr;
    BatchResult* results_dev = nullptr;
    int block = 256;
    int cus = 0;
    hipGraphExec_t graph = nullptr;
    int graph_block = 0;
    int graph_max_it = -1, graph_replace = -1, graph_jacobi = -1;
    int jacobi = 1; // preconditioner of the captured solve
    double graph_tol = -1.0;
};

namespace {

void release(sil_batched_poisson h)
{
    if (!h) return;
    for (void* ptr : {(void*)h->problems_dev, (void*)h->order_dev, (void*)h->queue, (void*)h->wx32, (void*)h->wy32,
                      (void*)h->dinv32,
                      (void*)h->r, (void*)h->p, (void*)h->q, (void*)h->wx64, (void*)h->wy64, (void*)h->b,
                      (void*)h->x, (void*)h->results_dev})
        if (ptr) (void)hipFree(ptr);
    if (h->graph) (void)hipGraphExecDestroy(h->graph);
    delete h;
}

template <typename T>
int dalloc(T** p, int64_t n)
{
    SIL_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * std::max<int64_t>(n, 1)));
    return SIL_OK;
}

int grid_for(sil_batched_poisson h, int block, int* grid)
{
    int per_cu = 0;
    int rc = batched_occupancy(block, &per_cu);
    if (rc) return rc;
    *grid = std::max(1, std::min(h->nprob, h->cus * std::max(per_cu, 1)));
    return SIL_OK;
}

int enqueue(sil_batched_poisson h, int block, double tol, int max_it, int replace, hipStream_t s)
{
    BatchLaunch L{};
    L.problems = h->problems_dev;
    L.order = h->order_dev;
    L.nprob = h->nprob;
    L.queue = h->queue;
    L.wx32 = h->wx32;
    L.wy32 = h->wy32;
    L.wx64 = h->wx64;
    L.wy64 = h->wy64;
    L.b = h->b;
    L.x = h->x;
    L.r = h->r;
    L.p = h->p;
    L.q = h->q;
    L.dinv32 = h->dinv32;
    L.jacobi = h->jacobi;
    L.tol = tol;
    L.max_it = max_it;
    L.replace = replace;
    L.results = h->results_dev;
    L.block = block;
    int rc = grid_for(h, block, &L.grid);
    if (rc) return rc;
    return batched_enqueue(L, s);
}

int capture(sil_batched_poisson h, int block, double tol, int max_it, int replace)
{
    if (h->graph && h->graph_block == block && h->graph_tol == tol && h->graph_max_it == max_it &&
        h->graph_replace == replace && h->graph_jacobi == h->jacobi)
        return SIL_OK;
    if (h->graph) {
        (void)hipGraphExecDestroy(h->graph);
        h->graph = nullptr;
    }
    hipStream_t cs;
    SIL_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        set_error("batched: capture failed: %s", hipGetErrorString(e));
        return SIL_ERR_HIP;
    }
    int rc = enqueue(h, block, tol, max_it, replace, cs);
    e = hipStreamEndCapture(cs, &g);
    if (rc == SIL_OK && e == hipSuccess) e = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
    if (e == hipSuccess || rc != SIL_OK) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    if (rc) return rc;
    if (e != hipSuccess) {
        set_error("batched: graph failed: %s", hipGetErrorString(e));
        return SIL_ERR_HIP;
    }
    h->graph_block = block;
    h->graph_tol = tol;
    h->graph_jacobi = h->jacobi;
    h->graph_max_it = max_it;
    h->graph_replace = replace;
    return SIL_OK;
}

}  // namespace

extern "C" {

int sil_batched_poisson2d_create(sil_batched_poisson* out, int32_t nproblems, const int32_t* sides,
                                 const double* kappa)
{
    SIL_REQUIRE(out && sides && kappa, SIL_ERR_INVALID, "batched_poisson2d_create: null argument");
    SIL_REQUIRE(nproblems >= 1, SIL_ERR_INVALID, "batched_poisson2d_create: need at least one problem");
    auto* h = new (std::nothrow) sil_batched_poisson_s();
    SIL_REQUIRE(h != nullptr, SIL_ERR_INVALID, "batched_poisson2d_create: out of host memory");
    h->nprob = nproblems;
    int64_t off = 0;
    for (int i = 0; i < nproblems; ++i) {
        if (sides[i] < 3 || sides[i] > 46340) {
            release(h);
            set_error("batched_poisson2d_create: side %d of problem %d outside 3..46340", sides[i], i);
            return SIL_ERR_INVALID;
        }
        h->problems.push_back(BatchProblem{off, sides[i], 0});
        off += (int64_t)sides[i] * sides[i];
        h->max_side = std::max(h->max_side, (int)sides[i]);
    }
    h->total = off;
    std::vector<int> order(nproblems);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sides[a] > sides[b]; });
    int rc;
    if ((rc = dalloc(&h->problems_dev, nproblems)) || (rc = dalloc(&h->order_dev, nproblems)) ||
        (rc = dalloc(&h->queue, 1)) || (rc = dalloc(&h->wx32, off)) || (rc = dalloc(&h->wy32, off)) ||
        (rc = dalloc(&h->r, off)) || (rc = dalloc(&h->p, off)) || (rc = dalloc(&h->q, off)) ||
        (rc = dalloc(&h->wx64, off)) || (rc = dalloc(&h->wy64, off)) || (rc = dalloc(&h->b, off)) ||
        (rc = dalloc(&h->x, off)) || (rc = dalloc(&h->dinv32, off)) || (rc = dalloc(&h->results_dev, nproblems))) {
        release(h);
        return rc;
    }
    hipError_t e = hipMemcpy(h->problems_dev, h->problems.data(), sizeof(BatchProblem) * nproblems,
                             hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->order_dev, order.data(), sizeof(int) * nproblems, hipMemcpyHostToDevice);
    int dev = 0;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) {
        release(h);
        set_error("batched_poisson2d_create: %s", hipGetErrorString(e));
        return SIL_ERR_HIP;
    }
    rc = batched_weights(h->problems_dev, nproblems, h->max_side, kappa, h->wx32, h->wy32, h->wx64, h->wy64,
                         h->dinv32, 0);
    if (rc == SIL_OK && hipDeviceSynchronize() != hipSuccess) rc = SIL_ERR_HIP;
    if (rc) {
        release(h);
        return rc;
    }
    *out = h;
    return SIL_OK;
}

int64_t sil_batched_poisson2d_offset(sil_batched_poisson h, int32_t problem)
{
    if (!h || problem < 0 || problem > h->nprob) return -1;
    return problem == h->nprob ? h->total : h->problems[problem].offset;
}

int sil_batched_poisson2d_set_block(sil_batched_poisson h, int32_t block)
{
    SIL_REQUIRE(h != nullptr, SIL_ERR_INVALID, "batched: null handle");
    SIL_REQUIRE(block == 256 || block == 512 || block == 1024, SIL_ERR_INVALID, "batched: workgroup size %d", block);
    h->block = block;
    return SIL_OK;
}

int sil_batched_poisson2d_solve(sil_batched_poisson h, const double* rhs, double* solution,
                                const sil_solver_settings* st, sil_batch_result* results, double* seconds,
                                void* stream)
{
    SIL_REQUIRE(h && rhs && solution && st, SIL_ERR_INVALID, "batched_poisson2d_solve: null argument");
    SIL_REQUIRE(st->precision == SIL_PRECISION_MIXED || st->precision == SIL_PRECISION_FP64, SIL_ERR_INVALID,
                "batched_poisson2d_solve: unknown precision");
    SIL_REQUIRE(st->preconditioner == SIL_PRECOND_IDENTITY || st->preconditioner == SIL_PRECOND_JACOBI,
                SIL_ERR_INVALID, "batched_poisson2d_solve: preconditioner must be identity or Jacobi");
    hipStream_t s = as_stream(stream);
    const int replace = st->replace_interval > 0 ? st->replace_interval : 400;
    h->jacobi = st->preconditioner == SIL_PRECOND_JACOBI ? 1 : 0;
    int rc = capture(h, h->block, st->relative_tolerance, st->max_iterations, replace);
    if (rc) return rc;
    SIL_CHECK_HIP(hipMemcpyAsync(h->b, rhs, sizeof(double) * h->total, hipMemcpyDeviceToDevice, s));
    SIL_CHECK_HIP(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();
    SIL_CHECK_HIP(hipGraphLaunch(h->graph, s));
    SIL_CHECK_HIP(hipStreamSynchronize(s));
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (seconds) *seconds = dt;
    SIL_CHECK_HIP(hipMemcpyAsync(solution, h->x, sizeof(double) * h->total, hipMemcpyDeviceToDevice, s));
    if (results) {
        std::vector<BatchResult> tmp(h->nprob);
        SIL_CHECK_HIP(hipMemcpyAsync(tmp.data(), h->results_dev, sizeof(BatchResult) * h->nprob,
                                     hipMemcpyDeviceToHost, s));
        SIL_CHECK_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < h->nprob; ++i) {
            results[i].iterations = tmp[i].iterations;
            results[i].converged = tmp[i].converged;
            results[i].true_relative_residual = tmp[i].true_relative_residual;
        }
    }
    SIL_CHECK_HIP(hipStreamSynchronize(s));
    return SIL_OK;
}

int sil_batched_poisson2d_autotune(sil_batched_poisson h, const double* rhs, int32_t probe_iterations,
                                   int32_t* chosen_block, double* seconds_per_candidate)
{
    SIL_REQUIRE(h && rhs && chosen_block, SIL_ERR_INVALID, "batched_poisson2d_autotune: null argument");
    const int cands[3] = {256, 512, 1024};
    double best = 1e300;
    int best_block = 256;
    SIL_CHECK_HIP(hipMemcpy(h->b, rhs, sizeof(double) * h->total, hipMemcpyDeviceToDevice));
    for (int c = 0; c < 3; ++c) {
        // a fixed number of iterations on every problem (tolerance 0): equal work per candidate
        int rc = capture(h, cands[c], 0.0, std::max(1, (int)probe_iterations), 1 << 30);
        if (rc) return rc;
        SIL_CHECK_HIP(hipGraphLaunch(h->graph, 0)); // warm-up
        SIL_CHECK_HIP(hipDeviceSynchronize());
        const auto t0 = std::chrono::steady_clock::now();
        SIL_CHECK_HIP(hipGraphLaunch(h->graph, 0));
        SIL_CHECK_HIP(hipDeviceSynchronize());
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (seconds_per_candidate) seconds_per_candidate[c] = dt;
        if (dt < best) {
            best = dt;
            best_block = cands[c];
        }
    }
    h->block = best_block;
    *chosen_block = best_block;
    return SIL_OK;
}

int sil_batched_poisson2d_destroy(sil_batched_poisson h)
{
    release(h);
    return SIL_OK;
}

}  // extern "C"
