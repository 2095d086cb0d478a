// extern "C" entry points of the batched persistent Poisson solver.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
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
    int* order_dev = nullptr;   // queue of the persistent kernel: the small problems, largest first
    int nsmall = 0;
    // wide path: problems of side >= wide_side, each over nbx workgroups
    int wide_side = 192;
    int nwide = 0, nbx = 1;
    int nbx_max = 1; // workgroups per wide problem the partial-sum array holds
    int* wide_ids_dev = nullptr;
    WideState* wide_state = nullptr;
    double* wide_part = nullptr;
    int* wide_done = nullptr;
    int* wide_done_host = nullptr; // pinned
    hipGraphExec_t wide_graph = nullptr;
    int wide_max_it = -1, wide_replace = -1, wide_jacobi = -1;
    double wide_tol = -1.0;
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
                      (void*)h->x, (void*)h->results_dev, (void*)h->wide_ids_dev, (void*)h->wide_state,
                      (void*)h->wide_part, (void*)h->wide_done})
        if (ptr) (void)hipFree(ptr);
    if (h->wide_done_host) (void)hipHostFree(h->wide_done_host);
    if (h->graph) (void)hipGraphExecDestroy(h->graph);
    if (h->wide_graph) (void)hipGraphExecDestroy(h->wide_graph);
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
    *grid = std::max(1, std::min(h->nsmall, h->cus * std::max(per_cu, 1)));
    return SIL_OK;
}

int enqueue(sil_batched_poisson h, int block, double tol, int max_it, int replace, hipStream_t s)
{
    BatchLaunch L{};
    L.problems = h->problems_dev;
    L.order = h->order_dev;
    L.nprob = h->nsmall;
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

WideLaunch wide_launch(sil_batched_poisson h, double tol, int max_it, int replace)
{
    WideLaunch W{};
    W.problems = h->problems_dev;
    W.ids = h->wide_ids_dev;
    W.nwide = h->nwide;
    W.nbx = h->nbx;
    W.wx32 = h->wx32;
    W.wy32 = h->wy32;
    W.wx64 = h->wx64;
    W.wy64 = h->wy64;
    W.b = h->b;
    W.x = h->x;
    W.r = h->r;
    W.p = h->p;
    W.q = h->q;
    W.dinv32 = h->dinv32;
    W.jacobi = h->jacobi;
    W.tol = tol;
    W.max_it = max_it;
    W.replace = replace;
    W.state = h->wide_state;
    W.part = h->wide_part;
    W.done = h->wide_done;
    W.results = h->results_dev;
    return W;
}

// iterations of the wide path per graph launch (the host reads the done
// flags between launches)
constexpr int kWideChunk = 8;

int capture_wide(sil_batched_poisson h, double tol, int max_it, int replace)
{
    if (h->wide_graph && h->wide_tol == tol && h->wide_max_it == max_it && h->wide_replace == replace &&
        h->wide_jacobi == h->jacobi)
        return SIL_OK;
    if (h->wide_graph) {
        (void)hipGraphExecDestroy(h->wide_graph);
        h->wide_graph = nullptr;
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
    const WideLaunch W = wide_launch(h, tol, max_it, replace);
    int rc = SIL_OK;
    for (int i = 0; i < kWideChunk && rc == SIL_OK; ++i) rc = wide_iteration(W, cs);
    e = hipStreamEndCapture(cs, &g);
    if (rc == SIL_OK && e == hipSuccess) e = hipGraphInstantiate(&h->wide_graph, g, nullptr, nullptr, 0);
    if (e == hipSuccess || rc != SIL_OK) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    if (rc) return rc;
    if (e != hipSuccess) {
        set_error("batched: graph failed: %s", hipGetErrorString(e));
        return SIL_ERR_HIP;
    }
    h->wide_tol = tol;
    h->wide_max_it = max_it;
    h->wide_replace = replace;
    h->wide_jacobi = h->jacobi;
    return SIL_OK;
}

// the wide path to completion on stream s
int solve_wide(sil_batched_poisson h, double tol, int max_it, int replace, hipStream_t s)
{
    int rc = capture_wide(h, tol, max_it, replace);
    if (rc) return rc;
    const WideLaunch W = wide_launch(h, tol, max_it, replace);
    if ((rc = wide_init(W, s))) return rc;
    for (int64_t done_it = 0; done_it < max_it; done_it += kWideChunk) {
        SIL_CHECK_HIP(hipMemcpyAsync(h->wide_done_host, h->wide_done, sizeof(int) * h->nwide,
                                     hipMemcpyDeviceToHost, s));
        SIL_CHECK_HIP(hipStreamSynchronize(s));
        bool all = true;
        for (int i = 0; i < h->nwide && all; ++i) all = h->wide_done_host[i] != 0;
        if (all) break;
        SIL_CHECK_HIP(hipGraphLaunch(h->wide_graph, s));
    }
    return wide_finish(W, s);
}

// splits the problems between the persistent kernel (side < wide_side,
// largest first) and the wide path, and sizes the wide path's arrays;
// drops the captured graphs (they hold the previous split)
int partition(sil_batched_poisson h, int wide_side)
{
    for (void* ptr : {(void*)h->order_dev, (void*)h->wide_ids_dev, (void*)h->wide_state, (void*)h->wide_part,
                      (void*)h->wide_done})
        if (ptr) (void)hipFree(ptr);
    h->order_dev = nullptr;
    h->wide_ids_dev = nullptr;
    h->wide_state = nullptr;
    h->wide_part = nullptr;
    h->wide_done = nullptr;
    if (h->wide_done_host) (void)hipHostFree(h->wide_done_host);
    h->wide_done_host = nullptr;
    if (h->graph) (void)hipGraphExecDestroy(h->graph);
    if (h->wide_graph) (void)hipGraphExecDestroy(h->wide_graph);
    h->graph = nullptr;
    h->wide_graph = nullptr;
    h->graph_block = 0;
    h->wide_side = wide_side;
    std::vector<int> order, wide;
    int64_t max_wide_nn = 0;
    for (int i = 0; i < h->nprob; ++i) {
        const int side = h->problems[i].side;
        if (side >= wide_side) {
            wide.push_back(i);
            max_wide_nn = std::max<int64_t>(max_wide_nn, (int64_t)side * side);
        } else {
            order.push_back(i);
        }
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return h->problems[a].side > h->problems[b].side; });
    h->nsmall = (int)order.size();
    h->nwide = (int)wide.size();
    // four nodes per thread of the largest wide problem, at most 256 workgroups per problem
    h->nbx = (int)std::min<int64_t>(256, std::max<int64_t>(1, (max_wide_nn + 1023) / 1024));
    h->nbx_max = h->nbx;
    int rc;
    if ((rc = dalloc(&h->order_dev, std::max(h->nsmall, 1))) || (rc = dalloc(&h->wide_ids_dev, std::max(h->nwide, 1))) ||
        (rc = dalloc(&h->wide_state, std::max(h->nwide, 1))) ||
        (rc = dalloc(&h->wide_part, (int64_t)std::max(h->nwide, 1) * h->nbx_max * 2)) ||
        (rc = dalloc(&h->wide_done, std::max(h->nwide, 1))))
        return rc;
    if (hipHostMalloc(reinterpret_cast<void**>(&h->wide_done_host), sizeof(int) * std::max(h->nwide, 1)) !=
        hipSuccess) {
        h->wide_done_host = nullptr;
        set_error("batched: pinned allocation failed");
        return SIL_ERR_HIP;
    }
    hipError_t e = hipSuccess;
    if (h->nsmall) e = hipMemcpy(h->order_dev, order.data(), sizeof(int) * h->nsmall, hipMemcpyHostToDevice);
    if (e == hipSuccess && h->nwide)
        e = hipMemcpy(h->wide_ids_dev, wide.data(), sizeof(int) * h->nwide, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error("batched: %s", hipGetErrorString(e));
        return SIL_ERR_HIP;
    }
    return SIL_OK;
}

// a fixed number of iterations of every problem (tolerance 0): equal work
// for every candidate configuration; seconds of the probe
int probe(sil_batched_poisson h, int iterations, double* seconds)
{
    const int replace = 400;
    int rc = h->nsmall ? capture(h, h->block, 0.0, iterations, replace) : SIL_OK;
    if (rc) return rc;
    for (int rep = 0; rep < 2; ++rep) { // a warm-up, then the timed run
        SIL_CHECK_HIP(hipDeviceSynchronize());
        const auto t0 = std::chrono::steady_clock::now();
        if (h->nwide && (rc = solve_wide(h, 0.0, iterations, replace, 0))) return rc;
        if (h->nsmall) SIL_CHECK_HIP(hipGraphLaunch(h->graph, 0));
        SIL_CHECK_HIP(hipDeviceSynchronize());
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
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
    int wide_side = 192;
    if (const char* e = std::getenv("SIL_BATCH_WIDE_SIDE")) wide_side = std::max(3, std::atoi(e));
    int rc;
    if ((rc = dalloc(&h->problems_dev, nproblems)) || (rc = dalloc(&h->queue, 1)) || (rc = dalloc(&h->wx32, off)) ||
        (rc = dalloc(&h->wy32, off)) || (rc = dalloc(&h->r, off)) || (rc = dalloc(&h->p, off)) ||
        (rc = dalloc(&h->q, off)) || (rc = dalloc(&h->wx64, off)) || (rc = dalloc(&h->wy64, off)) ||
        (rc = dalloc(&h->b, off)) || (rc = dalloc(&h->x, off)) || (rc = dalloc(&h->dinv32, off)) ||
        (rc = dalloc(&h->results_dev, nproblems))) {
        release(h);
        return rc;
    }
    // small problems: one workgroup each in the persistent kernel, largest
    // first; large ones: the wide path
    if ((rc = partition(h, wide_side))) {
        release(h);
        return rc;
    }
    hipError_t e = hipMemcpy(h->problems_dev, h->problems.data(), sizeof(BatchProblem) * nproblems,
                             hipMemcpyHostToDevice);
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
    int rc = h->nsmall ? capture(h, h->block, st->relative_tolerance, st->max_iterations, replace) : SIL_OK;
    if (rc) return rc;
    SIL_CHECK_HIP(hipMemcpyAsync(h->b, rhs, sizeof(double) * h->total, hipMemcpyDeviceToDevice, s));
    SIL_CHECK_HIP(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();
    if (h->nwide && (rc = solve_wide(h, st->relative_tolerance, st->max_iterations, replace, s))) return rc;
    if (h->nsmall) SIL_CHECK_HIP(hipGraphLaunch(h->graph, s));
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
    if (h->nsmall == 0) {
        // every problem takes the wide path: nothing to tune
        *chosen_block = h->block;
        if (seconds_per_candidate)
            for (int c = 0; c < 3; ++c) seconds_per_candidate[c] = 0.0;
        return SIL_OK;
    }
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

int sil_batched_poisson2d_autotune_shape(sil_batched_poisson h, const double* rhs, int32_t probe_iterations,
                                         sil_batch_shape* chosen)
{
    SIL_REQUIRE(h && rhs && chosen, SIL_ERR_INVALID, "batched_poisson2d_autotune_shape: null argument");
    SIL_REQUIRE(probe_iterations >= 1, SIL_ERR_INVALID, "batched_poisson2d_autotune_shape: probe iterations < 1");
    SIL_CHECK_HIP(hipMemcpy(h->b, rhs, sizeof(double) * h->total, hipMemcpyDeviceToDevice));
    // wide-path threshold, workgroups per wide problem, persistent workgroup size
    const int sides[] = {96, 128, 192, 256, 384, h->max_side + 1};
    sil_batch_shape best{};
    best.seconds = 1e300;
    std::vector<int> seen_splits;
    for (int ws : sides) {
        int nw = 0;
        for (const BatchProblem& p : h->problems) nw += p.side >= ws ? 1 : 0;
        if (std::find(seen_splits.begin(), seen_splits.end(), nw) != seen_splits.end()) continue;
        seen_splits.push_back(nw);
        int rc = partition(h, ws);
        if (rc) return rc;
        const int nbx0 = h->nbx_max;
        const int nbx_c[3] = {nbx0, std::max(1, nbx0 / 2), std::max(1, nbx0 / 4)};
        const int blocks[3] = {256, 512, 1024};
        for (int a = 0; a < (h->nwide ? 3 : 1); ++a) {
            if (a > 0 && nbx_c[a] == nbx_c[a - 1]) continue;
            h->nbx = nbx_c[a];
            if (h->wide_graph) (void)hipGraphExecDestroy(h->wide_graph);
            h->wide_graph = nullptr;
            for (int c = 0; c < (h->nsmall ? 3 : 1); ++c) {
                h->block = blocks[c];
                double dt = 0.0;
                if ((rc = probe(h, probe_iterations, &dt))) return rc;
                if (dt < best.seconds) {
                    best.seconds = dt;
                    best.wide_side = ws;
                    best.wide_workgroups = h->nbx;
                    best.block = h->block;
                    best.wide_problems = h->nwide;
                }
            }
        }
    }
    int rc = partition(h, best.wide_side);
    if (rc) return rc;
    h->nbx = std::min(best.wide_workgroups, h->nbx_max);
    h->block = best.block;
    *chosen = best;
    return SIL_OK;
}

int sil_batched_poisson2d_destroy(sil_batched_poisson h)
{
    release(h);
    return SIL_OK;
}

}  // extern "C"
