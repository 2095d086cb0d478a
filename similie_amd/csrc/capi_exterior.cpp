// extern "C" entry points of the sil::exterior operators.
#include <cstdarg>
#include <cstdlib>
#include <new>

#include "dec_kernels.hpp"

namespace sil {

namespace {
thread_local char g_last_error[1024] = "";
}

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
    va_end(ap);
}

int validate_view(const sil_view* v, const char* what)
{
    SIL_REQUIRE(v != nullptr, SIL_ERR_INVALID, "%s: null view", what);
    SIL_REQUIRE(v->ndim >= 1 && v->ndim <= SIL_MAX_DIM, SIL_ERR_INVALID, "%s: ndim %d outside 1..4", what, v->ndim);
    SIL_REQUIRE(v->nbatch >= 0 && v->nbatch <= SIL_MAX_DIM, SIL_ERR_INVALID, "%s: nbatch %d outside 0..4", what,
                v->nbatch);
    for (int a = 0; a < v->ndim; ++a)
        SIL_REQUIRE(v->extent[a] >= 0, SIL_ERR_INVALID, "%s: negative extent on axis %d", what, a);
    for (int a = 0; a < v->nbatch; ++a)
        SIL_REQUIRE(v->batch_extent[a] >= 0, SIL_ERR_INVALID, "%s: negative batch extent %d", what, a);
    return SIL_OK;
}

int64_t view_nodes(const sil_view& v)
{
    int64_t t = 1;
    for (int a = 0; a < v.ndim; ++a) t *= v.extent[a];
    return t;
}

int64_t view_batches(const sil_view& v)
{
    int64_t t = 1;
    for (int a = 0; a < v.nbatch; ++a) t *= v.batch_extent[a];
    return t;
}

bool view_is_index_last_dense(const sil_view& v, int ncomp)
{
    if (v.nbatch != 0 || v.comp_stride != 1) return false;
    int64_t expect = ncomp;
    for (int a = v.ndim - 1; a >= 0; --a) {
        if (v.extent[a] > 1 && v.stride[a] != expect) return false;
        expect *= v.extent[a];
    }
    return true;
}

}  // namespace sil

using namespace sil;

namespace {

// kout: degree of the output form as a function of the input degree k and
// the dimension (the kernels index C(N, k) input and C(N, kout) output
// components along the component axis)
enum class OutDegree { Up, Same, Down, Dual };

int check_form_args(int k, int n, const void* in, const sil_view* vin, const void* out, const sil_view* vout,
                    const char* what, OutDegree kout)
{
    int rc = validate_view(vin, what);
    if (rc) return rc;
    rc = validate_view(vout, what);
    if (rc) return rc;
    SIL_REQUIRE(vin->ndim == vout->ndim, SIL_ERR_INVALID, "%s: input and output grids differ in dimension", what);
    SIL_REQUIRE(n <= 0 || vin->ndim == n, SIL_ERR_INVALID, "%s: view dimension %d does not match the grid (%d)", what,
                vin->ndim, n);
    SIL_REQUIRE(k >= 0 && k <= vin->ndim, SIL_ERR_INVALID, "%s: degree %d outside 0..%d", what, k, vin->ndim);
    const int dim = vin->ndim;
    const int ko = kout == OutDegree::Up ? k + 1 : kout == OutDegree::Down ? k - 1 : kout == OutDegree::Dual ? dim - k : k;
    const int64_t cin = binom4(dim, k), cout = (ko >= 0 && ko <= dim) ? binom4(dim, ko) : 0;
    SIL_REQUIRE(vin->comp_extent == cin, SIL_ERR_INVALID,
                "%s: input holds %lld components, a %d-form in %d dimensions has %lld", what,
                (long long)vin->comp_extent, k, dim, (long long)cin);
    SIL_REQUIRE(cout == 0 || vout->comp_extent == cout, SIL_ERR_INVALID,
                "%s: output holds %lld components, a %d-form in %d dimensions has %lld", what,
                (long long)vout->comp_extent, ko, dim, (long long)cout);
    bool empty = view_nodes(*vout) * view_batches(*vout) == 0;
    SIL_REQUIRE(empty || (in != nullptr && out != nullptr), SIL_ERR_INVALID, "%s: null array", what);
    return SIL_OK;
}

int check_grid_view(const sil_grid g, const sil_view* v, const char* what)
{
    for (int a = 0; a < g->geom.ndim; ++a)
        SIL_REQUIRE(v->extent[a] == g->geom.ext[a], SIL_ERR_INVALID,
                    "%s: input extent %lld on axis %d differs from the grid's %lld", what, (long long)v->extent[a], a,
                    (long long)g->geom.ext[a]);
    SIL_REQUIRE(v->nbatch == 0, SIL_ERR_UNSUPPORTED, "%s: spectator dimensions are supported by the coboundary only",
                what);
    return SIL_OK;
}

bool fused_capable(const sil_grid g)
{
    return g->geom.rectilinear && g->geom.metric_kind != SIL_METRIC_SYMMETRIC;
}

FusedGeom fused_geom(const sil_grid g)
{
    FusedGeom f{};
    for (int a = 0; a < g->geom.ndim && a < 3; ++a) {
        f.ext[a] = g->geom.ext[a];
        f.spacing[a] = g->spacing[a];
    }
    f.metric = g->geom.metric_kind == SIL_METRIC_DIAGONAL ? g->geom.metric : nullptr;
    return f;
}

}  // namespace

extern "C" {

const char* sil_last_error(void) { return sil::g_last_error; }

int32_t sil_abi_version(void) { return 1; }

int sil_grid_create(sil_grid* out, int32_t ndim, const int64_t* extent, const double* const* axis_coords,
                    const double* nodal_positions, int32_t metric_kind, const double* metric)
{
    SIL_REQUIRE(out != nullptr && extent != nullptr, SIL_ERR_INVALID, "grid_create: null argument");
    SIL_REQUIRE(ndim >= 1 && ndim <= SIL_MAX_DIM, SIL_ERR_INVALID, "grid_create: ndim %d outside 1..4", ndim);
    SIL_REQUIRE((axis_coords != nullptr) != (nodal_positions != nullptr), SIL_ERR_INVALID,
                "grid_create: give exactly one of axis_coords and nodal_positions");
    SIL_REQUIRE(metric_kind >= SIL_METRIC_IDENTITY && metric_kind <= SIL_METRIC_SYMMETRIC, SIL_ERR_INVALID,
                "grid_create: unknown metric kind %d", metric_kind);
    SIL_REQUIRE(metric_kind == SIL_METRIC_IDENTITY || metric != nullptr, SIL_ERR_INVALID,
                "grid_create: metric array required for a non-identity metric");
    for (int a = 0; a < ndim; ++a) SIL_REQUIRE(extent[a] >= 1, SIL_ERR_INVALID, "grid_create: empty axis %d", a);
    auto* g = new (std::nothrow) sil_grid_s();
    SIL_REQUIRE(g != nullptr, SIL_ERR_INVALID, "grid_create: out of host memory");
    g->geom.ndim = ndim;
    g->geom.rectilinear = axis_coords != nullptr;
    g->geom.metric_kind = metric_kind;
    g->geom.metric = metric;
    g->geom.nodal = nodal_positions;
    for (int a = 0; a < ndim; ++a) g->geom.ext[a] = extent[a];
    if (axis_coords) {
        for (int a = 0; a < ndim; ++a) {
            const int64_t n = extent[a];
            const double* c = axis_coords[a];
            if (c == nullptr) {
                sil_grid_destroy(g);
                set_error("grid_create: null coordinate array for axis %d", a);
                return SIL_ERR_INVALID;
            }
            // per-node edge length: forward, backward at the last node, 0 on one node
            double* h = static_cast<double*>(std::malloc(sizeof(double) * n));
            for (int64_t i = 0; i < n; ++i)
                h[i] = (i + 1 < n) ? c[i + 1] - c[i] : (i > 0 ? c[i] - c[i - 1] : 0.0);
            hipError_t e1 = hipMalloc(&g->axis_storage[a], sizeof(double) * n);
            hipError_t e2 = hipMalloc(&g->spacing[a], sizeof(double) * n);
            hipError_t e3 = e1 == hipSuccess ? hipMemcpy(g->axis_storage[a], c, sizeof(double) * n, hipMemcpyHostToDevice)
                                             : e1;
            hipError_t e4 = e2 == hipSuccess ? hipMemcpy(g->spacing[a], h, sizeof(double) * n, hipMemcpyHostToDevice)
                                             : e2;
            std::free(h);
            if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
                sil_grid_destroy(g);
                set_error("grid_create: device allocation or copy failed");
                return SIL_ERR_HIP;
            }
            g->geom.axis[a] = g->axis_storage[a];
        }
    }
    *out = g;
    return SIL_OK;
}

int sil_grid_destroy(sil_grid grid)
{
    if (!grid) return SIL_OK;
    for (int a = 0; a < SIL_MAX_DIM; ++a) {
        if (grid->axis_storage[a]) (void)hipFree(grid->axis_storage[a]);
        if (grid->spacing[a]) (void)hipFree(grid->spacing[a]);
    }
    delete grid;
    return SIL_OK;
}

int sil_coboundary(int32_t k, const double* in, const sil_view* in_view, double* out, const sil_view* out_view,
                   void* stream)
{
    int rc = check_form_args(k, -1, in, in_view, out, out_view, "coboundary", OutDegree::Up);
    if (rc) return rc;
    SIL_REQUIRE(k < in_view->ndim, SIL_ERR_INVALID, "coboundary: degree %d has no coboundary in %d dimensions", k,
                in_view->ndim);
    SIL_REQUIRE(in_view->nbatch == out_view->nbatch, SIL_ERR_INVALID, "coboundary: spectator dimensions differ");
    for (int a = 0; a < in_view->nbatch; ++a)
        SIL_REQUIRE(in_view->batch_extent[a] == out_view->batch_extent[a], SIL_ERR_INVALID,
                    "coboundary: spectator extent %d differs", a);
    // 3D 0-form, dense index-last, same domain: dedicated kernel
    if (k == 0 && in_view->ndim == 3 && view_is_index_last_dense(*in_view, 1) &&
        view_is_index_last_dense(*out_view, 3) && in_view->extent[0] == out_view->extent[0] &&
        in_view->extent[1] == out_view->extent[1] && in_view->extent[2] == out_view->extent[2])
        return launch_coboundary_3d_0form(in, out, in_view->extent, as_stream(stream));
    return launch_coboundary(k, false, in, *in_view, out, *out_view, as_stream(stream));
}

int sil_transposed_coboundary(int32_t k, const double* in, const sil_view* in_view, double* out,
                              const sil_view* out_view, void* stream)
{
    int rc = check_form_args(k, -1, in, in_view, out, out_view, "transposed_coboundary", OutDegree::Up);
    if (rc) return rc;
    SIL_REQUIRE(k < in_view->ndim, SIL_ERR_INVALID, "transposed_coboundary: degree %d too large", k);
    SIL_REQUIRE(in_view->nbatch == out_view->nbatch, SIL_ERR_INVALID,
                "transposed_coboundary: spectator dimensions differ");
    for (int a = 0; a < in_view->nbatch; ++a)
        SIL_REQUIRE(in_view->batch_extent[a] == out_view->batch_extent[a], SIL_ERR_INVALID,
                    "transposed_coboundary: spectator extent %d differs", a);
    return launch_coboundary(k, true, in, *in_view, out, *out_view, as_stream(stream));
}

int sil_fill_discrete_hodge_star(sil_grid grid, int32_t k, double* hodge, void* stream)
{
    SIL_REQUIRE(grid != nullptr && hodge != nullptr, SIL_ERR_INVALID, "fill_discrete_hodge_star: null argument");
    SIL_REQUIRE(k >= 0 && k <= grid->geom.ndim, SIL_ERR_INVALID, "fill_discrete_hodge_star: degree %d invalid", k);
    return launch_fill_hodge(grid->geom, k, hodge, as_stream(stream));
}

int sil_hodge_star(sil_grid grid, int32_t k, const double* in, const sil_view* in_view, double* out,
                   const sil_view* out_view, void* stream)
{
    SIL_REQUIRE(grid != nullptr, SIL_ERR_INVALID, "hodge_star: null grid");
    int rc = check_form_args(k, grid->geom.ndim, in, in_view, out, out_view, "hodge_star", OutDegree::Dual);
    if (rc) return rc;
    if ((rc = check_grid_view(grid, in_view, "hodge_star"))) return rc;
    if ((rc = check_grid_view(grid, out_view, "hodge_star"))) return rc;
    return launch_hodge_apply(grid->geom, k, in, *in_view, out, *out_view, as_stream(stream));
}

int sil_codifferential(sil_grid grid, int32_t k, const double* in, const sil_view* in_view, double* out,
                       const sil_view* out_view, void* stream)
{
    SIL_REQUIRE(grid != nullptr, SIL_ERR_INVALID, "codifferential: null grid");
    int rc = check_form_args(k, grid->geom.ndim, in, in_view, out, out_view, "codifferential", OutDegree::Down);
    if (rc) return rc;
    SIL_REQUIRE(k >= 1, SIL_ERR_INVALID, "codifferential: degree must be at least 1");
    if ((rc = check_grid_view(grid, in_view, "codifferential"))) return rc;
    SIL_REQUIRE(out_view->nbatch == 0, SIL_ERR_UNSUPPORTED, "codifferential: spectator dimensions unsupported");
    for (int a = 0; a < grid->geom.ndim; ++a)
        SIL_REQUIRE(out_view->extent[a] <= grid->geom.ext[a], SIL_ERR_INVALID,
                    "codifferential: output extent exceeds the grid on axis %d", a);
    return launch_codifferential(grid->geom, k, in, *in_view, out, *out_view, as_stream(stream));
}

}  // extern "C"

struct sil_staged_codifferential_s {
    sil_grid grid;
    int k;
    sil_view in_view, out_view;
};

extern "C" {

int sil_make_staged_codifferential(sil_staged_codifferential* out, sil_grid grid, int32_t k, const sil_view* in_view,
                                   const sil_view* out_view)
{
    SIL_REQUIRE(out != nullptr && grid != nullptr, SIL_ERR_INVALID, "codifferential: null argument");
    int rc = check_form_args(k, grid->geom.ndim, (void*)1, in_view, (void*)1, out_view, "codifferential",
                             OutDegree::Down);
    if (rc) return rc;
    SIL_REQUIRE(k >= 1, SIL_ERR_INVALID, "codifferential: degree must be at least 1");
    if ((rc = check_grid_view(grid, in_view, "codifferential"))) return rc;
    SIL_REQUIRE(out_view->nbatch == 0, SIL_ERR_UNSUPPORTED, "codifferential: spectator dimensions unsupported");
    for (int a = 0; a < grid->geom.ndim; ++a)
        SIL_REQUIRE(out_view->extent[a] <= grid->geom.ext[a], SIL_ERR_INVALID,
                    "codifferential: output extent exceeds the grid on axis %d", a);
    auto* op = new (std::nothrow) sil_staged_codifferential_s();
    SIL_REQUIRE(op != nullptr, SIL_ERR_INVALID, "codifferential: out of host memory");
    op->grid = grid;
    op->k = k;
    op->in_view = *in_view;
    op->out_view = *out_view;
    *out = op;
    return SIL_OK;
}

int sil_staged_codifferential_run(sil_staged_codifferential op, double* out, const double* in, void* stream)
{
    SIL_REQUIRE(op != nullptr, SIL_ERR_INVALID, "codifferential: null staged operator");
    if (view_nodes(op->out_view) == 0) return SIL_OK;
    SIL_REQUIRE(in != nullptr && out != nullptr, SIL_ERR_INVALID, "codifferential: null array");
    return launch_codifferential(op->grid->geom, op->k, in, op->in_view, out, op->out_view, as_stream(stream));
}

int sil_staged_codifferential_destroy(sil_staged_codifferential op)
{
    delete op;
    return SIL_OK;
}

int sil_laplacian(sil_grid grid, int32_t k, const double* in, const sil_view* in_view, double* out,
                  const sil_view* out_view, void* stream)
{
    sil_staged_laplacian op = nullptr;
    int rc = sil_make_staged_laplacian(&op, grid, k, in_view, out_view);
    if (rc) return rc;
    rc = sil_staged_laplacian_run(op, out, in, stream);
    sil_staged_laplacian_destroy(op);
    return rc;
}

}  // extern "C"

struct sil_staged_laplacian_s {
    sil_grid grid;
    int k;
    sil_view in_view, out_view;
    int path; // 0 generic, 1 fused 2D 0-form, 2 fused 3D 1-form
    FusedTile tile;
};

namespace sil {
// Tile configuration of the fused 3D Laplacian; the environment overrides
// it (SIL_LAP3D_TILE="variant,ty,xchunk[,zw or planes in flight]").
FusedTile default_lap3d_tile()
{
    // staged march, 16-row tiles, two planes of loads in flight, 32 planes
    // per workgroup, non-temporal output stores: the fastest measured shape
    FusedTile t{FUSED_STAGE, 2, 16, 32, 1};
    if (const char* e = std::getenv("SIL_LAP3D_TILE")) {
        int v = 1, ty = 8, xc = 32, zw = 64;
        const int n = std::sscanf(e, "%d,%d,%d,%d", &v, &ty, &xc, &zw);
        if (n >= 3) {
            t.variant = v;
            t.ty = ty;
            t.xchunk = xc;
            t.tz = n == 4 ? zw : 64;
        }
    }
    if (const char* e = std::getenv("SIL_LAP3D_NT")) t.nt = std::atoi(e) ? 1 : 0;
    return t;
}
}  // namespace sil

extern "C" {

int sil_make_staged_laplacian(sil_staged_laplacian* out, sil_grid grid, int32_t k, const sil_view* in_view,
                              const sil_view* out_view)
{
    SIL_REQUIRE(out != nullptr && grid != nullptr, SIL_ERR_INVALID, "laplacian: null argument");
    int rc = check_form_args(k, grid->geom.ndim, (void*)1, in_view, (void*)1, out_view, "laplacian", OutDegree::Same);
    if (rc) return rc;
    if ((rc = check_grid_view(grid, in_view, "laplacian"))) return rc;
    SIL_REQUIRE(out_view->nbatch == 0, SIL_ERR_UNSUPPORTED, "laplacian: spectator dimensions unsupported");
    for (int a = 0; a < grid->geom.ndim; ++a)
        SIL_REQUIRE(out_view->extent[a] <= grid->geom.ext[a], SIL_ERR_INVALID,
                    "laplacian: output extent exceeds the grid on axis %d", a);
    auto* op = new (std::nothrow) sil_staged_laplacian_s();
    SIL_REQUIRE(op != nullptr, SIL_ERR_INVALID, "laplacian: out of host memory");
    op->grid = grid;
    op->k = k;
    op->in_view = *in_view;
    op->out_view = *out_view;
    op->path = 0;
    const int n = grid->geom.ndim;
    const int nk = static_cast<int>(binom4(n, k));
    if (fused_capable(grid) && view_is_index_last_dense(*in_view, nk) && view_is_index_last_dense(*out_view, nk)) {
        if (n == 2 && k == 0) op->path = 1;
        if (n == 3 && k == 1) op->path = 2;
    }
    op->tile = default_lap3d_tile();
    *out = op;
    return SIL_OK;
}

int sil_staged_laplacian_run(sil_staged_laplacian op, double* out, const double* in, void* stream)
{
    SIL_REQUIRE(op != nullptr, SIL_ERR_INVALID, "laplacian: null staged operator");
    if (view_nodes(op->out_view) == 0) return SIL_OK;
    SIL_REQUIRE(in != nullptr && out != nullptr, SIL_ERR_INVALID, "laplacian: null array");
    hipStream_t s = as_stream(stream);
    switch (op->path) {
    case 1:
        return launch_laplacian_2d_0form(fused_geom(op->grid), in, out, op->out_view.extent[0],
                                         op->out_view.extent[1], s);
    case 2:
        return launch_laplacian_3d_1form(fused_geom(op->grid), in, out, op->out_view.extent, op->tile, s);
    default:
        return launch_laplacian_generic(op->grid->geom, op->k, in, op->in_view, out, op->out_view, s);
    }
}

int sil_staged_laplacian_destroy(sil_staged_laplacian op)
{
    delete op;
    return SIL_OK;
}

int sil_staged_laplacian_autotune(sil_staged_laplacian op, double* out, const double* in, int32_t repeats,
                                  sil_lap3d_tile* chosen, void* stream)
{
    SIL_REQUIRE(op != nullptr && out != nullptr && in != nullptr, SIL_ERR_INVALID, "laplacian autotune: null argument");
    SIL_REQUIRE(repeats >= 1, SIL_ERR_INVALID, "laplacian autotune: repeats must be at least 1");
    auto report = [&](const FusedTile& t, double us) {
        if (!chosen) return;
        chosen->variant = t.variant;
        chosen->rows = t.ty;
        chosen->xchunk = t.xchunk;
        chosen->lanes_or_depth = t.tz;
        chosen->nontemporal = t.nt;
        chosen->microseconds = us;
    };
    // only the fused 3D 1-form path has shapes to choose; the environment's
    // explicit tile is kept
    if (op->path != 2 || std::getenv("SIL_LAP3D_TILE") != nullptr) {
        report(op->tile, 0.0);
        return SIL_OK;
    }
    hipStream_t s = as_stream(stream);
    // staged x-march (z neighbours by DPP): rows, planes in flight, planes per
    // workgroup; pipelined x-march: rows of 64 lanes
    const FusedTile candidates[] = {
        {FUSED_STAGE, 2, 16, 32, 1}, {FUSED_STAGE, 1, 16, 32, 1}, {FUSED_STAGE, 2, 12, 32, 1},
        {FUSED_STAGE, 2, 8, 32, 1},  {FUSED_STAGE, 2, 16, 24, 1}, {FUSED_STAGE, 2, 16, 40, 1},
        {FUSED_STAGE, 2, 16, 32, 0}, {FUSED_PIPE, 64, 12, 32, 1}, {FUSED_PIPE, 64, 16, 32, 0},
        {FUSED_PIPE, 64, 8, 32, 0},
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SIL_CHECK_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error("laplacian autotune: event creation failed");
        return SIL_ERR_HIP;
    }
    const FusedGeom g = fused_geom(op->grid);
    FusedTile best = op->tile;
    double best_us = 0.0;
    int rc = SIL_OK;
    for (const FusedTile& t : candidates) {
        if (launch_laplacian_3d_1form(g, in, out, op->out_view.extent, t, s) != SIL_OK) continue; // warm-up
        if (hipEventRecord(e0, s) != hipSuccess) {
            rc = SIL_ERR_HIP;
            break;
        }
        int ok = SIL_OK;
        for (int r = 0; r < repeats && ok == SIL_OK; ++r) ok = launch_laplacian_3d_1form(g, in, out, op->out_view.extent, t, s);
        float ms = 0.f;
        if (ok != SIL_OK || hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = SIL_ERR_HIP;
            break;
        }
        const double us = 1000.0 * ms / repeats;
        if (best_us == 0.0 || us < best_us) {
            best_us = us;
            best = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != SIL_OK) {
        set_error("laplacian autotune: timing failed");
        return rc;
    }
    op->tile = best;
    report(best, best_us);
    // the output holds the chosen kernel's result
    return launch_laplacian_3d_1form(g, in, out, op->out_view.extent, best, s);
}

}  // extern "C"
