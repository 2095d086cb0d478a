// extern "C" entry points of the linear operators and the CG solve.
#include <cmath>
#include <cstdlib>
#include <new>
#include <vector>

#include "operators.hpp"

namespace sil {
struct Workspace;
void destroy_workspace(Workspace* w);
int cg_solve(sil_operator op, const double* rhs, double* solution, const sil_solver_settings& st,
             sil_solver_diagnostics& dg, hipStream_t s);
int cg_solve_noalloc(sil_operator op, const double* rhs, double* solution, const sil_solver_settings& st,
                     sil_solver_diagnostics& dg, hipStream_t s);
int64_t noalloc_max_dofs();
int operator_apply64(sil_operator op, const double* in, double* out, hipStream_t s);
int operator_apply32(sil_operator op, const float* in, float* out, hipStream_t s);
int operator_diagonal(sil_operator op, double* d, hipStream_t s);
int newton_solve(sil_operator op, const double* rhs, double* x, const sil_solver_settings& st,
                 sil_solver_diagnostics& dg, hipStream_t s);
void destroy_newton(NewtonBuffers* nb);
int ms3d_ferro_flags(const double* frac, uint8_t* flags, int64_t n, hipStream_t s);
}  // namespace sil

using namespace sil;

namespace {

template <typename T>
int upload(T** dst, const std::vector<T>& src)
{
    SIL_CHECK_HIP(hipMalloc(reinterpret_cast<void**>(dst), sizeof(T) * src.size()));
    SIL_CHECK_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return SIL_OK;
}

// local_spacing (onelab_interface/magnetostatics_onelab.hpp:856-870)
double local_spacing(const double* c, int64_t n, int64_t i)
{
    if (n < 2) return 1.0;
    if (i == 0) return c[1] - c[0];
    if (i + 1 == n) return c[n - 1] - c[n - 2];
    return 0.5 * (c[i + 1] - c[i - 1]);
}

void free_operator(sil_operator op)
{
    if (!op) return;
    destroy_workspace(op->ws);
    auto f = [](void* p) {
        if (p) (void)hipFree(p);
    };
    for (int q = 0; q < 3; ++q) {
        f(op->ms.ls[q]);
        f(op->ms.inv_ls[q]);
        f(op->ms.coord[q]);
        f(op->ms.ls32[q]);
        f(op->ms.inv_ls32[q]);
    }
    f(op->ms.nu);
    f(op->ms.nu32);
    for (int q = 0; q < 2; ++q) f(op->m2.coord[q]);
    f(op->m2.nu);
    f(op->m2.nu32);
    f(op->p2.wx);
    f(op->p2.wy);
    f(op->p2.wx32);
    f(op->p2.wy32);
    csr_free(op->csr);
    f(const_cast<BhCurve*>(op->nl.curve));
    f(op->nl.ferro);
    f(op->nl.st_a);
    f(op->nl.st_c);
    for (int q = 0; q < 3; ++q) f(op->nl.st_b[q]);
    f(op->nl.st_a32);
    f(op->nl.st_c32);
    for (int q = 0; q < 3; ++q) f(op->nl.st_b32[q]);
    f(op->nl.st_pack);
    f(op->nl.st_pack32);
    if (op->jac) {
        // shares ms / nl with op: only its own solver workspace is freed
        destroy_workspace(op->jac->ws);
        delete op->jac;
    }
    destroy_newton(op->newton);
    delete op;
}

// a new operator owning an assembled matrix
int wrap_csr(const CsrData& m, sil_operator* out)
{
    auto* op = new (std::nothrow) sil_operator_s();
    if (!op) {
        CsrData tmp = m;
        csr_free(tmp);
        set_error("out of host memory");
        return SIL_ERR_INVALID;
    }
    op->kind = OP_CSR;
    op->n = m.n;
    op->csr = m;
    *out = op;
    return SIL_OK;
}

// InterpolatedNonlinearBHCurve's preconditions (validate_nonlinear_bh_curve)
int check_curve(const double* bh_b, int32_t nsamples)
{
    SIL_REQUIRE(nsamples >= 2 && nsamples <= kBhMax, SIL_ERR_INVALID,
                "invalid nonlinear B-H curve sample count %d (2..%d)", nsamples, kBhMax);
    for (int i = 0; i + 1 < nsamples; ++i)
        SIL_REQUIRE(bh_b[i + 1] > bh_b[i], SIL_ERR_INVALID, "B-H curve: induction samples must increase");
    return SIL_OK;
}

// the B-H curve, the ferromagnetic flags, the state arrays and the Jacobian
// operator of a nonlinear operator over nn nodes (InterpolatedNonlinearBHCurve,
// nonlinear_magnetostatics.py:26-55); on failure the caller frees op
int attach_material(sil_operator op, int64_t nn, const double* ferromagnetic_nodal, const double* bh_b,
                    const double* bh_h, int32_t nsamples, bool three_d)
{
    BhCurve c{};
    c.n = nsamples;
    for (int i = 0; i < nsamples; ++i) {
        c.b[i] = bh_b[i];
        c.h[i] = bh_h[i];
        c.q[i] = bh_b[i] * bh_b[i];
    }
    for (int i = 0; i + 1 < nsamples; ++i) c.dh_db[i] = (c.h[i + 1] - c.h[i]) / (c.b[i + 1] - c.b[i]);
    BhCurve* dc = nullptr;
    NlData& nl = op->nl;
    bool ok = hipMalloc(&dc, sizeof(BhCurve)) == hipSuccess;
    nl.curve = dc;
    ok = ok && hipMemcpy(dc, &c, sizeof(BhCurve), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMalloc(&nl.ferro, nn) == hipSuccess;
    if (three_d) {
        // per-field state: FP64 for the Jacobian diagonal and assembly (and
        // B_z for the FP64 Jacobian march), a and B_z in FP32 for the FP32
        // marches; the 2D operator reads only its packed record
        ok = ok && hipMalloc(&nl.st_a, sizeof(double) * nn) == hipSuccess;
        ok = ok && hipMalloc(&nl.st_c, sizeof(double) * nn) == hipSuccess;
        for (int q = 0; q < 3; ++q) ok = ok && hipMalloc(&nl.st_b[q], sizeof(double) * nn) == hipSuccess;
        ok = ok && hipMalloc(&nl.st_a32, sizeof(float) * nn) == hipSuccess;
        ok = ok && hipMalloc(&nl.st_b32[2], sizeof(float) * nn) == hipSuccess;
    }
    if (!ok) {
        set_error("nonlinear_create: device allocation failed");
        return SIL_ERR_HIP;
    }
    int rc = ms3d_ferro_flags(ferromagnetic_nodal, nl.ferro, nn, 0);
    if (rc == SIL_OK && hipDeviceSynchronize() != hipSuccess) {
        set_error("nonlinear_create: flag kernel failed");
        rc = SIL_ERR_HIP;
    }
    op->jac = rc == SIL_OK ? new (std::nothrow) sil_operator_s() : nullptr;
    if (rc == SIL_OK && !op->jac) {
        set_error("nonlinear_create: out of host memory");
        rc = SIL_ERR_INVALID;
    }
    return rc;
}

}  // namespace

extern "C" {

int sil_magnetostatics3d_create(sil_operator* out, int64_t nx, int64_t ny, int64_t nz, const double* x_coords,
                                const double* y_coords, const double* z_coords, const double* mu_nodal,
                                double gauge_penalty)
{
    SIL_REQUIRE(out && x_coords && y_coords && z_coords && mu_nodal, SIL_ERR_INVALID,
                "magnetostatics3d_create: null argument");
    SIL_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, SIL_ERR_INVALID, "magnetostatics3d_create: need ≥ 2 nodes per axis");
    auto* op = new (std::nothrow) sil_operator_s();
    SIL_REQUIRE(op != nullptr, SIL_ERR_INVALID, "magnetostatics3d_create: out of host memory");
    op->kind = OP_MAGNETOSTATICS3D;
    op->n = 3 * nx * ny * nz;
    Ms3dData& d = op->ms;
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    d.penalty = gauge_penalty;
    const double* cs[3] = {x_coords, y_coords, z_coords};
    const int64_t ns[3] = {nx, ny, nz};
    int rc = SIL_OK;
    for (int q = 0; q < 3 && rc == SIL_OK; ++q) {
        std::vector<double> ls(ns[q]), inv(ns[q]), co(cs[q], cs[q] + ns[q]);
        std::vector<float> ls32(ns[q]), inv32(ns[q]);
        for (int64_t i = 0; i < ns[q]; ++i) {
            ls[i] = local_spacing(cs[q], ns[q], i);
            inv[i] = 1.0 / ls[i];
            ls32[i] = (float)ls[i];
            inv32[i] = (float)inv[i];
        }
        if ((rc = upload(&d.ls[q], ls)) || (rc = upload(&d.inv_ls[q], inv)) || (rc = upload(&d.coord[q], co)) ||
            (rc = upload(&d.ls32[q], ls32)) || (rc = upload(&d.inv_ls32[q], inv32)))
            break;
    }
    const int64_t nn = nx * ny * nz;
    if (rc == SIL_OK) {
        hipError_t e1 = hipMalloc(&d.nu, sizeof(double) * nn);
        hipError_t e2 = hipMalloc(&d.nu32, sizeof(float) * nn);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            set_error("magnetostatics3d_create: device allocation failed");
            rc = SIL_ERR_HIP;
        } else {
            rc = ms3d_nu_from_mu(mu_nodal, d.nu, d.nu32, nn, 0);
            hipError_t e3 = hipDeviceSynchronize();
            if (rc == SIL_OK && e3 != hipSuccess) {
                set_error("magnetostatics3d_create: %s", hipGetErrorString(e3));
                rc = SIL_ERR_HIP;
            }
        }
    }
    if (rc != SIL_OK) {
        free_operator(op);
        return rc;
    }
    *out = op;
    return SIL_OK;
}

int sil_magnetostatics3d_nonlinear_create(sil_operator* out, int64_t nx, int64_t ny, int64_t nz,
                                          const double* x_coords, const double* y_coords, const double* z_coords,
                                          const double* mu_nodal, const double* ferromagnetic_nodal,
                                          const double* bh_b, const double* bh_h, int32_t nsamples,
                                          double gauge_penalty)
{
    SIL_REQUIRE(out && ferromagnetic_nodal && bh_b && bh_h, SIL_ERR_INVALID,
                "magnetostatics3d_nonlinear_create: null argument");
    int rc = check_curve(bh_b, nsamples);
    if (rc) return rc;
    SIL_REQUIRE(nx >= 3 && ny >= 3 && nz >= 3, SIL_ERR_INVALID,
                "magnetostatics3d_nonlinear_create: need >= 3 nodes per axis");
    sil_operator op = nullptr;
    rc = sil_magnetostatics3d_create(&op, nx, ny, nz, x_coords, y_coords, z_coords, mu_nodal, gauge_penalty);
    if (rc) return rc;
    op->kind = OP_MS3D_NONLINEAR;
    rc = attach_material(op, nx * ny * nz, ferromagnetic_nodal, bh_b, bh_h, nsamples, true);
    if (rc == SIL_OK && (hipMalloc(&op->nl.st_pack, sizeof(double) * 4 * nx * ny * nz) != hipSuccess ||
                         hipMalloc(&op->nl.st_pack32, sizeof(float) * 4 * nx * ny * nz) != hipSuccess)) {
        set_error("magnetostatics3d_nonlinear_create: device allocation failed");
        rc = SIL_ERR_HIP;
    }
    if (rc != SIL_OK) {
        free_operator(op);
        return rc;
    }
    op->jac->kind = OP_MS3D_JACOBIAN;
    op->jac->n = op->n;
    op->jac->ms = op->ms;
    op->jac->nl = op->nl;
    *out = op;
    return SIL_OK;
}

int sil_ms3d_switches_from_environment(sil_ms3d_switches* out)
{
    SIL_REQUIRE(out, SIL_ERR_INVALID, "ms3d_switches_from_environment: null argument");
    auto sign_of = [](const char* name) {
        const char* v = std::getenv(name);
        if (v == nullptr || v[0] == '\0') return 1.0;
        char* end = nullptr;
        const double parsed = std::strtod(v, &end);
        if (end == v || parsed == 0.0) return 1.0;
        return parsed < 0.0 ? -1.0 : 1.0;
    };
    const char* g = std::getenv("SIMILIE_DIVERGENCE_GAUGE_3D");
    out->y_response_sign = sign_of("SIMILIE_MAGNETIC_Y_RESPONSE_SIGN_3D");
    out->divergence_gauge = (g != nullptr && g[0] != '\0' && g[0] != '0') ? 1 : 0;
    out->gauge_sign = sign_of("SIMILIE_VECTOR_POTENTIAL_GAUGE_SIGN_3D");
    out->criterion = 1;
    return SIL_OK;
}

int sil_magnetostatics3d_set_switches(sil_operator op, const sil_ms3d_switches* sw)
{
    SIL_REQUIRE(op && sw, SIL_ERR_INVALID, "magnetostatics3d_set_switches: null argument");
    SIL_REQUIRE(op->kind == OP_MAGNETOSTATICS3D || op->kind == OP_MS3D_NONLINEAR, SIL_ERR_INVALID,
                "magnetostatics3d_set_switches: not a 3D magnetostatics operator");
    SIL_REQUIRE(sw->criterion >= 0 && sw->criterion <= 2, SIL_ERR_INVALID,
                "magnetostatics3d_set_switches: unknown criterion %d", sw->criterion);
    Ms3dData& d = op->ms;
    d.ysign = sw->y_response_sign < 0.0 ? -1.0 : 1.0;
    d.divg = sw->divergence_gauge != 0 ? 1 : 0;
    d.gauge_factor = d.penalty * (sw->gauge_sign < 0.0 ? -1.0 : 1.0);
    d.criterion = sw->criterion;
    // cached diagonals, factors and captured graphs hold the old operator
    destroy_workspace(op->ws);
    op->ws = nullptr;
    if (op->jac) {
        op->jac->ms = d;
        destroy_workspace(op->jac->ws);
        op->jac->ws = nullptr;
    }
    return SIL_OK;
}

int sil_magnetostatics3d_apply_jacobian(sil_operator op, const double* state, const double* in, double* out,
                                        void* stream)
{
    SIL_REQUIRE(op && state && in && out, SIL_ERR_INVALID, "magnetostatics3d_apply_jacobian: null argument");
    SIL_REQUIRE(op->kind == OP_MS3D_NONLINEAR, SIL_ERR_INVALID, "apply_jacobian: not a nonlinear operator");
    hipStream_t s = as_stream(stream);
    int rc = ms3d_state(op->ms, op->nl, state, s);
    if (rc) return rc;
    return ms3d_jacobian_apply64(op->ms, op->nl, in, out, DotEpilogue{nullptr}, nullptr, nullptr, s);
}

int sil_magnetostatics3d_jacobian_diagonal(sil_operator op, const double* state, double* diag, void* stream)
{
    SIL_REQUIRE(op && state && diag, SIL_ERR_INVALID, "magnetostatics3d_jacobian_diagonal: null argument");
    SIL_REQUIRE(op->kind == OP_MS3D_NONLINEAR, SIL_ERR_INVALID, "jacobian_diagonal: not a nonlinear operator");
    hipStream_t s = as_stream(stream);
    int rc = ms3d_state(op->ms, op->nl, state, s);
    if (rc) return rc;
    return ms3d_jacobian_diagonal(op->ms, op->nl, diag, s);
}

int sil_magnetostatics3d_post_process(sil_operator op, const double* potential, const double* cell_mu,
                                      const double* cell_nonlinear, double* induction, double* field,
                                      double* maxwell_stress, double* force_density, void* stream)
{
    SIL_REQUIRE(op && potential && cell_mu, SIL_ERR_INVALID, "magnetostatics3d_post_process: null argument");
    SIL_REQUIRE(op->kind == OP_MAGNETOSTATICS3D || op->kind == OP_MS3D_NONLINEAR, SIL_ERR_INVALID,
                "magnetostatics3d_post_process: not a magnetostatics operator");
    SIL_REQUIRE(!cell_nonlinear || op->kind == OP_MS3D_NONLINEAR, SIL_ERR_INVALID,
                "magnetostatics3d_post_process: nonlinear cells need a nonlinear operator (its B-H curve)");
    SIL_REQUIRE(op->ms.nx >= 2 && op->ms.ny >= 2 && op->ms.nz >= 2, SIL_ERR_INVALID,
                "magnetostatics3d_post_process: need ≥ 2 nodes per axis");
    const BhCurve* curve = op->kind == OP_MS3D_NONLINEAR ? op->nl.curve : nullptr;
    return ms3d_post_process(op->ms, curve, potential, cell_mu, cell_nonlinear, induction, field, maxwell_stress,
                             force_density, as_stream(stream));
}

int sil_magnetostatics2d_create(sil_operator* out, int64_t nx, int64_t ny, const double* x_coords,
                                const double* y_coords, const double* mu_nodal, int32_t criterion)
{
    SIL_REQUIRE(out && x_coords && y_coords && mu_nodal, SIL_ERR_INVALID, "magnetostatics2d_create: null argument");
    SIL_REQUIRE(nx >= 2 && ny >= 2, SIL_ERR_INVALID, "magnetostatics2d_create: need ≥ 2 nodes per axis");
    SIL_REQUIRE(criterion >= 0 && criterion <= 2, SIL_ERR_INVALID, "magnetostatics2d_create: unknown criterion %d",
                criterion);
    auto* op = new (std::nothrow) sil_operator_s();
    SIL_REQUIRE(op != nullptr, SIL_ERR_INVALID, "magnetostatics2d_create: out of host memory");
    op->kind = OP_MAGNETOSTATICS2D;
    op->n = nx * ny;
    Ms2dData& d = op->m2;
    d.nx = nx;
    d.ny = ny;
    d.criterion = criterion;
    d.terms = criterion == 2 ? 2.0 : 1.0;
    int rc = upload(&d.coord[0], std::vector<double>(x_coords, x_coords + nx));
    if (rc == SIL_OK) rc = upload(&d.coord[1], std::vector<double>(y_coords, y_coords + ny));
    if (rc == SIL_OK) {
        hipError_t e1 = hipMalloc(&d.nu, sizeof(double) * nx * ny);
        hipError_t e2 = hipMalloc(&d.nu32, sizeof(float) * nx * ny);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            set_error("magnetostatics2d_create: device allocation failed");
            rc = SIL_ERR_HIP;
        } else {
            rc = ms3d_nu_from_mu(mu_nodal, d.nu, d.nu32, nx * ny, 0);
            hipError_t e3 = hipDeviceSynchronize();
            if (rc == SIL_OK && e3 != hipSuccess) {
                set_error("magnetostatics2d_create: %s", hipGetErrorString(e3));
                rc = SIL_ERR_HIP;
            }
        }
    }
    if (rc != SIL_OK) {
        free_operator(op);
        return rc;
    }
    *out = op;
    return SIL_OK;
}

int sil_magnetostatics2d_nonlinear_create(sil_operator* out, int64_t nx, int64_t ny, const double* x_coords,
                                          const double* y_coords, const double* mu_nodal,
                                          const double* ferromagnetic_nodal, const double* bh_b, const double* bh_h,
                                          int32_t nsamples, int32_t criterion)
{
    SIL_REQUIRE(out && ferromagnetic_nodal && bh_b && bh_h, SIL_ERR_INVALID,
                "magnetostatics2d_nonlinear_create: null argument");
    int rc = check_curve(bh_b, nsamples);
    if (rc) return rc;
    sil_operator op = nullptr;
    rc = sil_magnetostatics2d_create(&op, nx, ny, x_coords, y_coords, mu_nodal, criterion);
    if (rc) return rc;
    op->kind = OP_MS2D_NONLINEAR;
    rc = attach_material(op, nx * ny, ferromagnetic_nodal, bh_b, bh_h, nsamples, false);
    if (rc == SIL_OK && (hipMalloc(&op->nl.st_pack, sizeof(double) * 4 * nx * ny) != hipSuccess ||
                         hipMalloc(&op->nl.st_pack32, sizeof(float) * 4 * nx * ny) != hipSuccess)) {
        set_error("magnetostatics2d_nonlinear_create: device allocation failed");
        rc = SIL_ERR_HIP;
    }
    if (rc != SIL_OK) {
        free_operator(op);
        return rc;
    }
    op->jac->kind = OP_MS2D_JACOBIAN;
    op->jac->n = op->n;
    op->jac->m2 = op->m2;
    op->jac->nl = op->nl;
    *out = op;
    return SIL_OK;
}

int sil_magnetostatics2d_apply_jacobian(sil_operator op, const double* state, const double* in, double* out,
                                        void* stream)
{
    SIL_REQUIRE(op && state && in && out, SIL_ERR_INVALID, "magnetostatics2d_apply_jacobian: null argument");
    SIL_REQUIRE(op->kind == OP_MS2D_NONLINEAR, SIL_ERR_INVALID, "apply_jacobian: not a 2D nonlinear operator");
    hipStream_t s = as_stream(stream);
    int rc = ms2d_state(op->m2, op->nl, state, s);
    if (rc) return rc;
    return ms2d_jacobian_apply64(op->m2, op->nl, in, out, DotEpilogue{nullptr}, nullptr, nullptr, s);
}

int sil_magnetostatics2d_jacobian_diagonal(sil_operator op, const double* state, double* diag, void* stream)
{
    SIL_REQUIRE(op && state && diag, SIL_ERR_INVALID, "magnetostatics2d_jacobian_diagonal: null argument");
    SIL_REQUIRE(op->kind == OP_MS2D_NONLINEAR, SIL_ERR_INVALID, "jacobian_diagonal: not a 2D nonlinear operator");
    hipStream_t s = as_stream(stream);
    int rc = ms2d_state(op->m2, op->nl, state, s);
    if (rc) return rc;
    return ms2d_jacobian_diagonal(op->m2, op->nl, diag, s);
}

int sil_magnetostatics2d_nodal_mu(int64_t nx, int64_t ny, const double* cell_mu, double mu0, double* mu_nodal,
                                  void* stream)
{
    SIL_REQUIRE(cell_mu && mu_nodal, SIL_ERR_INVALID, "magnetostatics2d_nodal_mu: null argument");
    SIL_REQUIRE(nx >= 2 && ny >= 2, SIL_ERR_INVALID, "magnetostatics2d_nodal_mu: need ≥ 2 nodes per axis");
    return ms2d_nodal_mu(nx, ny, cell_mu, mu0, mu_nodal, as_stream(stream));
}

int sil_magnetostatics2d_rhs(sil_operator op, const double* cell_current_z, double* rhs, void* stream)
{
    SIL_REQUIRE(op && cell_current_z && rhs, SIL_ERR_INVALID, "magnetostatics2d_rhs: null argument");
    SIL_REQUIRE(op->kind == OP_MAGNETOSTATICS2D || op->kind == OP_MS2D_NONLINEAR, SIL_ERR_INVALID,
                "magnetostatics2d_rhs: not a 2D magnetostatics operator");
    return ms2d_rhs(op->m2, cell_current_z, op->m2.criterion == 0, rhs, as_stream(stream));
}

int sil_magnetostatics2d_post_process(sil_operator op, const double* potential, const double* cell_mu,
                                      const double* cell_nonlinear, double* induction, double* field,
                                      double* maxwell_stress, double* force_density, void* stream)
{
    SIL_REQUIRE(op && potential && cell_mu, SIL_ERR_INVALID, "magnetostatics2d_post_process: null argument");
    SIL_REQUIRE(op->kind == OP_MAGNETOSTATICS2D || op->kind == OP_MS2D_NONLINEAR, SIL_ERR_INVALID,
                "magnetostatics2d_post_process: not a 2D magnetostatics operator");
    SIL_REQUIRE(!cell_nonlinear || op->kind == OP_MS2D_NONLINEAR, SIL_ERR_INVALID,
                "magnetostatics2d_post_process: nonlinear cells need a nonlinear operator (its B-H curve)");
    const BhCurve* curve = op->kind == OP_MS2D_NONLINEAR ? op->nl.curve : nullptr;
    return ms2d_post_process(op->m2, curve, potential, cell_mu, cell_nonlinear, induction, field, maxwell_stress,
                             force_density, as_stream(stream));
}

int sil_magnetostatics3d_nodal_mu(int64_t nx, int64_t ny, int64_t nz, const double* cell_mu, double mu0,
                                  double* mu_nodal, void* stream)
{
    SIL_REQUIRE(cell_mu && mu_nodal, SIL_ERR_INVALID, "magnetostatics3d_nodal_mu: null argument");
    SIL_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, SIL_ERR_INVALID, "magnetostatics3d_nodal_mu: need ≥ 2 nodes per axis");
    return ms3d_nodal_mu(nx, ny, nz, cell_mu, mu0, mu_nodal, as_stream(stream));
}

int sil_magnetostatics3d_rhs(sil_operator op, const double* cell_current, double* rhs, void* stream)
{
    SIL_REQUIRE(op && cell_current && rhs, SIL_ERR_INVALID, "magnetostatics3d_rhs: null argument");
    SIL_REQUIRE(op->kind == OP_MAGNETOSTATICS3D || op->kind == OP_MS3D_NONLINEAR, SIL_ERR_INVALID,
                "magnetostatics3d_rhs: not a magnetostatics operator");
    return ms3d_rhs(op->ms, cell_current, rhs, as_stream(stream));
}

int sil_poisson2d_create(sil_operator* out, int64_t nx, int64_t ny, const double* x_coords, const double* y_coords,
                         const double* kappa)
{
    SIL_REQUIRE(out && x_coords && y_coords, SIL_ERR_INVALID, "poisson2d_create: null argument");
    SIL_REQUIRE(nx >= 3 && ny >= 3, SIL_ERR_INVALID, "poisson2d_create: need ≥ 3 nodes per axis");
    auto* op = new (std::nothrow) sil_operator_s();
    SIL_REQUIRE(op != nullptr, SIL_ERR_INVALID, "poisson2d_create: out of host memory");
    op->kind = OP_POISSON2D;
    op->n = nx * ny;
    Poisson2dData& d = op->p2;
    d.nx = nx;
    d.ny = ny;
    std::vector<double> hx(nx), hy(ny);
    for (int64_t i = 0; i < nx; ++i) hx[i] = i + 1 < nx ? x_coords[i + 1] - x_coords[i] : x_coords[i] - x_coords[i - 1];
    for (int64_t j = 0; j < ny; ++j) hy[j] = j + 1 < ny ? y_coords[j + 1] - y_coords[j] : y_coords[j] - y_coords[j - 1];
    double *hxd = nullptr, *hyd = nullptr;
    int rc = upload(&hxd, hx);
    if (rc == SIL_OK) rc = upload(&hyd, hy);
    const int64_t nn = nx * ny;
    if (rc == SIL_OK) {
        hipError_t e = hipMalloc(&d.wx, sizeof(double) * nn);
        if (e == hipSuccess) e = hipMalloc(&d.wy, sizeof(double) * nn);
        if (e == hipSuccess) e = hipMalloc(&d.wx32, sizeof(float) * nn);
        if (e == hipSuccess) e = hipMalloc(&d.wy32, sizeof(float) * nn);
        if (e != hipSuccess) {
            set_error("poisson2d_create: device allocation failed");
            rc = SIL_ERR_HIP;
        }
    }
    if (rc == SIL_OK) rc = p2d_weights(d, hxd, hyd, kappa, 0);
    if (rc == SIL_OK) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            set_error("poisson2d_create: %s", hipGetErrorString(e));
            rc = SIL_ERR_HIP;
        }
    }
    if (hxd) (void)hipFree(hxd);
    if (hyd) (void)hipFree(hyd);
    if (rc != SIL_OK) {
        free_operator(op);
        return rc;
    }
    *out = op;
    return SIL_OK;
}

int sil_magnetostatics3d_assemble(sil_operator op, sil_operator* csr_out, void* stream)
{
    SIL_REQUIRE(op && csr_out, SIL_ERR_INVALID, "magnetostatics3d_assemble: null argument");
    SIL_REQUIRE(op->kind == OP_MAGNETOSTATICS3D, SIL_ERR_INVALID,
                "magnetostatics3d_assemble: not a magnetostatics operator");
    hipStream_t s = as_stream(stream);
    CsrData m{};
    int rc = csr_assemble_ms3d(op->ms, m, s);
    if (rc) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) {
        csr_free(m);
        set_error("magnetostatics3d_assemble: assembly kernels failed");
        return SIL_ERR_HIP;
    }
    return wrap_csr(m, csr_out);
}

int sil_magnetostatics3d_assemble_jacobian(sil_operator op, const double* state, sil_operator* csr_out, void* stream)
{
    SIL_REQUIRE(op && state && csr_out, SIL_ERR_INVALID, "magnetostatics3d_assemble_jacobian: null argument");
    SIL_REQUIRE(op->kind == OP_MS3D_NONLINEAR, SIL_ERR_INVALID,
                "magnetostatics3d_assemble_jacobian: not a nonlinear magnetostatics operator");
    hipStream_t s = as_stream(stream);
    int rc = ms3d_state(op->ms, op->nl, state, s);
    if (rc) return rc;
    CsrData m{};
    if ((rc = csr_assemble_ms3d_jacobian(op->ms, op->nl, m, s))) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) {
        csr_free(m);
        set_error("magnetostatics3d_assemble_jacobian: assembly kernels failed");
        return SIL_ERR_HIP;
    }
    return wrap_csr(m, csr_out);
}

int sil_magnetostatics2d_assemble(sil_operator op, sil_operator* csr_out, void* stream)
{
    SIL_REQUIRE(op && csr_out, SIL_ERR_INVALID, "magnetostatics2d_assemble: null argument");
    SIL_REQUIRE(op->kind == OP_MAGNETOSTATICS2D, SIL_ERR_INVALID,
                "magnetostatics2d_assemble: not a linear 2D magnetostatics operator");
    hipStream_t s = as_stream(stream);
    CsrData m{};
    int rc = csr_assemble_ms2d(op->m2, m, s);
    if (rc) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) {
        csr_free(m);
        set_error("magnetostatics2d_assemble: assembly kernels failed");
        return SIL_ERR_HIP;
    }
    return wrap_csr(m, csr_out);
}

int sil_magnetostatics2d_assemble_jacobian(sil_operator op, const double* state, sil_operator* csr_out, void* stream)
{
    SIL_REQUIRE(op && state && csr_out, SIL_ERR_INVALID, "magnetostatics2d_assemble_jacobian: null argument");
    SIL_REQUIRE(op->kind == OP_MS2D_NONLINEAR, SIL_ERR_INVALID,
                "magnetostatics2d_assemble_jacobian: not a nonlinear 2D magnetostatics operator");
    hipStream_t s = as_stream(stream);
    int rc = ms2d_state(op->m2, op->nl, state, s);
    if (rc) return rc;
    CsrData m{};
    if ((rc = csr_assemble_ms2d_jacobian(op->m2, op->nl, m, s))) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) {
        csr_free(m);
        set_error("magnetostatics2d_assemble_jacobian: assembly kernels failed");
        return SIL_ERR_HIP;
    }
    return wrap_csr(m, csr_out);
}

int sil_csr_create(sil_operator* out, int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* cols,
                   const double* vals, void* stream)
{
    SIL_REQUIRE(out && row_ptr && (nnz == 0 || (cols && vals)), SIL_ERR_INVALID, "csr_create: null argument");
    CsrData m{};
    int rc = csr_import(n, nnz, row_ptr, cols, vals, m, as_stream(stream));
    if (rc) return rc;
    return wrap_csr(m, out);
}

int64_t sil_csr_nnz(sil_operator op) { return op && op->kind == OP_CSR ? op->csr.nnz : -1; }

int sil_csr_copy(sil_operator op, int32_t* row_ptr, int32_t* cols, double* vals, void* stream)
{
    SIL_REQUIRE(op && row_ptr && cols && vals, SIL_ERR_INVALID, "csr_copy: null argument");
    SIL_REQUIRE(op->kind == OP_CSR, SIL_ERR_INVALID, "csr_copy: not an assembled operator");
    hipStream_t s = as_stream(stream);
    const CsrData& m = op->csr;
    SIL_CHECK_HIP(hipMemcpyAsync(row_ptr, m.row_ptr, sizeof(int32_t) * (m.n + 1), hipMemcpyDefault, s));
    if (m.nnz > 0) {
        SIL_CHECK_HIP(hipMemcpyAsync(cols, m.col, sizeof(int32_t) * m.nnz, hipMemcpyDefault, s));
        SIL_CHECK_HIP(hipMemcpyAsync(vals, m.val, sizeof(double) * m.nnz, hipMemcpyDefault, s));
    }
    SIL_CHECK_HIP(hipStreamSynchronize(s));
    return SIL_OK;
}

int sil_spd_isai_create(sil_operator a, sil_operator* g_out, void* stream)
{
    SIL_REQUIRE(a && g_out, SIL_ERR_INVALID, "spd_isai_create: null argument");
    SIL_REQUIRE(a->kind == OP_CSR || a->kind == OP_MAGNETOSTATICS3D || a->kind == OP_MAGNETOSTATICS2D,
                SIL_ERR_INVALID, "spd_isai_create: needs an assembled operator or a linear magnetostatics operator");
    hipStream_t s = as_stream(stream);
    CsrData assembled{};
    const CsrData* m = &a->csr;
    int rc;
    if (a->kind == OP_MAGNETOSTATICS3D) {
        if ((rc = csr_assemble_ms3d(a->ms, assembled, s))) return rc;
        m = &assembled;
    } else if (a->kind == OP_MAGNETOSTATICS2D) {
        if ((rc = csr_assemble_ms2d(a->m2, assembled, s))) return rc;
        m = &assembled;
    }
    CsrData g{}, gt{};
    rc = spd_isai_generate(*m, g, gt, s);
    csr_free(assembled);
    if (rc) return rc;
    csr_free(gt);
    return wrap_csr(g, g_out);
}

int64_t sil_operator_size(sil_operator op) { return op ? op->n : -1; }

int sil_operator_apply(sil_operator op, const double* in, double* out, void* stream)
{
    SIL_REQUIRE(op && in && out, SIL_ERR_INVALID, "operator_apply: null argument");
    if (op->kind == OP_MS3D_NONLINEAR) {
        // F(in); the operator keeps in's state (a later Jacobian is taken there)
        hipStream_t s = as_stream(stream);
        int rc = ms3d_state(op->ms, op->nl, in, s);
        if (rc) return rc;
        return ms3d_state_apply64(op->ms, op->nl, in, out, s);
    }
    if (op->kind == OP_MS2D_NONLINEAR) {
        hipStream_t s = as_stream(stream);
        int rc = ms2d_state(op->m2, op->nl, in, s);
        if (rc) return rc;
        return ms2d_state_apply64(op->m2, op->nl, in, out, s);
    }
    return operator_apply64(op, in, out, as_stream(stream));
}

int sil_operator_apply_f32(sil_operator op, const float* in, float* out, void* stream)
{
    SIL_REQUIRE(op && in && out, SIL_ERR_INVALID, "operator_apply_f32: null argument");
    return operator_apply32(op, in, out, as_stream(stream));
}

int sil_operator_diagonal(sil_operator op, double* diag, void* stream)
{
    SIL_REQUIRE(op && diag, SIL_ERR_INVALID, "operator_diagonal: null argument");
    return operator_diagonal(op, diag, as_stream(stream));
}

int sil_operator_destroy(sil_operator op)
{
    free_operator(op);
    return SIL_OK;
}

void sil_default_solver_settings(sil_solver_settings* s)
{
    if (!s) return;
    // StrongFormulationSolverSettings defaults (minimize_strong_formulation_residual.hpp:119-134)
    s->max_iterations = 2000;
    s->relative_tolerance = 1.0e-12;
    s->precision = SIL_PRECISION_FP64;
    s->preconditioner = SIL_PRECOND_JACOBI;
    s->check_interval = 0;
    s->replace_interval = 0;
}

int sil_minimize_strong_formulation_residual(sil_operator op, const double* rhs, double* solution,
                                             const sil_solver_settings* settings,
                                             sil_solver_diagnostics* diagnostics, void* stream)
{
    SIL_REQUIRE(op && rhs && solution && settings && diagnostics, SIL_ERR_INVALID,
                "minimize_strong_formulation_residual: null argument");
    SIL_REQUIRE(settings->relative_tolerance >= 0.0, SIL_ERR_INVALID, "negative relative tolerance");
    SIL_REQUIRE(settings->precision == SIL_PRECISION_FP64 || settings->precision == SIL_PRECISION_MIXED,
                SIL_ERR_INVALID, "unknown precision mode %d", settings->precision);
    SIL_REQUIRE(settings->preconditioner == SIL_PRECOND_IDENTITY || settings->preconditioner == SIL_PRECOND_JACOBI ||
                    settings->preconditioner == SIL_PRECOND_SPD_ISAI,
                SIL_ERR_INVALID, "unknown preconditioner %d", settings->preconditioner);
    *diagnostics = sil_solver_diagnostics{};
    if (op->kind == OP_MS3D_NONLINEAR || op->kind == OP_MS2D_NONLINEAR) {
        // SpdIsai: generated from the Jacobian assembled at every Newton step
        return newton_solve(op, rhs, solution, *settings, *diagnostics, as_stream(stream));
    }
    return cg_solve(op, rhs, solution, *settings, *diagnostics, as_stream(stream));
}

int sil_minimize_strong_formulation_residual_noalloc(sil_operator op, const double* rhs, double* solution,
                                                     const sil_solver_settings* settings,
                                                     sil_solver_diagnostics* diagnostics, void* stream)
{
    SIL_REQUIRE(op && rhs && solution && settings && diagnostics, SIL_ERR_INVALID,
                "minimize_strong_formulation_residual_noalloc: null argument");
    SIL_REQUIRE(static_cast<const void*>(rhs) != static_cast<const void*>(solution), SIL_ERR_INVALID,
                "minimize_strong_formulation_residual_noalloc: the solution (the iterate) must not alias rhs");
    SIL_REQUIRE(settings->relative_tolerance >= 0.0, SIL_ERR_INVALID, "negative relative tolerance");
    *diagnostics = sil_solver_diagnostics{};
    return cg_solve_noalloc(op, rhs, solution, *settings, *diagnostics, as_stream(stream));
}

int64_t sil_noalloc_max_unknowns(void) { return noalloc_max_dofs(); }

}  // extern "C"
