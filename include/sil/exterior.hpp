// sil/exterior.hpp — C++20 host interface of the MI355X DEC library.
//
// Keeps the call surface of the reference's sil::exterior
// (src/similie/exterior/coboundary.hpp:614-709, hodge_star.hpp:453-494,
// codifferential.hpp:510-656, laplacian.hpp:327-358, 814-951) over
// non-owning strided views instead of DDC chunk spans, and forwards to the C
// ABI in sil_hip.h on a HIP stream (the reference's Kokkos execution-space
// instance).  Errors throw sil::Error (the C ABI itself never throws).
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../sil_hip.h"

namespace sil {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const noexcept { return status_; }

private:
    int status_;
};

inline void check(int status)
{
    if (status != SIL_OK) throw Error(status, sil_last_error());
}

// Non-owning view of a k-form sampled on an N-dimensional node grid, in the
// reference's Tensor layout conventions.
template <int N>
struct FormView {
    double* data = nullptr;
    std::array<int64_t, N> extent{};
    std::array<int64_t, N> stride{};
    int64_t comp_stride = 1;
    int64_t comp_extent = 1; // components along the component axis: C(N, k) for a k-form
    std::vector<std::pair<int64_t, int64_t>> spectators; // (extent, stride)

    sil_view view() const
    {
        sil_view v{};
        v.ndim = N;
        for (int a = 0; a < N; ++a) {
            v.extent[a] = extent[a];
            v.stride[a] = stride[a];
        }
        v.comp_stride = comp_stride;
        v.comp_extent = comp_extent;
        v.nbatch = static_cast<int32_t>(spectators.size());
        for (size_t b = 0; b < spectators.size(); ++b) {
            v.batch_extent[b] = spectators[b].first;
            v.batch_stride[b] = spectators[b].second;
        }
        return v;
    }
};

// layout_right with the component index last: [x][y]...[c]
template <int N>
FormView<N> index_last(double* data, std::array<int64_t, N> extent, int64_t ncomp)
{
    FormView<N> f;
    f.data = data;
    f.extent = extent;
    int64_t s = ncomp;
    for (int a = N - 1; a >= 0; --a) {
        f.stride[a] = s;
        s *= extent[a];
    }
    f.comp_stride = 1;
    f.comp_extent = ncomp;
    return f;
}

// component index first: [c][x][y]... (tests/exterior/derivative.cpp:19-58)
template <int N>
FormView<N> index_first(double* data, std::array<int64_t, N> extent, int64_t ncomp = 1)
{
    FormView<N> f;
    f.data = data;
    f.extent = extent;
    int64_t s = 1;
    for (int a = N - 1; a >= 0; --a) {
        f.stride[a] = s;
        s *= extent[a];
    }
    f.comp_stride = s;
    f.comp_extent = ncomp;
    return f;
}

namespace exterior {

// Node-grid geometry (the reference's position and metric tensors).
template <int N>
class Grid {
public:
    // rectilinear: one host coordinate array per axis; metric: device pointer
    // laid out per `metric_kind` (nullptr for the identity)
    Grid(std::array<std::vector<double>, N> const& axes, int32_t metric_kind = SIL_METRIC_IDENTITY,
         const double* metric = nullptr)
    {
        std::array<int64_t, N> ext{};
        std::array<const double*, N> ptrs{};
        for (int a = 0; a < N; ++a) {
            ext[a] = static_cast<int64_t>(axes[a].size());
            ptrs[a] = axes[a].data();
        }
        check(sil_grid_create(&handle_, N, ext.data(), ptrs.data(), nullptr, metric_kind, metric));
        extent_ = ext;
    }
    // general positions: device [node][N]
    Grid(std::array<int64_t, N> extent, const double* nodal_positions, int32_t metric_kind = SIL_METRIC_IDENTITY,
         const double* metric = nullptr)
    {
        check(sil_grid_create(&handle_, N, extent.data(), nullptr, nodal_positions, metric_kind, metric));
        extent_ = extent;
    }
    Grid(const Grid&) = delete;
    Grid& operator=(const Grid&) = delete;
    ~Grid() { sil_grid_destroy(handle_); }
    sil_grid handle() const { return handle_; }
    std::array<int64_t, N> extent() const { return extent_; }

private:
    sil_grid handle_ = nullptr;
    std::array<int64_t, N> extent_{};
};

// deriv / coboundary<TagToAdd, CochainTag>(exec, out, in)
template <int N>
FormView<N> coboundary(void* stream, FormView<N> out, FormView<N> in, int degree)
{
    sil_view vi = in.view(), vo = out.view();
    check(sil_coboundary(degree, in.data, &vi, out.data, &vo, stream));
    return out;
}

template <int N>
FormView<N> deriv(void* stream, FormView<N> out, FormView<N> in, int degree)
{
    return coboundary<N>(stream, out, in, degree);
}

template <int N>
FormView<N> transposed_coboundary(void* stream, FormView<N> out, FormView<N> in, int degree)
{
    sil_view vi = in.view(), vo = out.view();
    check(sil_transposed_coboundary(degree, in.data, &vi, out.data, &vo, stream));
    return out;
}

// fill_discrete_hodge_star<I(k), J(N-k)>(exec, hodge, metric, position):
// [node][N^k][C(N, N-k)]
template <int N>
double* fill_discrete_hodge_star(void* stream, Grid<N> const& grid, int degree, double* hodge)
{
    check(sil_fill_discrete_hodge_star(grid.handle(), degree, hodge, stream));
    return hodge;
}

template <int N>
FormView<N> hodge_star(void* stream, Grid<N> const& grid, FormView<N> out, FormView<N> in, int degree)
{
    sil_view vi = in.view(), vo = out.view();
    check(sil_hodge_star(grid.handle(), degree, in.data, &vi, out.data, &vo, stream));
    return out;
}

template <int N>
FormView<N> codifferential(void* stream, Grid<N> const& grid, FormView<N> out, FormView<N> in, int degree)
{
    sil_view vi = in.view(), vo = out.view();
    check(sil_codifferential(grid.handle(), degree, in.data, &vi, out.data, &vo, stream));
    return out;
}

template <int N>
FormView<N> laplacian(void* stream, Grid<N> const& grid, FormView<N> out, FormView<N> in, int degree)
{
    sil_view vi = in.view(), vo = out.view();
    check(sil_laplacian(grid.handle(), degree, in.data, &vi, out.data, &vo, stream));
    return out;
}

// make_staged_laplacian(exec, out, in, metric, position).run(out, in)
template <int N>
class StagedLaplacian {
public:
    StagedLaplacian(Grid<N> const& grid, FormView<N> out, FormView<N> in, int degree)
        : in_(in), out_(out)
    {
        sil_view vi = in.view(), vo = out.view();
        check(sil_make_staged_laplacian(&handle_, grid.handle(), degree, &vi, &vo));
    }
    StagedLaplacian(const StagedLaplacian&) = delete;
    StagedLaplacian& operator=(const StagedLaplacian&) = delete;
    ~StagedLaplacian() { sil_staged_laplacian_destroy(handle_); }
    // the handle keeps the layouts it was built for: views of another shape
    // or other strides are refused
    FormView<N> run(void* stream, FormView<N> out, FormView<N> in)
    {
        if (!same_layout(in, in_) || !same_layout(out, out_))
            throw Error(SIL_ERR_INVALID, "staged laplacian called with layouts it was not built for");
        check(sil_staged_laplacian_run(handle_, out.data, in.data, stream));
        return out;
    }

private:
    static bool same_layout(FormView<N> const& a, FormView<N> const& b)
    {
        return a.extent == b.extent && a.stride == b.stride && a.comp_stride == b.comp_stride &&
               a.comp_extent == b.comp_extent && a.spectators == b.spectators;
    }

    FormView<N> in_, out_;
    sil_staged_laplacian handle_ = nullptr;
};

template <int N>
StagedLaplacian<N> make_staged_laplacian(Grid<N> const& grid, FormView<N> out, FormView<N> in, int degree)
{
    return StagedLaplacian<N>(grid, out, in, degree);
}

// StagedCodifferential(exec, tensor, metric, position).run(out, tensor)
// (codifferential.hpp:321-508)
template <int N>
class StagedCodifferential {
public:
    StagedCodifferential(Grid<N> const& grid, FormView<N> out, FormView<N> in, int degree)
        : in_(in), out_(out)
    {
        sil_view vi = in.view(), vo = out.view();
        check(sil_make_staged_codifferential(&handle_, grid.handle(), degree, &vi, &vo));
    }
    StagedCodifferential(const StagedCodifferential&) = delete;
    StagedCodifferential& operator=(const StagedCodifferential&) = delete;
    ~StagedCodifferential() { sil_staged_codifferential_destroy(handle_); }
    FormView<N> run(void* stream, FormView<N> out, FormView<N> in)
    {
        if (!same_layout(in, in_) || !same_layout(out, out_))
            throw Error(SIL_ERR_INVALID, "staged codifferential called with layouts it was not built for");
        check(sil_staged_codifferential_run(handle_, out.data, in.data, stream));
        return out;
    }

private:
    static bool same_layout(FormView<N> const& a, FormView<N> const& b)
    {
        return a.extent == b.extent && a.stride == b.stride && a.comp_stride == b.comp_stride &&
               a.comp_extent == b.comp_extent && a.spectators == b.spectators;
    }

    FormView<N> in_, out_;
    sil_staged_codifferential handle_ = nullptr;
};

template <int N>
StagedCodifferential<N> make_staged_codifferential(Grid<N> const& grid, FormView<N> out, FormView<N> in,
                                                   int degree)
{
    return StagedCodifferential<N>(grid, out, in, degree);
}

}  // namespace exterior

namespace solvers {

// StrongFormulationSolverSettings / Diagnostics
// (src/similie/solvers/minimize_strong_formulation_residual.hpp:119-144)
struct StrongFormulationSolverSettings : sil_solver_settings {
    StrongFormulationSolverSettings() { sil_default_solver_settings(this); }
};
using StrongFormulationSolverDiagnostics = sil_solver_diagnostics;

// minimize_strong_formulation_residual(exec, operator_model, rhs, solution, settings)
inline StrongFormulationSolverDiagnostics minimize_strong_formulation_residual(
        void* stream, sil_operator op, const double* rhs, double* solution,
        StrongFormulationSolverSettings const& settings = {})
{
    StrongFormulationSolverDiagnostics d{};
    check(sil_minimize_strong_formulation_residual(op, rhs, solution, &settings, &d, stream));
    return d;
}

// build_matrix(gko_exec, operator_model) (minimize_strong_formulation_residual.hpp:623-633) over
// assemble_matrix_data (onelab_interface/magnetostatics_onelab.hpp:3115-3536): the operator
// assembled into a device CSR matrix, as a new operator the caller destroys with
// sil_operator_destroy; solving on it is the use_matrix_free = false path
inline sil_operator build_matrix(void* stream, sil_operator operator_model)
{
    sil_operator m = nullptr;
    check(sil_magnetostatics3d_assemble(operator_model, &m, stream));
    return m;
}

}  // namespace solvers

}  // namespace sil
