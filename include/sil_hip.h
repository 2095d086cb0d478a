/*
 * sil_hip.h — C ABI of the MI355X discrete-exterior-calculus library
 * (libsil_hip.so).
 *
 * Drop-in boundary for SimiLie's sil::exterior operators and its
 * strong-formulation CG solve.  Every entry point is `extern "C"`, takes
 * plain pointers and sizes, never throws, returns SIL_OK (0) or a negative
 * status with a message available from sil_last_error(), and runs
 * asynchronously on the caller's HIP stream (passed as `void*`; NULL is the
 * default stream) — the analogue of the Kokkos execution-space instance the
 * reference launches on.  Array pointers are device pointers owned by the
 * caller unless stated otherwise.
 *
 * Array layout follows the reference's Tensor over
 * DiscreteDomain<DDim..., Index> (src/similie/tensor/tensor_impl.hpp
 * :587-602): a sil_view gives per-axis strides and a component stride, so
 * both index-last (layout_right, component fastest) and index-first
 * (tests/exterior/derivative.cpp:19-58) storage are accepted, plus spectator
 * (batch) dimensions for the coboundary (derivative.cpp:889-990).  Component
 * order of a k-form is the lexicographic order of the k-subsets of the axes
 * (antisymmetric_tensor.hpp:72-98).
 */
#ifndef SIL_HIP_H
#define SIL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SIL_MAX_DIM 4

enum sil_status {
    SIL_OK = 0,
    SIL_ERR_INVALID = -1,     /* bad argument (shape, rank, null pointer) */
    SIL_ERR_UNSUPPORTED = -2, /* valid request outside what the library implements */
    SIL_ERR_HIP = -3,         /* a HIP runtime call failed */
    SIL_ERR_NOT_CONVERGED = -4
};

enum sil_metric_kind {
    SIL_METRIC_IDENTITY = 0,  /* TensorIdentityIndex: no storage */
    SIL_METRIC_DIAGONAL = 1,  /* TensorDiagonalIndex: [node][ndim] */
    SIL_METRIC_SYMMETRIC = 2  /* TensorSymmetricIndex: [node][ndim(ndim+1)/2], XX XY XZ YY YZ ZZ */
};

enum sil_precision {
    SIL_PRECISION_FP64 = 0,  /* FP64 throughout */
    SIL_PRECISION_MIXED = 1  /* FP32 operator applications, FP64 residual replacement, dots and scalars */
};

enum sil_preconditioner {
    SIL_PRECOND_IDENTITY = 0,
    SIL_PRECOND_JACOBI = 1,
    SIL_PRECOND_LEARNED = 2, /* the U-Net of a sil_unet handle, inside flexible CG (equal-size batches only) */
    SIL_PRECOND_SPD_ISAI = 3 /* PreconditionerType::SpdIsai: M = Gᵀ G, G on the pattern of tril(A) (FP64 only) */
};

/* Strided view of a k-form sampled on a node grid. */
typedef struct sil_view {
    int32_t ndim;   /* grid (form) axes, 1..4 */
    int32_t nbatch; /* spectator dimensions, 0..4 */
    int64_t extent[SIL_MAX_DIM];
    int64_t stride[SIL_MAX_DIM]; /* element strides of the grid axes */
    int64_t comp_stride;         /* element stride between components */
    int64_t batch_extent[SIL_MAX_DIM];
    int64_t batch_stride[SIL_MAX_DIM];
    int64_t comp_extent; /* components the array holds along its component axis (1 when it has none);
                            every entry point requires C(N, k) for a k-form */
} sil_view;

const char* sil_last_error(void);
int32_t sil_abi_version(void);

/* ------------------------------------------------------------------ */
/* geometry                                                            */
/* ------------------------------------------------------------------ */

typedef struct sil_grid_s* sil_grid;

/* Node grid geometry: the reference's position and metric tensors
 * (fill_discrete_hodge_star's `metric`, `position` arguments,
 * src/similie/exterior/hodge_star.hpp:461-465).
 *  axis_coords: host arrays, axis_coords[a][i] = coordinate of node i on
 *               axis a (rectilinear grids); or NULL when nodal_positions
 *               is given.
 *  nodal_positions: device [node][ndim] positions (general grids) or NULL.
 *  metric: device pointer per metric_kind (NULL for the identity); the
 *          grid keeps the pointer, the caller keeps the memory alive. */
int sil_grid_create(sil_grid* out, int32_t ndim, const int64_t* extent, const double* const* axis_coords,
                    const double* nodal_positions, int32_t metric_kind, const double* metric);
int sil_grid_destroy(sil_grid grid);

/* ------------------------------------------------------------------ */
/* sil::exterior operators                                             */
/* ------------------------------------------------------------------ */

/* deriv / coboundary (src/similie/exterior/coboundary.hpp:614-654,
 * 698-709): k-form `in` → (k+1)-form `out`, forward differences with samples
 * clamped into the input domain (evaluators.hpp:30-46).  out extents may be
 * any size (the reference's output tensor domain). */
int sil_coboundary(int32_t k, const double* in, const sil_view* in_view, double* out, const sil_view* out_view,
                   void* stream);

/* transposed_coboundary (coboundary.hpp:656-696): k-form → (k+1)-form,
 * backward differences with zero outside the input domain, negated. */
int sil_transposed_coboundary(int32_t k, const double* in, const sil_view* in_view, double* out,
                              const sil_view* out_view, void* stream);

/* fill_discrete_hodge_star<I(k indices), J(N-k indices)>
 * (hodge_star.hpp:453-494): per-node coefficient tensor in the reference
 * layout [node][N^k source tuple][C(N, N-k) target], index-last. */
int sil_fill_discrete_hodge_star(sil_grid grid, int32_t k, double* hodge, void* stream);

/* Discrete ⋆ applied to a k-form: the tensor_prod of the filled star with
 * the form (codifferential.hpp:584) — k-form → dual (N-k)-form, on the grid's
 * node domain. */
int sil_hodge_star(sil_grid grid, int32_t k, const double* in, const sil_view* in_view, double* out,
                   const sil_view* out_view, void* stream);

/* codifferential (codifferential.hpp:551-656): k-form on the grid's node
 * domain → (k-1)-form on out_view's extents (≤ grid extents). */
int sil_codifferential(sil_grid grid, int32_t k, const double* in, const sil_view* in_view, double* out,
                       const sil_view* out_view, void* stream);

/* StagedCodifferential (codifferential.hpp:321-508): the request validated
 * once, then run(out, in) any number of times on arrays of the same views.
 * The Hodge coefficients are evaluated inside the kernel from the grid's
 * geometry (the reference stores the two star tensors instead). */
typedef struct sil_staged_codifferential_s* sil_staged_codifferential;
int sil_make_staged_codifferential(sil_staged_codifferential* out, sil_grid grid, int32_t k, const sil_view* in_view,
                                   const sil_view* out_view);
int sil_staged_codifferential_run(sil_staged_codifferential op, double* out, const double* in, void* stream);
int sil_staged_codifferential_destroy(sil_staged_codifferential op);

/* laplacian (laplacian.hpp:814-951): Hodge Laplacian δd + dδ of a k-form on
 * the grid's node domain, written on out_view's extents (the reference's
 * callers use the node domain minus the last node of every axis).  Fused
 * single-pass kernels serve 2D 0-forms and 3D 1-forms on rectilinear grids
 * with identity or diagonal metrics; every other case runs the general
 * per-node kernel. */
int sil_laplacian(sil_grid grid, int32_t k, const double* in, const sil_view* in_view, double* out,
                  const sil_view* out_view, void* stream);

/* make_staged_laplacian(...).run(out, in) (laplacian.hpp:312-358, 599-647):
 * validates the request once and fixes the kernel and tile configuration. */
typedef struct sil_staged_laplacian_s* sil_staged_laplacian;
int sil_make_staged_laplacian(sil_staged_laplacian* out, sil_grid grid, int32_t k, const sil_view* in_view,
                              const sil_view* out_view);
int sil_staged_laplacian_run(sil_staged_laplacian op, double* out, const double* in, void* stream);
int sil_staged_laplacian_destroy(sil_staged_laplacian op);

/* The autotuner of the fused 3D 1-form Laplacian (no reference
 * counterpart: the reference has one kernel shape): times each candidate
 * kernel shape `repeats` times on the staged operator's own arrays (one
 * warm-up run each, HIP events on `stream`), keeps the fastest for later
 * runs, leaves its result in `out` and reports it.  variant 2 = pipelined
 * x-march (lanes_or_depth: z-lanes per row), 3 = staged x-march
 * (lanes_or_depth: planes of loads in flight).  Other paths, or a tile
 * fixed by SIL_LAP3D_TILE, are reported unchanged (microseconds 0).
 * Synchronises `stream`. */
typedef struct sil_lap3d_tile {
    int32_t variant;
    int32_t rows;
    int32_t xchunk;
    int32_t lanes_or_depth;
    int32_t nontemporal;
    double microseconds;
} sil_lap3d_tile;
int sil_staged_laplacian_autotune(sil_staged_laplacian op, double* out, const double* in, int32_t repeats,
                                  sil_lap3d_tile* chosen, void* stream);

/* ------------------------------------------------------------------ */
/* linear operators and the CG solve                                   */
/* ------------------------------------------------------------------ */

typedef struct sil_operator_s* sil_operator;

/* MagnetostaticsOperator3D (onelab_interface/magnetostatics_onelab.hpp
 * :1524-2789) with precomputed-stencil semantics (:2241-2396): DOF order
 * 3·(i + nx·(j + ny·k)) + c, identity rows on boundary nodes.  Coordinates
 * are host arrays (copied); mu_nodal is a device array of nx·ny·nz nodal
 * permeabilities (copied). */
int sil_magnetostatics3d_create(sil_operator* out, int64_t nx, int64_t ny, int64_t nz, const double* x_coords,
                                const double* y_coords, const double* z_coords, const double* mu_nodal,
                                double gauge_penalty);

/* MagnetostaticsOperator3D with MaterialMagnetostaticsHamiltonian
 * (magnetostatics_onelab.hpp:957-1011, 5108-5121): nodes whose
 * ferromagnetic fraction (ferromagnetic_nodal, device, the mean over the
 * adjacent cells of the nonlinear-material flag, :5026-5063) exceeds 0.5
 * follow NonlinearMagnetostaticsHamiltonian with the InterpolatedNonlinear-
 * BHCurve of (bh_b, bh_h) (host arrays, 2..64 samples, bh_b increasing;
 * src/similie/physics/magnetostatics/nonlinear_magnetostatics.py:14-227), the
 * others are linear with mu_nodal.  sil_operator_apply gives F(x) and keeps
 * x's material state; sil_minimize_strong_formulation_residual runs the
 * reference's Newton branch (minimize_strong_formulation_residual.hpp
 * :1383-1569; identity or Jacobi; with SIL_PRECISION_MIXED the inner CG
 * applies the Jacobian in FP32 and converges on its FP64 residual, F(x)
 * stays FP64).  Needs >= 3 nodes per axis. */
int sil_magnetostatics3d_nonlinear_create(sil_operator* out, int64_t nx, int64_t ny, int64_t nz,
                                          const double* x_coords, const double* y_coords, const double* z_coords,
                                          const double* mu_nodal, const double* ferromagnetic_nodal,
                                          const double* bh_b, const double* bh_h, int32_t nsamples,
                                          double gauge_penalty);

/* The 3D operator's switches (magnetostatics_onelab.hpp:796-828, read there
 * from the environment at every application, and the 3D right-hand side's
 * criterion, :4930-4935):
 *   y_response_sign  −1 flips the sign of the y moment's response (<0 → −1,
 *                    anything else → +1);
 *   divergence_gauge nonzero: the penalty term P·w^A·A is replaced by the
 *                    divergence gauge factor·δᵀ W δ A (:1893-1941,
 *                    2370-2392: δ the codifferential of the interior
 *                    potential at an identity metric, W = ℓx ℓy ℓz per node,
 *                    factor = P · gauge_sign);
 *   gauge_sign       sign of that factor (<0 → −1, else +1);
 *   criterion        0 PotentialTemporalDerivative (the right-hand side is
 *                    zero), 1 MomentsTemporalDerivative (the default),
 *                    2 both. */
typedef struct sil_ms3d_switches {
    double y_response_sign;
    int32_t divergence_gauge;
    double gauge_sign;
    int32_t criterion;
} sil_ms3d_switches;

/* the defaults, then SIMILIE_MAGNETIC_Y_RESPONSE_SIGN_3D,
 * SIMILIE_DIVERGENCE_GAUGE_3D and SIMILIE_VECTOR_POTENTIAL_GAUGE_SIGN_3D
 * parsed as the reference parses them (magnetostatics_onelab.hpp:796-828:
 * a sign is +1 unless the variable holds a number < 0; the gauge is on when
 * the variable is set, non-empty and does not start with '0') */
int sil_ms3d_switches_from_environment(sil_ms3d_switches* out);

/* sets the switches of a linear or nonlinear 3D magnetostatics operator
 * (every later application, diagonal, assembly, right-hand side and solve);
 * drops the operator's cached solver state */
int sil_magnetostatics3d_set_switches(sil_operator op, const sil_ms3d_switches* switches);

/* Air-gap diagnostics of a solution (magnetostatics_onelab.hpp:4666-4720 2D,
 * 5246-5307 3D, 460-519): from the post-processing's cell fields (cell_b,
 * cell_h: device [cells][3], cells x fastest) and a device mask per cell
 * (bit 0: in the diagnostic region, bit 1: a positively oriented
 * conductor).  Upper air-gap faces are the +y faces of region cells whose +y
 * neighbour is outside the region (measure dx·dz, 2D dx·length_z); flux =
 * Σ measure·B_y, traction T·n = (B_x H_y, B_y H_y − ½B·H, B_z H_y) summed
 * with its magnitude, replaced by the x-mirror projection (y, z parts of
 * pairs at ±x) when the faces are mirror-symmetric; current = Σ dx·dy·J_z
 * over conductor cells (3D: the layer k = ncz/2).  Sums in the reference's
 * cell order.  Synchronises `stream`. */
typedef struct sil_air_gap_result {
    int64_t faces;
    int64_t current_cells;
    double surface_measure;
    double flux;
    double traction[3];
    double traction_magnitude;
    double current;
} sil_air_gap_result;

int sil_magnetostatics3d_air_gap_diagnostics(sil_operator op, const double* cell_b, const double* cell_h,
                                             const uint8_t* cell_mask, const double* cell_current,
                                             sil_air_gap_result* out, void* stream);
/* 2D: cell_current_z is [cells]; z_value is the z of every face (the
 * reference's grid z) */
int sil_magnetostatics2d_air_gap_diagnostics(sil_operator op, const double* cell_b, const double* cell_h,
                                             const uint8_t* cell_mask, const double* cell_current_z,
                                             double length_z, double z_value, sil_air_gap_result* out,
                                             void* stream);

/* apply_jacobian (magnetostatics_onelab.hpp:2918-3113): out = J(state)·in,
 * the derivative of the nonlinear operator at the potential `state`. */
int sil_magnetostatics3d_apply_jacobian(sil_operator op, const double* state, const double* in, double* out,
                                        void* stream);

/* the diagonal of J(state) (the Jacobi preconditioner of each Newton step,
 * which the reference takes from the assembled Jacobian, :1434-1438) */
int sil_magnetostatics3d_jacobian_diagonal(sil_operator op, const double* state, double* diag, void* stream);

/* Nodal permeability = mean over adjacent cells (magnetostatics_onelab.hpp
 * :5029-5066); cell_mu is [(nx-1)(ny-1)(nz-1)] with cells ordered
 * ci + (nx-1)(cj + (ny-1)ck). */
int sil_magnetostatics3d_nodal_mu(int64_t nx, int64_t ny, int64_t nz, const double* cell_mu, double mu0,
                                  double* mu_nodal, void* stream);

/* Right-hand side ⅛·|dual face|·J summed over the cells around each interior
 * node (magnetostatics_onelab.hpp:4919-4975); cell_current is [cell][3]. */
int sil_magnetostatics3d_rhs(sil_operator op, const double* cell_current, double* rhs, void* stream);

/* Cell post-processing of a potential on the operator's grid
 * (fill_post_process_fields_on_cell_domain_3d, make_cell_post_process_fields
 * and fill_force_density_on_hexahedral_grid_xy_slices,
 * magnetostatics_onelab.hpp:3984-4046, 4048-4168, 4215-4254).  Per cell
 * (ci + (nx-1)(cj + (ny-1)ck)):
 *   induction [cell][3]  B_m = moment_m / face area (the potential's
 *                        forward_value stencil, all nodes read);
 *   field [cell][3]      H = h(|B|)/|B|·B where cell_nonlinear != 0 (the
 *                        operator's B-H curve, H = 0 at B = 0), B/μ_cell
 *                        elsewhere;
 *   maxwell_stress [cell][6]  B_xH_x − ½B·H, B_yH_y − ½B·H, B_zH_z − ½B·H,
 *                        B_xH_y, B_xH_z, B_yH_z;
 *   force_density [cell][3]   per xy slice, −δ of (T_xx, T_xy), (T_xy, T_yy),
 *                        (T_xz, T_yz) on the cell-centre grid, identity
 *                        metric.
 * op: a linear or nonlinear magnetostatics operator; cell_nonlinear (device
 * [cell], or NULL for all-linear) needs a nonlinear operator.  Any output
 * may be NULL. */
int sil_magnetostatics3d_post_process(sil_operator op, const double* potential, const double* cell_mu,
                                      const double* cell_nonlinear, double* induction, double* field,
                                      double* maxwell_stress, double* force_density, void* stream);

/* MagnetostaticsOperator2D (onelab_interface/magnetostatics_onelab.hpp
 * :1013-1522): A_z on an nx×ny rectilinear node grid, one unknown per node
 * (row i + nx·j).  Interior rows gather ∂H/∂B at the interior sampled nodes
 * s of the moments B_x = A(s+e_y) − A(s), B_y = A(s) − A(s+e_x) (no Hodge
 * weight; boundary values of A are read); boundary rows are the identity.
 * criterion: 0 potential, 1 moments (the reference's default), 2 both
 * temporal derivatives (the row sum counted twice).  mu_nodal: device [node]. */
int sil_magnetostatics2d_create(sil_operator* out, int64_t nx, int64_t ny, const double* x_coords,
                                const double* y_coords, const double* mu_nodal, int32_t criterion);

/* The same operator with MaterialMagnetostaticsHamiltonian (:957-1011): nodes
 * whose ferromagnetic fraction exceeds 0.5 follow the interpolated B-H curve
 * (2..64 increasing samples); Newton in
 * sil_minimize_strong_formulation_residual, as for the 3D operator. */
int sil_magnetostatics2d_nonlinear_create(sil_operator* out, int64_t nx, int64_t ny, const double* x_coords,
                                          const double* y_coords, const double* mu_nodal,
                                          const double* ferromagnetic_nodal, const double* bh_b, const double* bh_h,
                                          int32_t nsamples, int32_t criterion);

/* J(state)·in and the diagonal of J(state) of a 2D nonlinear operator
 * (apply_jacobian for the 2D operator, :2790-2920) */
int sil_magnetostatics2d_apply_jacobian(sil_operator op, const double* state, const double* in, double* out,
                                        void* stream);
int sil_magnetostatics2d_jacobian_diagonal(sil_operator op, const double* state, double* diag, void* stream);

/* mean μ of the cells around each node (:4467-4500); cell c = ci + (nx-1)·cj */
int sil_magnetostatics2d_nodal_mu(int64_t nx, int64_t ny, const double* cell_mu, double mu0, double* mu_nodal,
                                  void* stream);

/* ¼·area·J_z summed over the cells around each interior node (:4413-4445);
 * zero everywhere for the potential criterion */
int sil_magnetostatics2d_rhs(sil_operator op, const double* cell_current_z, double* rhs, void* stream);

/* Cell post-processing of a 2D solution (fill_post_process_fields_on_cell_domain
 * and fill_force_density_on_quadrilateral_grid, :3866-3982, 4048-4213): B =
 * (ΔyA/h_y, −ΔxA/h_x, 0) reconstructed per cell, H, the Maxwell stress and
 * −δ of the stress one-forms on the cell-centre grid; layouts as for
 * sil_magnetostatics3d_post_process with cell = ci + (nx-1)·cj. */
int sil_magnetostatics2d_post_process(sil_operator op, const double* potential, const double* cell_mu,
                                      const double* cell_nonlinear, double* induction, double* field,
                                      double* maxwell_stress, double* force_density, void* stream);

/* Dirichlet-zero 2D Poisson operator d0ᵀ ⋆1^κ d0 on the interior nodes of a
 * rectilinear nx×ny grid (x slowest): the ⋆0-weighted DEC 0-form Laplacian of
 * the reference, with optional nodal coefficient κ (edge value = mean of its
 * two nodes).  Vectors span all nx·ny nodes; boundary entries stay zero. */
int sil_poisson2d_create(sil_operator* out, int64_t nx, int64_t ny, const double* x_coords,
                         const double* y_coords, const double* kappa);

/* assemble_matrix_data + csr_from_matrix_data
 * (onelab_interface/magnetostatics_onelab.hpp:3115-3536,
 * src/similie/solvers/minimize_strong_formulation_residual.hpp:565-646):
 * the magnetostatics operator assembled on the device into a new operator
 * holding a CSR matrix (gko::matrix::Csr<double, int32>: int32 row pointers
 * and columns, columns ascending within a row, FP64 values) whose apply is
 * the sparse product.  Solving on it is the reference's
 * use_matrix_free = false path.  Synchronises `stream`. */
int sil_magnetostatics3d_assemble(sil_operator op, sil_operator* csr_out, void* stream);

/* assemble_matrix_data(model, state) (onelab_interface/magnetostatics_onelab
 * .hpp:3115-3536): the Jacobian of a nonlinear magnetostatics operator at
 * `state` (a potential), assembled on the device as above — the matrix the
 * reference builds at every Newton step to generate its preconditioner
 * (minimize_strong_formulation_residual.hpp:1432-1433).  Rows reach at most
 * 23 columns.  Synchronises `stream`. */
int sil_magnetostatics3d_assemble_jacobian(sil_operator op, const double* state, sil_operator* csr_out, void* stream);

/* assemble_matrix_data(model) / assemble_matrix_data(model, state) of the 2D
 * operator (onelab_interface/magnetostatics_onelab.hpp:3554-3748) + the
 * CSR conversion (minimize_strong_formulation_residual.hpp:565-646): the
 * linear 2D magnetostatics operator, or the Jacobian of a nonlinear one at
 * `state`, assembled on the device into a new operator holding a CSR matrix
 * as above (boundary rows: 1 on the diagonal; interior rows at most 7
 * columns).  Synchronises `stream`. */
int sil_magnetostatics2d_assemble(sil_operator op, sil_operator* csr_out, void* stream);
int sil_magnetostatics2d_assemble_jacobian(sil_operator op, const double* state, sil_operator* csr_out, void* stream);

/* gko::matrix::Csr<double, int32>::create(exec, size, values, columns,
 * row_ptrs) (minimize_strong_formulation_residual.hpp:613-620): an operator
 * over a caller's square CSR matrix (host or device arrays, copied).  Rows
 * must run from 0 to nnz with strictly ascending columns in [0, n).
 * Synchronises `stream`. */
int sil_csr_create(sil_operator* out, int64_t n, int64_t nnz, const int32_t* row_ptr, const int32_t* cols,
                   const double* vals, void* stream);

/* stored entries of an assembled operator; -1 for a matrix-free one */
int64_t sil_csr_nnz(sil_operator op);

/* copies the CSR arrays of an assembled operator into caller buffers (host
 * or device) of n+1, nnz and nnz entries.  Synchronises `stream`. */
int sil_csr_copy(sil_operator op, int32_t* row_ptr, int32_t* cols, double* vals, void* stream);

/* gko::preconditioner::SpdIsai::generate (minimize_strong_formulation_residual
 * .hpp:963-967): the factor G of M = Gᵀ G ≈ A⁻¹ as a new assembled operator
 * (pattern of tril(A); row i solves A[J,J] y = e with J its stored columns ≤ i
 * and is scaled by 1/√y_i).  `a` is an assembled operator or the
 * linear 3D or 2D magnetostatics operator (assembled for the purpose, as
 * the reference does).  Rows may hold at most 32 entries in their lower triangle.
 * Synchronises `stream`. */
int sil_spd_isai_create(sil_operator a, sil_operator* g_out, void* stream);

int64_t sil_operator_size(sil_operator op);
int sil_operator_apply(sil_operator op, const double* in, double* out, void* stream);
int sil_operator_apply_f32(sil_operator op, const float* in, float* out, void* stream);
int sil_operator_diagonal(sil_operator op, double* diag, void* stream);
int sil_operator_destroy(sil_operator op);

/* StrongFormulationSolverSettings / Diagnostics
 * (src/similie/solvers/minimize_strong_formulation_residual.hpp:119-144) */
typedef struct sil_solver_settings {
    int32_t max_iterations;
    double relative_tolerance;
    int32_t precision;      /* enum sil_precision */
    int32_t preconditioner; /* enum sil_preconditioner */
    int32_t check_interval; /* iterations between host convergence checks (0 = auto) */
    int32_t replace_interval; /* mixed precision: FP64 residual replacement period (0 = auto) */
} sil_solver_settings;

typedef struct sil_solver_diagnostics {
    int32_t iterations;
    int32_t converged;
    double initial_residual_l2;
    double final_residual_l2;       /* recurrence (or replaced) residual */
    double final_relative_residual;
    double true_relative_residual;  /* FP64 ‖b − A x‖2 / ‖b‖2 recomputed at exit */
    double duration;                /* seconds, host wall clock around the solve */
    int32_t outer_iterations;       /* Newton steps of a nonlinear operator (0 for a linear one) */
    int32_t verified;               /* 1 when the FP64 residual recomputed at exit meets the tolerance
                                       (‖b − A x‖2 ≤ tol·‖b‖2; Newton: the FP64 nonlinear residual) */
} sil_solver_diagnostics;

void sil_default_solver_settings(sil_solver_settings* settings);

/* minimize_strong_formulation_residual for a linear operator
 * (minimize_strong_formulation_residual.hpp:1308-1382): x0 = 0, CG until
 * ‖r‖2 ≤ tol·‖b‖2 or max_iterations.  `op` is matrix-free
 * (use_matrix_free = true) or assembled (false); the SpdIsai preconditioner
 * is generated once per operator, from its assembled matrix.  Synchronises
 * `stream` on return. */
int sil_minimize_strong_formulation_residual(sil_operator op, const double* rhs, double* solution,
                                             const sil_solver_settings* settings,
                                             sil_solver_diagnostics* diagnostics, void* stream);

/* No counterpart in the reference.  The same FP64 CG (identity or Jacobi
 * preconditioning; Poisson 2D and linear magnetostatics 2D / 3D operators)
 * for callers that cannot allocate device memory: code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc.  It calls
 * no allocating routine and instantiates no graph: its vectors and partial
 * sums live in a file-scope device array of the library (about 29 MB on
 * every device it runs on, held until the process ends), so problems of at
 * most sil_noalloc_max_unknowns() unknowns (2^19) are accepted and larger
 * ones are refused with SIL_ERR_UNSUPPORTED.  `solution` is the iterate,
 * `rhs` is read in place.  Every call on one device shares that array:
 * calls that overlap on one device race.  Synchronises `stream` on return. */
int sil_minimize_strong_formulation_residual_noalloc(sil_operator op, const double* rhs, double* solution,
                                                     const sil_solver_settings* settings,
                                                     sil_solver_diagnostics* diagnostics, void* stream);
int64_t sil_noalloc_max_unknowns(void);

/* ------------------------------------------------------------------ */
/* input formats (host only: no device memory, no stream)              */
/* ------------------------------------------------------------------ */

/* gmsh::parse_supported_msh2_mesh + build_structured_grid
 * (onelab_interface/gmsh_structured_grid.hpp:239-553): a Gmsh MSH2 ASCII
 * mesh made of quadrilaterals (2D) or hexahedra (3D) on a full rectilinear
 * grid; points, lines and triangles are skipped.  The grid keeps the sorted
 * axis coordinates (merged within 1e-12) and the physical tag of every cell,
 * x fastest. */
typedef struct sil_structured_grid_s* sil_structured_grid;
int sil_msh2_read(const char* path, sil_structured_grid* out);
int sil_msh2_parse(const char* text, int64_t length, sil_structured_grid* out);
/* ndim = 2 or 3; nodes[3] = nodes per axis (1 for the missing axis) */
int sil_structured_grid_shape(sil_structured_grid grid, int32_t* ndim, int64_t* nodes);
int sil_structured_grid_coords(sil_structured_grid grid, int32_t axis, double* out);
/* z of the plane of a quadrilateral mesh */
double sil_structured_grid_plane(sil_structured_grid grid);
int sil_structured_grid_cell_tags(sil_structured_grid grid, int32_t* out);
int sil_structured_grid_destroy(sil_structured_grid grid);

/* .silpro problem files (onelab_interface/onelab_interface.hpp:160-560):
 * sections of `Key value;` entries, # and // comments, quoted strings. */
typedef struct sil_silpro_s* sil_silpro;
int sil_silpro_read(const char* path, sil_silpro* out);
int sil_silpro_parse(const char* text, int64_t length, sil_silpro* out);
/* value at a dotted path, e.g. "Solver.RelativeTolerance" (NUL-terminated
 * into buf of `size` bytes) */
int sil_silpro_value(sil_silpro h, const char* path, char* buf, int64_t size);
/* index-th value of a section, in key order (e.g. the tags of
 * "Preprocess.MagneticMaterialTags") */
int sil_silpro_section_value(sil_silpro h, const char* section, int32_t index, char* buf, int64_t size);
int sil_silpro_destroy(sil_silpro h);

enum sil_physics {
    SIL_PHYSICS_LINEAR_MAGNETOSTATICS = 0,
    SIL_PHYSICS_NONLINEAR_MAGNETOSTATICS = 1,
    SIL_PHYSICS_SCALAR_FIELD_WITH_POWER_COUPLING = 2
};

/* PreconditionerType (src/similie/solvers/minimize_strong_formulation_residual.hpp:54-63) */
enum sil_silpro_preconditioner {
    SIL_SILPRO_PRECOND_IDENTITY = 0,
    SIL_SILPRO_PRECOND_JACOBI = 1,
    SIL_SILPRO_PRECOND_SPD_ISAI = 2,
    SIL_SILPRO_PRECOND_SYMMETRIC_GAUSS_SEIDEL = 3,
    SIL_SILPRO_PRECOND_SSOR = 4,
    SIL_SILPRO_PRECOND_CHEBYSHEV_JACOBI = 5,
    SIL_SILPRO_PRECOND_IR_JACOBI = 6,
    SIL_SILPRO_PRECOND_GENERAL_ISAI = 7
};

/* the Problem / Solver sections of a .silpro file: MinimizeStrongFormulation-
 * ResidualProblem (onelab_interface.hpp:96-111) with its defaults */
typedef struct sil_silpro_settings {
    int32_t physics;                 /* enum sil_physics */
    int64_t max_iterations;          /* 10000 */
    double relative_tolerance;       /* 1e-10 */
    int32_t jacobi_max_block_size;   /* 1 */
    int32_t use_matrix_free;         /* 1 */
    int32_t criterion;               /* 0 potential, 1 moments (default), 2 both temporal derivatives */
    int32_t preconditioner;          /* enum sil_silpro_preconditioner, Jacobi */
    double vector_potential_gauge_penalty; /* 20 */
    double sor_relaxation_factor;    /* 1.2 */
    int32_t chebyshev_iterations;    /* 8 */
    double chebyshev_lower_bound;    /* 0.05 */
    double chebyshev_upper_bound;    /* 1.5 */
    int32_t ir_iterations;           /* 8 */
    double ir_relaxation_factor;     /* 1 */
} sil_silpro_settings;

int sil_silpro_problem_settings(sil_silpro h, sil_silpro_settings* out);
/* the solver settings of this library for a parsed problem; the
 * preconditioners not implemented here (Gauss-Seidel, SSOR, Chebyshev, IR,
 * GeneralIsai, block Jacobi) are an error */
int sil_silpro_solver_settings(const sil_silpro_settings* in, sil_solver_settings* out);

/* the cell inputs of the magnetostatics preprocess
 * (magnetostatics_onelab.hpp:4870-4897): μ0, no current and linear by
 * default; magnetic-material tags take core_mu (and the nonlinear flag when
 * enabled); otherwise ±x and then ±z conductor tags take ±current density. */
typedef struct sil_cell_materials {
    double mu0, core_mu, current_density_x, current_density_z;
    int32_t use_nonlinear_magnetic_material;
    const int32_t* magnetic_material_tags;
    int32_t n_magnetic_material_tags;
    const int32_t* positive_x_tags;
    int32_t n_positive_x_tags;
    const int32_t* negative_x_tags;
    int32_t n_negative_x_tags;
    const int32_t* positive_z_tags;
    int32_t n_positive_z_tags;
    const int32_t* negative_z_tags;
    int32_t n_negative_z_tags;
} sil_cell_materials;

/* host arrays: cell_mu[ncells], cell_current[ncells][3], nonlinear_flags[ncells] (0/1) */
int sil_magnetostatics_cell_inputs(const int32_t* cell_tags, int64_t ncells, const sil_cell_materials* materials,
                                   double* cell_mu, double* cell_current, double* nonlinear_flags);

/* ------------------------------------------------------------------ */
/* batched persistent solves                                           */
/* ------------------------------------------------------------------ */

/* No reference counterpart (the reference runs one Ginkgo solve per
 * operator): many independent Dirichlet-zero 2D Poisson problems of uneven
 * size on unit-spaced square grids, solved in one persistent launch captured
 * in a hipGraph.  Problem p has sides[p]² nodes (x slowest) stored at offset
 * sil_batched_poisson2d_offset(h, p) of every concatenated array; kappa is
 * the nodal coefficient (device, concatenated; edge value = mean of its two
 * nodes).  The solve is CG (settings.preconditioner: Jacobi or identity)
 * with FP32 operator applications and recurrence and FP64 x, dots and
 * residual replacement; a problem converges only on its FP64 residual
 * ‖b − A x‖2 ≤ tol·‖b‖2. */
typedef struct sil_batch_result {
    int32_t iterations;
    int32_t converged;
    double true_relative_residual;
} sil_batch_result;

typedef struct sil_batched_poisson_s* sil_batched_poisson;
int sil_batched_poisson2d_create(sil_batched_poisson* out, int32_t nproblems, const int32_t* sides,
                                 const double* kappa);
int64_t sil_batched_poisson2d_offset(sil_batched_poisson h, int32_t problem);
/* workgroup size of the persistent kernel (256, 512 or 1024) */
int sil_batched_poisson2d_set_block(sil_batched_poisson h, int32_t block);
/* times every workgroup size on `probe_iterations` iterations of every
 * problem and keeps the fastest */
int sil_batched_poisson2d_autotune(sil_batched_poisson h, const double* rhs, int32_t probe_iterations,
                                   int32_t* chosen_block, double* seconds_per_candidate);
/* the wider autotuner: the wide-path threshold (problems of at least
 * wide_side nodes per side take the multi-workgroup path; candidates 96,
 * 128, 192, 256, 384 and none, one per distinct split), the workgroups per
 * wide problem (the default of four nodes per thread, its half and quarter)
 * and the persistent kernel's workgroup size, each timed on
 * `probe_iterations` iterations of every problem; keeps the fastest */
typedef struct sil_batch_shape {
    int32_t wide_side;
    int32_t wide_workgroups;
    int32_t block;
    int32_t wide_problems;
    double seconds;
} sil_batch_shape;
int sil_batched_poisson2d_autotune_shape(sil_batched_poisson h, const double* rhs, int32_t probe_iterations,
                                         sil_batch_shape* chosen);
/* results: host array of nproblems entries (or NULL); seconds: device time
 * of the graph launch (or NULL) */
int sil_batched_poisson2d_solve(sil_batched_poisson h, const double* rhs, double* solution,
                                const sil_solver_settings* settings, sil_batch_result* results, double* seconds,
                                void* stream);
int sil_batched_poisson2d_destroy(sil_batched_poisson h);

/* ------------------------------------------------------------------ */
/* learned preconditioner                                              */
/* ------------------------------------------------------------------ */

/* No reference counterpart.  A 4-level convolutional U-Net (32 channels,
 * 3x3 kernels, ReLU, 2x2 average pooling, nearest upsampling with additive
 * skips) with BF16 weights and activations and FP32 accumulation on the
 * matrix cores.  Input [B][H][W][2] FP32 (residual / rms, normalised log κ),
 * output [B][H][W] FP32.  Parameters are one FP32 vector of
 * sil_unet_param_count() entries: conv_in W[9][32][2] b[32]; twelve 32->32
 * layers W[9][32][32] ([tap][out][in], tap = 3 dy + dx) b[32]; conv_out
 * W[9][32] b[1].  H and W must be multiples of 8.  All pointers are device
 * pointers; every call is asynchronous on `stream`. */
typedef struct sil_unet_s* sil_unet;
int64_t sil_unet_param_count(void);
int sil_unet_create(sil_unet* out, int32_t max_batch, int32_t height, int32_t width, uint64_t seed);
int sil_unet_destroy(sil_unet net);
int sil_unet_get_params(sil_unet net, float* params, void* stream);
int sil_unet_set_params(sil_unet net, const float* params, void* stream);
int sil_unet_forward(sil_unet net, int32_t batch, const float* input, float* output, void* stream);
/* the same output as sil_unet_forward, bit for bit, for inference: the
 * input layer and layer 1 run as one kernel and the input layer's
 * activation is not stored (no sil_unet_backward may follow) */
int sil_unet_infer(sil_unet net, int32_t batch, const float* input, float* output, void* stream);
/* the stored BF16 activation `which` of the preceding sil_unet_forward, as
 * FP32 [batch][H_l][W_l][32] (verification of single layers): 0 a0, 1 e0,
 * 2 a1, 3 e1, 4 a2, 5 e2, 6 a3, 7 b3, 8 c2, 9 f2, 10 c1, 11 f1, 12 c0 — the
 * outputs of conv_in and layers 1..12 in order; levels 0 0 1 1 2 2 3 3 2 2 1
 * 1 0 (H_l = H >> l) */
int sil_unet_activation(sil_unet net, int32_t which, int32_t batch, float* output, void* stream);
/* parameter gradient of <dout, forward(input)>; uses the activations of the
 * preceding sil_unet_forward on the same input */
int sil_unet_backward(sil_unet net, int32_t batch, const float* input, const float* dout, float* grad,
                      void* stream);
/* direction loss of an output N against the error e of the residual
 * r = input[..., 0] = A e: loss[b] = 1 - (r.N)^2 / ((e.r)(N.A N)), A the
 * Dirichlet-zero 5-point operator with edge weights wx, wy [B][H][W];
 * dout = d(mean loss)/dN.  Square images only. */
int sil_unet_loss(sil_unet net, int32_t batch, const float* output, const float* input, const float* err,
                  const float* wx, const float* wy, float* dout, float* loss, void* stream);
/* forward, loss, backward and one Adam step (β 0.9 / 0.999, ε 1e-8) */
int sil_unet_train_step(sil_unet net, int32_t batch, const float* input, const float* err, const float* wx,
                        const float* wy, float learning_rate, float* loss, void* stream);

/* ------------------------------------------------------------------ */
/* equal-size batches of 2D Poisson problems                           */
/* ------------------------------------------------------------------ */

/* No reference counterpart.  B Dirichlet-zero problems on n x n unit-spaced
 * grids, arrays [B][n][n] (x slowest), nodal coefficient kappa (edge value
 * = mean of its two nodes).  Methods, chosen by the settings:
 *   JACOBI / IDENTITY: CG; precision FP64, or MIXED (FP32 directions and
 *   operator applications, FP64 iterate, residual, dots and scalars);
 *   LEARNED: flexible CG (truncation 1) whose preconditioner is a two-level
 *   cycle: damped Jacobi pre-smoothing, the network's correction of the
 *   remaining residual scaled to its A-optimal length, damped Jacobi
 *   post-smoothing (`net` must be built for batch >= B of n x n); a problem
 *   whose FP64 residual norm has not fallen by 0.8 over five consecutive
 *   replacement intervals continues with the Jacobi preconditioner (plain
 *   CG is the fallback).
 * A problem converges only on an FP64 residual b - A x recomputed every
 * replace_interval iterations (even; 0 = 10 for CG, 4 for flexible CG).
 * results (host, B entries) report iterations and the FP64 relative
 * residual of every problem at exit; seconds is the host wall time of the
 * solve.  An optional record stores copies of r (FP32) and x (FP64) of all problems
 * before every graph of `every` iterations (the training data of the
 * learned preconditioner). */
typedef struct sil_solve_record {
    int32_t every;       /* even, >= 2; overrides replace_interval */
    int32_t max_records;
    float* r;            /* device [max_records][B][n][n], FP32 residuals */
    double* x;           /* device [max_records][B][n][n], FP64 iterates */
    int32_t* count;      /* host: records written */
} sil_solve_record;

typedef struct sil_poisson2d_batch_s* sil_poisson2d_batch;
int sil_poisson2d_batch_create(sil_poisson2d_batch* out, int32_t batch, int32_t n, const double* kappa);
int sil_poisson2d_batch_destroy(sil_poisson2d_batch h);
int sil_poisson2d_batch_apply(sil_poisson2d_batch h, const double* x, double* y, void* stream);
/* FP32 edge weights and normalised log κ (any pointer may be NULL) */
int sil_poisson2d_batch_weights(sil_poisson2d_batch h, float* wx, float* wy, float* logk, void* stream);
int sil_poisson2d_batch_solve(sil_poisson2d_batch h, const double* rhs, double* solution,
                              const sil_solver_settings* settings, sil_unet net, sil_batch_result* results,
                              double* seconds, const sil_solve_record* record, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SIL_HIP_H */
