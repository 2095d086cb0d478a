// Linear operators behind sil_operator and the solver workspace (host side).
#pragma once

#include "common.hpp"

namespace sil {

enum OperatorKind {
    OP_MAGNETOSTATICS3D = 1,
    OP_POISSON2D = 2,
    OP_CSR = 3,
    OP_MS3D_NONLINEAR = 4,
    OP_MS3D_JACOBIAN = 5,
    OP_MAGNETOSTATICS2D = 6,
    OP_MS2D_NONLINEAR = 7,
    OP_MS2D_JACOBIAN = 8
};

struct Workspace;     // CG vectors, scalars and captured graphs (krylov.hip)
struct NewtonBuffers; // Newton vectors (nonlinear.hip)

struct Ms3dData {
    int64_t nx, ny, nz;
    double penalty;
    // switches (magnetostatics_onelab.hpp:796-828): sign of the y moment's
    // response; divergence gauge factor·δᵀWδ (factor = penalty · gauge
    // sign) in place of the penalty term; criterion 0 (potential temporal
    // derivative) zeroes the right-hand side
    double ysign = 1.0;
    int divg = 0;
    double gauge_factor = 0.0;
    int criterion = 1;
    double* ls[3];      // local_spacing per axis (device)
    double* inv_ls[3];  // 1 / local_spacing
    double* coord[3];   // coordinates (device)
    double* nu;         // 1/μ per node (device, FP64)
    float* nu32;        // FP32 copy
    float* ls32[3];
    float* inv_ls32[3];
};

// nonlinear material: interpolated B-H curve (≤ 64 samples, device copy)
// and the per-node state of the last operator evaluation
constexpr int kBhMax = 64;
struct BhCurve {
    int n;
    double b[kBhMax], h[kBhMax], q[kBhMax], dh_db[kBhMax];
};

// a node's material state as one aligned record: a, c and the first two
// components of B0 (the Jacobian's response reads it in one or two vector
// loads)
template <typename T>
struct alignas(4 * sizeof(T)) NodeState {
    T a, c, bx, by;
};

struct NlData {
    const BhCurve* curve = nullptr; // device
    uint8_t* ferro = nullptr;       // per node: ferromagnetic fraction > 0.5
    // state of the last evaluation of F at x, per node: H = w·(a·δB + c·B0·(B0·δB))
    // (3D only: the 2D operator keeps its state in the packed record alone)
    double* st_a = nullptr;
    double* st_c = nullptr;
    double* st_b[3] = {nullptr, nullptr, nullptr};
    // FP32 copies for mixed-precision Jacobian applications: a, and B_z
    // (st_b32[2]); no FP32 kernel reads c, B_x or B_y apart from the packed
    // record, so st_c32 and st_b32[0..1] stay unallocated
    float* st_a32 = nullptr;
    float* st_c32 = nullptr;
    float* st_b32[3] = {nullptr, nullptr, nullptr};
    // the state packed per node as NodeState (a, c, B_x, B_y), FP64 and FP32;
    // the 3D Jacobian takes B_z from st_b[2]
    double* st_pack = nullptr;
    float* st_pack32 = nullptr;
};

// 2D magnetostatics on A_z (one unknown per node, i + nx·j): the moment
// stencils B_x = A_z(s+e_y) − A_z(s), B_y = A_z(s) − A_z(s+e_x) sampled at
// interior nodes, no Hodge weights; `terms` is 2 when the criterion adds the
// potential and the moments parts (each equal), else 1
struct Ms2dData {
    int64_t nx, ny;
    int criterion;      // 0 potential, 1 moments, 2 both temporal derivatives
    double terms;
    double* coord[2];   // coordinates (device)
    double* nu;         // 1/μ per node (device)
    float* nu32;
};

struct Poisson2dData {
    int64_t nx, ny;
    double* wx;   // edge weight (i,j)→(i+1,j), per node (device)
    double* wy;   // edge weight (i,j)→(i,j+1)
    float* wx32;
    float* wy32;
};

// Optional epilogue of an operator application: per-block partial sums of
// ⟨in, out⟩ written to `partials[blockIdx]` (deterministic reduction later).
struct DotEpilogue {
    double* partials; // null: no dot
};

// CG "p = z + β p_old" prologue fused into an application: when enabled the
// kernel forms p from (r, p_old, β, diag) for every node it reads, writes p
// for its own rows, and applies the operator to the new p.
template <typename T>
struct PUpdate {
    const T* r;       // residual
    const T* p_old;
    T* p_new;
    const double* beta;  // device scalar
    const T* inv_diag;   // Jacobi, in the recurrence's arithmetic, or null
    const int* done;     // device flag: skip when set
};

int ms3d_apply64(const Ms3dData& d, const double* in, double* out, DotEpilogue dot, const PUpdate<double>* pu,
                 int* nblocks_out, hipStream_t s);
int ms3d_apply32(const Ms3dData& d, const float* in, float* out, DotEpilogue dot, const PUpdate<float>* pu,
                 int* nblocks_out, hipStream_t s);
// FP32 q = A p with the partials of ⟨p, q⟩, ⟨q, D⁻¹ r⟩ and ⟨q, D⁻¹ q⟩ at
// partials, + pstride, + 2·pstride (inv_diag null: identity); nl: the
// Jacobian at the stored state.  Supported when ms3d_single_pass_ok.
bool ms3d_single_pass_ok(const Ms3dData& d);
int ms3d_qdots32(const Ms3dData& d, const NlData* nl, const float* p, float* q, const float* r,
                 const float* inv_diag, double* partials, int64_t pstride, const int* done, int* nb_out,
                 hipStream_t s);
int ms3d_diagonal(const Ms3dData& d, double* diag, hipStream_t s);
// adds the divergence gauge's diagonal (nothing without the switch)
int ms3d_gauge_diagonal_add(const Ms3dData& d, double* diag, hipStream_t s);
int ms3d_nodal_mu(int64_t nx, int64_t ny, int64_t nz, const double* cell_mu, double mu0, double* mu, hipStream_t s);
int ms3d_rhs(const Ms3dData& d, const double* cell_current, double* rhs, hipStream_t s);
int ms3d_nu_from_mu(const double* mu, double* nu, float* nu32, int64_t n, hipStream_t s);

// nonlinear magnetostatics (nonlinear.hip / operators.hip)
int ms3d_state(const Ms3dData& d, const NlData& nl, const double* x, hipStream_t s);       // state of x
int ms3d_state_apply64(const Ms3dData& d, const NlData& nl, const double* x, double* out, hipStream_t s); // F(x), state given
int ms3d_jacobian_apply64(const Ms3dData& d, const NlData& nl, const double* in, double* out, DotEpilogue dot,
                          const PUpdate<double>* pu, int* nblocks_out, hipStream_t s);
int ms3d_jacobian_apply32(const Ms3dData& d, const NlData& nl, const float* in, float* out, DotEpilogue dot,
                          const PUpdate<float>* pu, int* nblocks_out, hipStream_t s);
int ms3d_jacobian_diagonal(const Ms3dData& d, const NlData& nl, double* diag, hipStream_t s);

// cell B, H, Maxwell stress and force density of a potential (postprocess.hip);
// curve may be null when cell_nonlinear is
int ms3d_post_process(const Ms3dData& d, const BhCurve* curve, const double* potential, const double* cell_mu,
                      const double* cell_nonlinear, double* induction, double* field, double* stress,
                      double* force, hipStream_t s);

// 2D magnetostatics (magnetostatics2d.hip); nl uses st_pack / st_pack32
int ms2d_apply64(const Ms2dData& d, const double* in, double* out, DotEpilogue dot, const PUpdate<double>* pu,
                 int* nblocks_out, hipStream_t s);
int ms2d_apply32(const Ms2dData& d, const float* in, float* out, DotEpilogue dot, const PUpdate<float>* pu,
                 int* nblocks_out, hipStream_t s);
int ms2d_diagonal(const Ms2dData& d, double* diag, hipStream_t s);
int ms2d_nodal_mu(int64_t nx, int64_t ny, const double* cell_mu, double mu0, double* mu, hipStream_t s);
int ms2d_rhs(const Ms2dData& d, const double* cell_current_z, bool zero, double* rhs, hipStream_t s);
int ms2d_state(const Ms2dData& d, const NlData& nl, const double* x, hipStream_t s);
int ms2d_state_apply64(const Ms2dData& d, const NlData& nl, const double* x, double* out, hipStream_t s);
int ms2d_jacobian_apply64(const Ms2dData& d, const NlData& nl, const double* in, double* out, DotEpilogue dot,
                          const PUpdate<double>* pu, int* nblocks_out, hipStream_t s);
int ms2d_jacobian_apply32(const Ms2dData& d, const NlData& nl, const float* in, float* out, DotEpilogue dot,
                          const PUpdate<float>* pu, int* nblocks_out, hipStream_t s);
int ms2d_jacobian_diagonal(const Ms2dData& d, const NlData& nl, double* diag, hipStream_t s);
int ms2d_post_process(const Ms2dData& d, const BhCurve* curve, const double* potential, const double* cell_mu,
                      const double* cell_nonlinear, double* induction, double* field, double* stress,
                      double* force, hipStream_t s);

// p_new = z + β p_old over n rows (the CG direction) as its own pass
int direction_pass64(const PUpdate<double>& pu, int64_t n, hipStream_t s);
int direction_pass32(const PUpdate<float>& pu, int64_t n, hipStream_t s);

// ------------------------------------------------- assembled (CSR) form ---

// square CSR matrix, int32 row pointers and columns (ascending within a
// row), FP64 values — gko::matrix::Csr<double, gko::int32>
struct CsrData {
    int64_t n = 0, nnz = 0;
    int32_t* row_ptr = nullptr; // n + 1
    int32_t* col = nullptr;
    double* val = nullptr;
    int lanes = 16;             // lanes per row in the sparse product
    int unroll = 1;             // loads in flight per lane and row
    int groups = 1;             // rows per lane and step
    int max_blocks = 8192;      // workgroups of the sparse product
};

int csr_assemble_ms3d(const Ms3dData& d, CsrData& out, hipStream_t s);
// the Jacobian of the nonlinear operator at the state of its last evaluation
int csr_assemble_ms3d_jacobian(const Ms3dData& d, const NlData& nl, CsrData& out, hipStream_t s);
// the 2D operator (linear) and its Jacobian at the state of the last evaluation
int csr_assemble_ms2d(const Ms2dData& d, CsrData& out, hipStream_t s);
int csr_assemble_ms2d_jacobian(const Ms2dData& d, const NlData& nl, CsrData& out, hipStream_t s);
int csr_import(int64_t n, int64_t nnz, const int32_t* rp, const int32_t* col, const double* val, CsrData& out,
               hipStream_t s);
void csr_free(CsrData& a);
// y = A x; with w set, per-block ⟨w, y⟩ partials (nb_out = their count);
// returns at once when *done is set
int csr_spmv(const CsrData& a, const double* x, double* y, const double* w, double* partials, const int* done,
             int* nb_out, hipStream_t s);
int csr_apply64(const CsrData& a, const double* in, double* out, DotEpilogue dot, const PUpdate<double>* pu,
                int* nb_out, hipStream_t s);
int csr_diagonal(const CsrData& a, double* d, hipStream_t s);
// SPD sparse approximate inverse M = Gᵀ G: G (pattern of tril(A)) and Gᵀ
int spd_isai_generate(const CsrData& a, CsrData& g, CsrData& gt, hipStream_t s);

int p2d_apply64(const Poisson2dData& d, const double* in, double* out, DotEpilogue dot, const PUpdate<double>* pu,
                int* nblocks_out, hipStream_t s);
int p2d_apply32(const Poisson2dData& d, const float* in, float* out, DotEpilogue dot, const PUpdate<float>* pu,
                int* nblocks_out, hipStream_t s);
int p2d_diagonal(const Poisson2dData& d, double* diag, hipStream_t s);
int p2d_weights(Poisson2dData& d, const double* hx_host, const double* hy_host, const double* kappa, hipStream_t s);

// number of partial-sum slots an application with a dot epilogue may write
int64_t max_dot_blocks();

}  // namespace sil

struct sil_operator_s {
    int kind;
    int64_t n;
    sil::Ms3dData ms;
    sil::Poisson2dData p2;
    sil::Ms2dData m2;
    sil::CsrData csr;
    sil::NlData nl;
    sil::Workspace* ws;
    sil_operator_s* jac; // nonlinear operator: its Jacobian at the current state (shares ms / nl)
    sil::NewtonBuffers* newton;
};
