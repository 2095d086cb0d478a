// Assembled (CSR) operators.
//
// * The 3D magnetostatics operator assembled into a CSR matrix on the device
//   (assemble_matrix_data, onelab_interface/magnetostatics_onelab.hpp
//   :3115-3536, packed as csr_from_matrix_data does,
//   src/similie/solvers/minimize_strong_formulation_residual.hpp:565-621).
//   The reference gathers every row through precomputed transposed-moment
//   lists (≤ 4 slots per DOF, filled by atomics) and merges repeated columns
//   with a linear search over a 128-entry local array.  Here the row pattern
//   is fixed by the stencil: for a row of component C at node n, the (moment,
//   tap) pairs that reach it, the sampled nodes s = n − d_tap they come from
//   and the ≤ 13 columns they produce are compile-time tables, so a thread
//   accumulates its row in registers in one pass with no search, adds the
//   contributions in the order the serial transposed lists hold them
//   (sampled node ascending, then moment), and writes the row with ascending
//   columns.  One thread handles the three rows of a node; a count pass, a
//   scan and a fill pass produce row pointers, columns and values.
// * y = A x with L lanes per row (L from the mean row length) and an
//   optional per-block partial sum of ⟨w, y⟩.
// * The Jacobi diagonal (gko::preconditioner::Jacobi, max block size 1, the
//   reference's default preconditioner, :921-930).
// * SpdIsai (:963-967): Ginkgo v1.11.0's factorized sparse approximate inverse on
//   the pattern of tril(A) — for row i, J = stored columns ≤ i, solve
//   A[J,J] y = e_|J| and scale by 1/√y_|J|; with A[J,J] = L Lᵀ this is
//   G[i,J] = L⁻ᵀ e_|J|, one Cholesky factorisation and one back substitution
//   per row, in registers.  Gᵀ is formed explicitly so that both products of
//   z = Gᵀ (G r) are gathers.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>

#include "operators.hpp"

namespace sil {

namespace {

constexpr int kAsmBlock = 128;
constexpr int kSpmvBlock = 256;
constexpr int kSpmvMaxBlocks = 8192;
constexpr int kFsaiMaxRow = 32;
constexpr int kXcds = 8; // MI355X: 8 XCDs, each with its own L2

// ----------------------------------------------------------- stencil table ---

struct Tap {
    int di, dj, dk, comp, coeff;
};

// Moment stencils of the induction (magnetostatics_quantities.hpp:173-196):
// B_x = (dA)_yz, B_y = −(dA)_xz, B_z = (dA)_xy as forward differences at s.
constexpr Tap kMoment[3][4] = {
    {{0, 1, 0, 2, 1}, {0, 0, 0, 2, -1}, {0, 0, 1, 1, -1}, {0, 0, 0, 1, 1}},
    {{1, 0, 0, 2, -1}, {0, 0, 0, 2, 1}, {0, 0, 1, 0, 1}, {0, 0, 0, 0, -1}},
    {{1, 0, 0, 1, 1}, {0, 0, 0, 1, -1}, {0, 1, 0, 0, -1}, {0, 0, 0, 0, 1}},
};

// column key of a relative offset (each in −1..1) and component: ascending
// keys are ascending DOF indices whenever nx, ny ≥ 3
constexpr int key_of(int di, int dj, int dk, int comp) { return (((dk + 1) * 3 + (dj + 1)) * 3 + (di + 1)) * 3 + comp; }

// rank of the sampled node s = n − d in linear-index order: n−e_z, n−e_y, n−e_x, n
constexpr int sampled_rank(const Tap& t) { return t.dk ? 0 : (t.dj ? 1 : (t.di ? 2 : 3)); }

struct RowPlan {
    int pair_m[4], pair_t[4]; // (moment, tap) pairs reaching the row, in slot order
    int npairs;
    int slot_of_key[81];      // -1: column not in the pattern
    int key_of_slot[32];
    int nslots;
    int diag_slot;
};

// JAC: the row of the Jacobian at a state, whose response couples every
// moment of a sampled node with the three moments there (∂²H/∂B_m∂B_m2):
// each pair reaches the taps of all three moment stencils at its sampled node.
// DIVG: the divergence gauge's columns — the codifferential rows n = m and
// n = m + e_C each reach (n − e_d, d) and (n, d).
constexpr RowPlan make_plan(int C, bool jac = false, bool divg = false)
{
    RowPlan p{};
    int n = 0;
    for (int rank = 0; rank < 4; ++rank)
        for (int m = 0; m < 3; ++m)
            for (int t = 0; t < 4; ++t)
                if (kMoment[m][t].comp == C && sampled_rank(kMoment[m][t]) == rank) {
                    p.pair_m[n] = m;
                    p.pair_t[n] = t;
                    ++n;
                }
    p.npairs = n;
    bool used[81] = {};
    used[key_of(0, 0, 0, C)] = true;
    for (int q = 0; q < n; ++q) {
        const Tap& t = kMoment[p.pair_m[q]][p.pair_t[q]];
        for (int m2 = 0; m2 < 3; ++m2) {
            if (!jac && m2 != p.pair_m[q]) continue;
            for (int u = 0; u < 4; ++u) {
                const Tap& v = kMoment[m2][u];
                used[key_of(v.di - t.di, v.dj - t.dj, v.dk - t.dk, v.comp)] = true;
            }
        }
    }
    if (divg)
        for (int o = 0; o < 2; ++o)
            for (int d = 0; d < 3; ++d) {
                const int ox = o && C == 0, oy = o && C == 1, oz = o && C == 2;
                used[key_of(ox - (d == 0), oy - (d == 1), oz - (d == 2), d)] = true;
                used[key_of(ox, oy, oz, d)] = true;
            }
    int s = 0;
    for (int k = 0; k < 81; ++k) {
        p.slot_of_key[k] = used[k] ? s : -1;
        if (used[k]) p.key_of_slot[s++] = k;
    }
    p.nslots = s;
    p.diag_slot = p.slot_of_key[key_of(0, 0, 0, C)];
    return p;
}

template <int C, bool JAC = false, bool DIVG = false>
struct Plan {
    static constexpr RowPlan v = make_plan(C, JAC, DIVG);
    static_assert(v.npairs == 4 && v.nslots == (JAC ? 23 : 13) + (DIVG ? 2 : 0),
                  "curl-curl row pattern (13 entries at most, 23 for the Jacobian; the divergence gauge adds "
                  "(m ± e_C, C))");
};

// largest row of a plan
constexpr int row_max(bool jac, bool divg) { return (jac ? 23 : 13) + (divg ? 2 : 0); }

struct MsAsm {
    int64_t nx, ny, nz;
    double penalty;
    const double* ls[3];
    const double* nu;
    double ysign = 1.0;
    // divergence gauge: factor and coordinates (cell edges)
    double gauge_factor = 0.0;
    const double* coord[3] = {nullptr, nullptr, nullptr};
    // Jacobian at a state: ∂²H/∂B_m∂B_m2 = a δ_m,m2 + c B0_m B0_m2 per node
    const double* st_a = nullptr;
    const double* st_c = nullptr;
    const double* st_b[3] = {nullptr, nullptr, nullptr};
};

__device__ __forceinline__ bool interior(const MsAsm& a, int64_t i, int64_t j, int64_t k)
{
    return i > 0 && j > 0 && k > 0 && i + 1 < a.nx && j + 1 < a.ny && k + 1 < a.nz;
}

// one interior row (component C at node (i,j,k)): COUNT writes the number of
// stored entries, otherwise the row is written at col/val[rp[r] − base]
__device__ __forceinline__ double asm_cell_edge(const double* c, int64_t n, int64_t i)
{
    return i + 1 < n ? c[i + 1] - c[i] : c[i] - c[i - 1];
}

template <int C, bool COUNT, bool JAC = false, bool DIVG = false>
__device__ __forceinline__ void assemble_row(const MsAsm& a, int64_t i, int64_t j, int64_t k, int64_t* __restrict__ cnt,
                                             const int32_t* __restrict__ rp, int32_t* __restrict__ col,
                                             double* __restrict__ val, int32_t base)
{
#pragma clang fp contract(off)
    constexpr RowPlan P = Plan<C, JAC, DIVG>::v;
    const int64_t node = i + a.nx * (j + a.ny * k);
    const int64_t r = 3 * node + C;
    double acc[P.nslots];
#pragma unroll
    for (int s = 0; s < P.nslots; ++s) acc[s] = 0.0;
    uint32_t present = 0;
    if constexpr (!DIVG) {
        const double lx = a.ls[0][i], ly = a.ls[1][j], lz = a.ls[2][k];
        const double wa = C == 0 ? ly * lz / lx : (C == 1 ? lx * lz / ly : lx * ly / lz);
        const double g = a.penalty * wa;
        if (g != 0.0) {
            acc[P.diag_slot] += g;
            present |= 1u << P.diag_slot;
        }
    } else {
        // the divergence gauge (:3204-3227): for the codifferential rows
        // n = m (coefficient −1/e_C(m)²) and n = m + e_C (1/(e_C(n) e_C(m))),
        // factor·w(n)·ρ·c over row n's entries (n − e_d, d) then (n, d)
        const int64_t ext[3] = {a.nx, a.ny, a.nz}, idx[3] = {i, j, k};
        const double ec = asm_cell_edge(a.coord[C], ext[C], idx[C]);
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            const int64_t n3[3] = {i + (o && C == 0), j + (o && C == 1), k + (o && C == 2)};
            const double rho = o == 0 ? -1.0 / (ec * ec) : 1.0 / (asm_cell_edge(a.coord[C], ext[C], n3[C]) * ec);
            const double w = a.ls[0][n3[0]] * a.ls[1][n3[1]] * a.ls[2][n3[2]];
            double e[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) e[d] = asm_cell_edge(a.coord[d], ext[d], n3[d]);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                int64_t q[3] = {n3[0], n3[1], n3[2]};
                q[d] -= 1;
                if (!interior(a, q[0], q[1], q[2])) continue;
                const double cf = 1.0 / (e[d] * asm_cell_edge(a.coord[d], ext[d], q[d]));
                const double v = a.gauge_factor * w * rho * cf;
                if (v == 0.0) continue;
                const int slot = P.slot_of_key[key_of((int)(q[0] - i), (int)(q[1] - j), (int)(q[2] - k), d)];
                acc[slot] += v;
                present |= 1u << slot;
            }
            if (interior(a, n3[0], n3[1], n3[2])) {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double v = a.gauge_factor * w * rho * (-1.0 / (e[d] * e[d]));
                    if (v == 0.0) continue;
                    const int slot =
                        P.slot_of_key[key_of((int)(n3[0] - i), (int)(n3[1] - j), (int)(n3[2] - k), d)];
                    acc[slot] += v;
                    present |= 1u << slot;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < P.npairs; ++q) {
        const int m = P.pair_m[q];
        const Tap t = kMoment[m][P.pair_t[q]];
        const int64_t si = i - t.di, sj = j - t.dj, sk = k - t.dk;
        if (!interior(a, si, sj, sk)) continue;
        const double sx = a.ls[0][si], sy = a.ls[1][sj], sz = a.ls[2][sk];
        const double hf = m == 0 ? sx / (sy * sz) : (m == 1 ? a.ysign * (sy / (sx * sz)) : sz / (sx * sy));
        const int64_t sn = si + a.nx * (sj + a.ny * sk);
        if constexpr (!JAC) {
            const double jc = hf * a.nu[sn];
            if (jc == 0.0) continue;
            const double rjc = t.coeff * jc;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const Tap v = kMoment[m][u];
                if (!interior(a, si + v.di, sj + v.dj, sk + v.dk)) continue;
                const int slot = P.slot_of_key[key_of(v.di - t.di, v.dj - t.dj, v.dk - t.dk, v.comp)];
                acc[slot] += rjc * v.coeff;
                present |= 1u << slot;
            }
        } else {
            const double sa = a.st_a[sn], sc = a.st_c[sn];
            const double b0[3] = {a.st_b[0][sn], a.st_b[1][sn], a.st_b[2][sn]};
#pragma unroll
            for (int m2 = 0; m2 < 3; ++m2) {
                const double jmm = (m == m2 ? sa : 0.0) + sc * b0[m] * b0[m2];
                const double jc = hf * jmm;
                if (jc == 0.0) continue;
                const double rjc = t.coeff * jc;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const Tap v = kMoment[m2][u];
                    if (!interior(a, si + v.di, sj + v.dj, sk + v.dk)) continue;
                    const int slot = P.slot_of_key[key_of(v.di - t.di, v.dj - t.dj, v.dk - t.dk, v.comp)];
                    acc[slot] += rjc * v.coeff;
                    present |= 1u << slot;
                }
            }
        }
    }
    if constexpr (COUNT) {
        cnt[r] = __popc(present);
    } else {
        int32_t out = rp[r] - base;
#pragma unroll
        for (int s = 0; s < P.nslots; ++s) {
            if (!((present >> s) & 1u)) continue;
            const int key = P.key_of_slot[s];
            const int comp = key % 3, di = (key / 3) % 3 - 1, dj = (key / 9) % 3 - 1, dk = key / 27 - 1;
            col[out] = static_cast<int32_t>(3 * (node + di + a.nx * (dj + a.ny * dk)) + comp);
            val[out] = acc[s];
            ++out;
        }
    }
}

// count pass: entries per row
template <bool JAC, bool DIVG>
__global__ __launch_bounds__(kAsmBlock) void k_ms3d_count(MsAsm a, int64_t* __restrict__ cnt)
{
    const int64_t nn = a.nx * a.ny * a.nz;
    for (int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; node < nn;
         node += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = node % a.nx, j = (node / a.nx) % a.ny, k = node / (a.nx * a.ny);
        if (!interior(a, i, j, k)) {
            for (int c = 0; c < 3; ++c) cnt[3 * node + c] = 1;
            continue;
        }
        assemble_row<0, true, JAC, DIVG>(a, i, j, k, cnt, nullptr, nullptr, nullptr, 0);
        assemble_row<1, true, JAC, DIVG>(a, i, j, k, cnt, nullptr, nullptr, nullptr, 0);
        assemble_row<2, true, JAC, DIVG>(a, i, j, k, cnt, nullptr, nullptr, nullptr, 0);
    }
}

// fill pass: a workgroup assembles the rows of kAsmBlock consecutive nodes —
// one contiguous stretch of the CSR arrays — into LDS, then writes the
// stretch with consecutive lanes on consecutive entries (the per-thread rows
// written directly are scattered 4- and 8-byte stores: partial-line writes)
template <bool JAC, bool DIVG>
__global__ __launch_bounds__(kAsmBlock) void k_ms3d_fill(MsAsm a, const int32_t* __restrict__ rp,
                                                         int32_t* __restrict__ col, double* __restrict__ val)
{
    constexpr int kAsmRowMax = row_max(JAC, DIVG);
    constexpr int kNodes = JAC ? kAsmBlock / 2 : kAsmBlock; // nodes per pass (LDS)
    __shared__ int32_t s_col[kNodes * 3 * kAsmRowMax];
    __shared__ double s_val[kNodes * 3 * kAsmRowMax];
    const int64_t nn = a.nx * a.ny * a.nz;
    for (int64_t node0 = (int64_t)blockIdx.x * kNodes; node0 < nn; node0 += (int64_t)gridDim.x * kNodes) {
        const int64_t node_end = node0 + kNodes < nn ? node0 + kNodes : nn;
        const int32_t beg = rp[3 * node0], len = rp[3 * node_end] - beg;
        const int64_t node = node0 + threadIdx.x;
        if (threadIdx.x < kNodes && node < node_end) {
            const int64_t i = node % a.nx, j = (node / a.nx) % a.ny, k = node / (a.nx * a.ny);
            if (!interior(a, i, j, k)) {
                for (int c = 0; c < 3; ++c) {
                    const int64_t r = 3 * node + c;
                    s_col[rp[r] - beg] = static_cast<int32_t>(r);
                    s_val[rp[r] - beg] = 1.0;
                }
            } else {
                assemble_row<0, false, JAC, DIVG>(a, i, j, k, nullptr, rp, s_col, s_val, beg);
                assemble_row<1, false, JAC, DIVG>(a, i, j, k, nullptr, rp, s_col, s_val, beg);
                assemble_row<2, false, JAC, DIVG>(a, i, j, k, nullptr, rp, s_col, s_val, beg);
            }
        }
        __syncthreads();
        for (int32_t q = threadIdx.x; q < len; q += kAsmBlock) {
            col[beg + q] = s_col[q];
            val[beg + q] = s_val[q];
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------- helpers ---

inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 16)
{
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, cap)));
}

// row pointers from per-row counts (int64): rp32[0] = 0, rp32[r+1] = Σ_{q≤r}
// counts; fails when the total exceeds the int32 index range
__global__ void k_rowptr_narrow(const int64_t* __restrict__ incl, int32_t* __restrict__ rp, int64_t n)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r <= n; r += (int64_t)gridDim.x * blockDim.x)
        rp[r] = r == 0 ? 0 : static_cast<int32_t>(incl[r - 1]);
}

int rowptr_from_counts(const int64_t* counts, int64_t n, int32_t* rp, int64_t* nnz_out, hipStream_t s)
{
    int64_t* incl = nullptr;
    void* tmp = nullptr;
    size_t bytes = 0;
    int rc = SIL_OK;
    if (hipcub::DeviceScan::InclusiveSum(nullptr, bytes, counts, incl, static_cast<int>(n), s) != hipSuccess) {
        set_error("row pointer scan: temporary size query failed");
        return SIL_ERR_HIP;
    }
    SIL_CHECK_HIP(hipMalloc(&incl, sizeof(int64_t) * std::max<int64_t>(n, 1)));
    if (hipMalloc(&tmp, bytes) != hipSuccess) {
        (void)hipFree(incl);
        set_error("row pointer scan: allocation failed");
        return SIL_ERR_HIP;
    }
    int64_t total = 0;
    if (n > 0 && hipcub::DeviceScan::InclusiveSum(tmp, bytes, counts, incl, static_cast<int>(n), s) != hipSuccess) {
        set_error("row pointer scan failed");
        rc = SIL_ERR_HIP;
    }
    if (rc == SIL_OK && n > 0 &&
        (hipMemcpyAsync(&total, incl + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess)) {
        set_error("row pointer scan: reading the total failed");
        rc = SIL_ERR_HIP;
    }
    if (rc == SIL_OK && total > INT32_MAX) {
        set_error("CSR: %lld stored entries exceed the int32 index range", (long long)total);
        rc = SIL_ERR_INVALID;
    }
    if (rc == SIL_OK) {
        hipLaunchKernelGGL(k_rowptr_narrow, dim3(grid_for(n + 1, 256)), dim3(256), 0, s, incl, rp, n);
        if (hipGetLastError() != hipSuccess) {
            set_error("row pointer kernel launch failed");
            rc = SIL_ERR_HIP;
        }
    }
    if (rc == SIL_OK && hipStreamSynchronize(s) != hipSuccess) {
        set_error("row pointer kernel failed");
        rc = SIL_ERR_HIP;
    }
    (void)hipFree(tmp);
    (void)hipFree(incl);
    *nnz_out = total;
    return rc;
}

// ------------------------------------------------------------ sparse product ---

struct CsrView {
    int64_t n;
    const int32_t* rp;
    const int32_t* col;
    const double* val;
};

inline CsrView view(const CsrData& a) { return CsrView{a.n, a.row_ptr, a.col, a.val}; }

// y = A x with L lanes per row; with `w` set, per-block partial sums of
// ⟨w, y⟩ go to partials[blockIdx].  A workgroup takes R groups of 256/L rows
// per step; each lane first issues the K column and value loads of its share
// of each of its R rows (independent of one another, streamed past the
// caches: the matrix is read once per product), then the R·K gathers of x,
// so a wave keeps 2·R·K loads per lane in flight instead of one row-pointer →
// column → x chain; entries beyond K·L take a plain loop.
template <int L, int K, int R>
__global__ __launch_bounds__(kSpmvBlock) void k_csr_spmv(CsrView A, const double* __restrict__ x,
                                                         double* __restrict__ y, const double* __restrict__ w,
                                                         double* __restrict__ partials, const int* __restrict__ done)
{
    if (done && *done) return;
    constexpr int kRows = kSpmvBlock / L;
    constexpr int kStep = kRows * R;
    const int lane = threadIdx.x % L;
    // workgroups are dealt to the XCDs round-robin: XCD x takes the x-th
    // eighth of the rows, so the x planes one L2 serves stay few
    const int xcd = blockIdx.x % kXcds;
    const int64_t local = blockIdx.x / kXcds, per = gridDim.x / kXcds;
    const int64_t span = ((A.n + kXcds - 1) / kXcds + kStep - 1) / kStep * kStep;
    const int64_t lo = xcd * span, hi = lo + span < A.n ? lo + span : A.n;
    double dot = 0.0;
    for (int64_t base = lo + local * kStep; base < hi; base += per * kStep) {
        int64_t row[R];
        int32_t b[R], e[R];
#pragma unroll
        for (int g = 0; g < R; ++g) {
            row[g] = base + g * kRows + threadIdx.x / L;
            b[g] = row[g] < hi ? A.rp[row[g]] : 0;
            e[g] = row[g] < hi ? A.rp[row[g] + 1] : 0;
        }
        int32_t c[R][K];
        double a[R][K];
#pragma unroll
        for (int g = 0; g < R; ++g)
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int32_t q = b[g] + lane + k * L;
                c[g][k] = q < e[g] ? __builtin_nontemporal_load(&A.col[q]) : -1;
                a[g][k] = q < e[g] ? __builtin_nontemporal_load(&A.val[q]) : 0.0;
            }
        double v[R];
#pragma unroll
        for (int g = 0; g < R; ++g) {
            v[g] = 0.0;
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (c[g][k] >= 0) v[g] += a[g][k] * x[c[g][k]];
            for (int32_t q = b[g] + lane + K * L; q < e[g]; q += L) v[g] += A.val[q] * x[A.col[q]];
        }
#pragma unroll
        for (int g = 0; g < R; ++g) {
#pragma unroll
            for (int off = L / 2; off > 0; off >>= 1) v[g] += __shfl_xor(v[g], off, L);
            if (lane == 0 && row[g] < hi) {
                y[row[g]] = v[g];
                if (w) dot += w[row[g]] * v[g];
            }
        }
    }
    if (partials) {
        using BR = hipcub::BlockReduce<double, kSpmvBlock>;
        __shared__ typename BR::TempStorage tmp;
        const double sum = BR(tmp).Sum(dot);
        if (threadIdx.x == 0) partials[blockIdx.x] = sum;
    }
}

__global__ void k_csr_diag(CsrView A, double* __restrict__ d)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < A.n; r += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int32_t q = A.rp[r]; q < A.rp[r + 1]; ++q)
            if (A.col[q] == r) v = A.val[q];
        d[r] = v;
    }
}

// flag bit 0: row pointers not monotone / out of [0, nnz]; bit 1: column
// out of range; bit 2: columns not strictly ascending
__global__ void k_csr_check(int64_t n, int64_t nnz, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                            int* __restrict__ flag)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = rp[r], e = rp[r + 1];
        if (b < 0 || e < b || e > nnz || (r == 0 && b != 0) || (r == n - 1 && e != nnz)) {
            atomicOr(flag, 1);
            continue;
        }
        for (int32_t q = b; q < e; ++q) {
            if (col[q] < 0 || col[q] >= n) atomicOr(flag, 2);
            if (q > b && col[q] <= col[q - 1]) atomicOr(flag, 4);
        }
    }
}

// ------------------------------------------------------------------ FSAI ---

// lower counts m_i = |{stored j ≤ i}|; flag bit 0: diagonal missing; the
// block maxima of m_i go to maxm
__global__ __launch_bounds__(256) void k_fsai_count(CsrView A, int64_t* __restrict__ cnt, int* __restrict__ flag,
                                                    int* __restrict__ maxm)
{
    int local_max = 0;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < A.n; r += (int64_t)gridDim.x * blockDim.x) {
        int32_t q = A.rp[r];
        const int32_t e = A.rp[r + 1];
        while (q < e && A.col[q] <= r) ++q;
        const int m = q - A.rp[r];
        if (m == 0 || A.col[q - 1] != r) atomicOr(flag, 1);
        cnt[r] = m;
        local_max = std::max(local_max, m);
    }
    using BR = hipcub::BlockReduce<int, 256>;
    __shared__ typename BR::TempStorage tmp;
    const int bm = BR(tmp).Reduce(local_max, hipcub::Max());
    if (threadIdx.x == 0) atomicMax(maxm, bm);
}

// G row i: Cholesky of S = A[J,J] (entry (a,b), b ≤ a, from row J[a]) and
// Lᵀ g = e_m.  flag bit 1: a local system is not positive definite.
template <int M>
__global__ __launch_bounds__(64) void k_fsai_generate(CsrView A, const int32_t* __restrict__ grp,
                                                      int32_t* __restrict__ gcol, double* __restrict__ gval,
                                                      int* __restrict__ flag)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const int m = grp[i + 1] - grp[i];
    const int32_t ab = A.rp[i];
    int J[M];
    double S[M * (M + 1) / 2]; // packed lower triangle, (a, b) at a(a+1)/2 + b
#pragma unroll
    for (int a = 0; a < M; ++a) J[a] = a < m ? A.col[ab + a] : INT32_MAX;
#pragma unroll
    for (int q = 0; q < M * (M + 1) / 2; ++q) S[q] = 0.0;
#pragma unroll
    for (int a = 0; a < M; ++a) {
        if (a < m) {
            const int32_t e = A.rp[J[a] + 1];
            for (int32_t q = A.rp[J[a]]; q < e; ++q) {
                const int32_t c = A.col[q];
                if (c > J[a]) break;
                const double v = A.val[q];
#pragma unroll
                for (int b = 0; b <= a; ++b)
                    if (c == J[b]) S[a * (a + 1) / 2 + b] = v;
            }
        }
    }
    bool ok = true;
#pragma unroll
    for (int a = 0; a < M; ++a) {
        if (a < m) {
#pragma unroll
            for (int b = 0; b < a; ++b) {
                double v = S[a * (a + 1) / 2 + b];
#pragma unroll
                for (int c = 0; c < b; ++c) v -= S[a * (a + 1) / 2 + c] * S[b * (b + 1) / 2 + c];
                S[a * (a + 1) / 2 + b] = v / S[b * (b + 1) / 2 + b];
            }
            double d = S[a * (a + 1) / 2 + a];
#pragma unroll
            for (int c = 0; c < a; ++c) d -= S[a * (a + 1) / 2 + c] * S[a * (a + 1) / 2 + c];
            ok = ok && d > 0.0;
            S[a * (a + 1) / 2 + a] = ok ? sqrt(d) : 1.0;
        }
    }
    if (!ok) atomicOr(flag, 2);
    double g[M];
#pragma unroll
    for (int a = M - 1; a >= 0; --a) {
        if (a >= m) {
            g[a] = 0.0;
            continue;
        }
        double v = a == m - 1 ? 1.0 : 0.0;
#pragma unroll
        for (int c = a + 1; c < M; ++c)
            if (c < m) v -= S[c * (c + 1) / 2 + a] * g[c];
        g[a] = v / S[a * (a + 1) / 2 + a];
    }
    const int32_t gb = grp[i];
#pragma unroll
    for (int a = 0; a < M; ++a) {
        if (a < m) {
            gcol[gb + a] = J[a];
            gval[gb + a] = g[a];
        }
    }
}

// transpose: column counts, then scatter through per-row cursors, then sort
// each row by column (rows of Gᵀ are short)
__global__ void k_t_count(CsrView G, int64_t* __restrict__ cnt)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < G.n; r += (int64_t)gridDim.x * blockDim.x)
        for (int32_t q = G.rp[r]; q < G.rp[r + 1]; ++q)
            atomicAdd(reinterpret_cast<unsigned long long*>(&cnt[G.col[q]]), 1ull);
}

__global__ void k_t_fill(CsrView G, int32_t* __restrict__ cursor, int32_t* __restrict__ tcol,
                         double* __restrict__ tval)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < G.n; r += (int64_t)gridDim.x * blockDim.x)
        for (int32_t q = G.rp[r]; q < G.rp[r + 1]; ++q) {
            const int32_t pos = atomicAdd(&cursor[G.col[q]], 1);
            tcol[pos] = static_cast<int32_t>(r);
            tval[pos] = G.val[q];
        }
}

__global__ void k_sort_rows(int64_t n, const int32_t* __restrict__ rp, int32_t* __restrict__ col,
                            double* __restrict__ val)
{
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = rp[r], e = rp[r + 1];
        for (int32_t q = b + 1; q < e; ++q) {
            const int32_t kc = col[q];
            const double kv = val[q];
            int32_t p = q - 1;
            while (p >= b && col[p] > kc) {
                col[p + 1] = col[p];
                val[p + 1] = val[p];
                --p;
            }
            col[p + 1] = kc;
            val[p + 1] = kv;
        }
    }
}

int alloc_csr(CsrData& a, int64_t n, int64_t nnz)
{
    a.n = n;
    a.nnz = nnz;
    SIL_CHECK_HIP(hipMalloc(&a.row_ptr, sizeof(int32_t) * (n + 1)));
    SIL_CHECK_HIP(hipMalloc(&a.col, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
    SIL_CHECK_HIP(hipMalloc(&a.val, sizeof(double) * std::max<int64_t>(nnz, 1)));
    return SIL_OK;
}

// lanes per row and loads per lane of the sparse product from the mean row
// length; SIL_CSR_SPMV="lanes,loads,row groups[,max workgroups]" overrides
// (tuning)
void choose_lanes(CsrData& a)
{
    const double mean = a.n > 0 ? static_cast<double>(a.nnz) / static_cast<double>(a.n) : 1.0;
    // measured on the 192³ magnetostatics matrix (12.6 entries per row: 8
    // lanes × 2 loads) and its SpdIsai factor (6.8 per row: 4 lanes × 2)
    a.lanes = mean > 10.0 ? 8 : (mean > 5.0 ? 4 : (mean > 2.5 ? 4 : 2));
    a.unroll = mean > 5.0 ? 2 : 1;
    a.groups = 1;
    a.max_blocks = kSpmvMaxBlocks;
    if (const char* e = std::getenv("SIL_CSR_SPMV")) {
        int l = 0, k = 0, r = 0, b = 0;
        const int got = std::sscanf(e, "%d,%d,%d,%d", &l, &k, &r, &b);
        if (got >= 3) {
            a.lanes = l;
            a.unroll = k;
            a.groups = r;
        }
        if (got == 4 && b > 0) a.max_blocks = b;
    }
}

int read_flag(const int* dflag, int* out, hipStream_t s)
{
    SIL_CHECK_HIP(hipMemcpyAsync(out, dflag, sizeof(int), hipMemcpyDeviceToHost, s));
    SIL_CHECK_HIP(hipStreamSynchronize(s));
    return SIL_OK;
}

}  // namespace

void csr_free(CsrData& a)
{
    for (void* p : {static_cast<void*>(a.row_ptr), static_cast<void*>(a.col), static_cast<void*>(a.val)})
        if (p) (void)hipFree(p);
    a = CsrData{};
}

namespace {
// the linear operator (nl == null) or the Jacobian at the state of the last
// evaluation of the nonlinear operator
int assemble_ms3d(const Ms3dData& d, const NlData* nl, CsrData& out, hipStream_t s)
{
    SIL_REQUIRE(3 * d.nx * d.ny * d.nz < INT32_MAX, SIL_ERR_INVALID,
                "assemble: %lld rows exceed the int32 index range", (long long)(3 * d.nx * d.ny * d.nz));
    const int64_t nn = d.nx * d.ny * d.nz, n = 3 * nn;
    MsAsm a{d.nx, d.ny, d.nz, d.penalty, {d.ls[0], d.ls[1], d.ls[2]}, d.nu};
    a.ysign = d.ysign;
    a.gauge_factor = d.gauge_factor;
    for (int q = 0; q < 3; ++q) a.coord[q] = d.coord[q];
    const bool jac = nl != nullptr, divg = d.divg != 0;
    if (jac) {
        a.st_a = nl->st_a;
        a.st_c = nl->st_c;
        for (int q = 0; q < 3; ++q) a.st_b[q] = nl->st_b[q];
    }
    int64_t* cnt = nullptr;
    SIL_CHECK_HIP(hipMalloc(&cnt, sizeof(int64_t) * n));
    CsrData m{};
    int rc = SIL_OK;
    const unsigned g = grid_for(nn, kAsmBlock);
    if (jac && divg)
        hipLaunchKernelGGL((k_ms3d_count<true, true>), dim3(g), dim3(kAsmBlock), 0, s, a, cnt);
    else if (jac)
        hipLaunchKernelGGL((k_ms3d_count<true, false>), dim3(g), dim3(kAsmBlock), 0, s, a, cnt);
    else if (divg)
        hipLaunchKernelGGL((k_ms3d_count<false, true>), dim3(g), dim3(kAsmBlock), 0, s, a, cnt);
    else
        hipLaunchKernelGGL((k_ms3d_count<false, false>), dim3(g), dim3(kAsmBlock), 0, s, a, cnt);
    if (hipGetLastError() != hipSuccess) {
        set_error("assemble: count kernel launch failed");
        rc = SIL_ERR_HIP;
    }
    int64_t nnz = 0;
    if (rc == SIL_OK && hipMalloc(&m.row_ptr, sizeof(int32_t) * (n + 1)) != hipSuccess) {
        set_error("assemble: allocation failed");
        rc = SIL_ERR_HIP;
    }
    if (rc == SIL_OK) rc = rowptr_from_counts(cnt, n, m.row_ptr, &nnz, s);
    (void)hipFree(cnt);
    if (rc == SIL_OK) {
        m.n = n;
        m.nnz = nnz;
        if (hipMalloc(&m.col, sizeof(int32_t) * std::max<int64_t>(nnz, 1)) != hipSuccess ||
            hipMalloc(&m.val, sizeof(double) * std::max<int64_t>(nnz, 1)) != hipSuccess) {
            set_error("assemble: allocation of %lld entries failed", (long long)nnz);
            rc = SIL_ERR_HIP;
        }
    }
    if (rc == SIL_OK) {
        const dim3 gj(grid_for(nn, kAsmBlock / 2));
        if (jac && divg)
            hipLaunchKernelGGL((k_ms3d_fill<true, true>), gj, dim3(kAsmBlock), 0, s, a, m.row_ptr, m.col, m.val);
        else if (jac)
            hipLaunchKernelGGL((k_ms3d_fill<true, false>), gj, dim3(kAsmBlock), 0, s, a, m.row_ptr, m.col, m.val);
        else if (divg)
            hipLaunchKernelGGL((k_ms3d_fill<false, true>), dim3(g), dim3(kAsmBlock), 0, s, a, m.row_ptr, m.col,
                               m.val);
        else
            hipLaunchKernelGGL((k_ms3d_fill<false, false>), dim3(g), dim3(kAsmBlock), 0, s, a, m.row_ptr, m.col,
                               m.val);
        if (hipGetLastError() != hipSuccess) {
            set_error("assemble: fill kernel launch failed");
            rc = SIL_ERR_HIP;
        }
    }
    if (rc != SIL_OK) {
        csr_free(m);
        return rc;
    }
    choose_lanes(m);
    out = m;
    return SIL_OK;
}
}  // namespace

namespace {
// --------------------------------------------------------- 2D assembly ---
//
// assemble_matrix_data of MagnetostaticsOperator2D (magnetostatics_onelab
// .hpp:3554-3748).  An interior row r gathers, for every transposed slot
// (potential criterion: moment-x slots, then moment-y slots; sampled rows
// ascending) and then every outer slot (moments criterion, coefficients
// negated twice), the block of its sampled node s: with h = ∂²H/∂B∂B(s) at
// the state's induction, rc·h_{x,m}·c over the moment-x stencil of s and
// rc·h_{y,m}·c over its moment-y stencil (m the slot's moment), adding each
// nonzero value to its column — summed with an earlier one in the row or
// appended — in that order; zero sums are dropped and columns sorted.  The
// slots of r are fixed by the stencils (B_x(s) = A(s+e_y) − A(s), B_y(s) =
// A(s) − A(s+e_x)): moment x from s = r − e_y (+1) and s = r (−1), moment
// y from s = r − e_x (−1) and s = r (+1), interior sampled nodes only.
struct Ms2Asm {
    int64_t nx, ny;
    int criterion;
    const double* nu;              // linear Hamiltonian (no state)
    const NodeState<double>* st;   // Jacobian: (a, c, B0_x, B0_y) per node
};

__device__ __forceinline__ bool interior2(const Ms2Asm& a, int64_t s)
{
    const int64_t i = s % a.nx, j = s / a.nx;
    return i > 0 && j > 0 && i + 1 < a.nx && j + 1 < a.ny;
}

// the stored entries of interior row r, ascending columns; returns the count
template <bool JAC>
__device__ __forceinline__ int ms2_row_entries(const Ms2Asm& a, int64_t r, int64_t col[8], double val[8])
{
#pragma clang fp contract(off)
    int count = 0;
    auto add = [&](int64_t c, double v) {
        for (int q = 0; q < count; ++q)
            if (col[q] == c) {
                val[q] += v;
                return;
            }
        col[count] = c;
        val[count] = v;
        ++count;
    };
    // sampled node and stencil coefficient of each slot, per moment
    const int64_t sx[2] = {r - a.nx, r}, sy[2] = {r - 1, r};
    const double cx[2] = {1.0, -1.0}, cy[2] = {-1.0, 1.0};
    for (int part = 0; part < 2; ++part) {
        if (part == 0 && !(a.criterion == 0 || a.criterion == 2)) continue;
        if (part == 1 && !(a.criterion == 1 || a.criterion == 2)) continue;
        for (int m = 0; m < 2; ++m) {
            for (int k = 0; k < 2; ++k) {
                const int64_t s = m == 0 ? sx[k] : sy[k];
                if (s < 0 || !interior2(a, s)) continue;
                // the outer slot holds −c and is applied with a minus sign
                const double rc = part == 0 ? (m == 0 ? cx[k] : cy[k]) : -(-(m == 0 ? cx[k] : cy[k]));
                double hx, hy; // h_{x,m}, h_{y,m}
                if constexpr (JAC) {
                    const NodeState<double> q = a.st[s];
                    const double bb[2] = {q.bx, q.by};
                    hx = (m == 0 ? q.a : 0.0) + q.c * bb[m] * bb[0];
                    hy = (m == 1 ? q.a : 0.0) + q.c * bb[m] * bb[1];
                } else {
                    hx = m == 0 ? a.nu[s] : 0.0;
                    hy = m == 1 ? a.nu[s] : 0.0;
                }
                // moment-x stencil of s: A(s) −1, A(s+e_y) +1
                const double v0 = rc * (hx * -1.0), v1 = rc * (hx * 1.0);
                if (v0 != 0.0) add(s, v0);
                if (v1 != 0.0) add(s + a.nx, v1);
                // moment-y stencil of s: A(s) +1, A(s+e_x) −1
                const double w0 = rc * (hy * 1.0), w1 = rc * (hy * -1.0);
                if (w0 != 0.0) add(s, w0);
                if (w1 != 0.0) add(s + 1, w1);
            }
        }
    }
    int kept = 0;
    for (int q = 0; q < count; ++q)
        if (val[q] != 0.0) {
            col[kept] = col[q];
            val[kept] = val[q];
            ++kept;
        }
    for (int p = 1; p < kept; ++p) {
        const int64_t kc = col[p];
        const double kv = val[p];
        int b = p - 1;
        while (b >= 0 && col[b] > kc) {
            col[b + 1] = col[b];
            val[b + 1] = val[b];
            --b;
        }
        col[b + 1] = kc;
        val[b + 1] = kv;
    }
    return kept;
}

template <bool JAC>
__global__ __launch_bounds__(256) void k_ms2d_count(Ms2Asm a, int64_t* __restrict__ cnt)
{
    const int64_t n = a.nx * a.ny;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        if (!interior2(a, r)) {
            cnt[r] = 1;
            continue;
        }
        int64_t col[8];
        double val[8];
        cnt[r] = ms2_row_entries<JAC>(a, r, col, val);
    }
}

template <bool JAC>
__global__ __launch_bounds__(256) void k_ms2d_fill(Ms2Asm a, const int32_t* __restrict__ rp, int32_t* __restrict__ col,
                                                   double* __restrict__ val)
{
    const int64_t n = a.nx * a.ny;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t o = rp[r];
        if (!interior2(a, r)) {
            col[o] = static_cast<int32_t>(r);
            val[o] = 1.0;
            continue;
        }
        int64_t c[8];
        double v[8];
        const int k = ms2_row_entries<JAC>(a, r, c, v);
        for (int q = 0; q < k; ++q) {
            col[o + q] = static_cast<int32_t>(c[q]);
            val[o + q] = v[q];
        }
    }
}

int assemble_ms2d(const Ms2dData& d, const NlData* nl, CsrData& out, hipStream_t s)
{
    const int64_t n = d.nx * d.ny;
    SIL_REQUIRE(n < INT32_MAX, SIL_ERR_INVALID, "assemble: %lld rows exceed the int32 index range", (long long)n);
    Ms2Asm a{d.nx, d.ny, d.criterion, d.nu, nl ? reinterpret_cast<const NodeState<double>*>(nl->st_pack) : nullptr};
    int64_t* cnt = nullptr;
    SIL_CHECK_HIP(hipMalloc(&cnt, sizeof(int64_t) * std::max<int64_t>(n, 1)));
    CsrData m{};
    int rc = SIL_OK;
    const unsigned g = grid_for(n, 256);
    if (nl)
        hipLaunchKernelGGL(k_ms2d_count<true>, dim3(g), dim3(256), 0, s, a, cnt);
    else
        hipLaunchKernelGGL(k_ms2d_count<false>, dim3(g), dim3(256), 0, s, a, cnt);
    if (hipGetLastError() != hipSuccess) {
        set_error("assemble 2D: count kernel launch failed");
        rc = SIL_ERR_HIP;
    }
    int64_t nnz = 0;
    if (rc == SIL_OK && hipMalloc(&m.row_ptr, sizeof(int32_t) * (n + 1)) != hipSuccess) {
        set_error("assemble 2D: allocation failed");
        rc = SIL_ERR_HIP;
    }
    if (rc == SIL_OK) rc = rowptr_from_counts(cnt, n, m.row_ptr, &nnz, s);
    (void)hipFree(cnt);
    if (rc == SIL_OK) {
        m.n = n;
        m.nnz = nnz;
        if (hipMalloc(&m.col, sizeof(int32_t) * std::max<int64_t>(nnz, 1)) != hipSuccess ||
            hipMalloc(&m.val, sizeof(double) * std::max<int64_t>(nnz, 1)) != hipSuccess) {
            set_error("assemble 2D: allocation of %lld entries failed", (long long)nnz);
            rc = SIL_ERR_HIP;
        }
    }
    if (rc == SIL_OK) {
        if (nl)
            hipLaunchKernelGGL(k_ms2d_fill<true>, dim3(g), dim3(256), 0, s, a, m.row_ptr, m.col, m.val);
        else
            hipLaunchKernelGGL(k_ms2d_fill<false>, dim3(g), dim3(256), 0, s, a, m.row_ptr, m.col, m.val);
        if (hipGetLastError() != hipSuccess) {
            set_error("assemble 2D: fill kernel launch failed");
            rc = SIL_ERR_HIP;
        }
    }
    if (rc != SIL_OK) {
        csr_free(m);
        return rc;
    }
    choose_lanes(m);
    out = m;
    return SIL_OK;
}
}  // namespace

int csr_assemble_ms2d(const Ms2dData& d, CsrData& out, hipStream_t s) { return assemble_ms2d(d, nullptr, out, s); }

int csr_assemble_ms2d_jacobian(const Ms2dData& d, const NlData& nl, CsrData& out, hipStream_t s)
{
    return assemble_ms2d(d, &nl, out, s);
}

int csr_assemble_ms3d(const Ms3dData& d, CsrData& out, hipStream_t s) { return assemble_ms3d(d, nullptr, out, s); }

int csr_assemble_ms3d_jacobian(const Ms3dData& d, const NlData& nl, CsrData& out, hipStream_t s)
{
    return assemble_ms3d(d, &nl, out, s);
}

int csr_import(int64_t n, int64_t nnz, const int32_t* rp, const int32_t* col, const double* val, CsrData& out,
               hipStream_t s)
{
    SIL_REQUIRE(n > 0 && n < INT32_MAX && nnz >= 0 && nnz <= INT32_MAX, SIL_ERR_INVALID,
                "csr_create: size %lld / %lld entries outside the int32 index range", (long long)n, (long long)nnz);
    CsrData m{};
    int rc = alloc_csr(m, n, nnz);
    int* flag = nullptr;
    if (rc == SIL_OK && hipMalloc(&flag, sizeof(int)) != hipSuccess) {
        set_error("csr_create: allocation failed");
        rc = SIL_ERR_HIP;
    }
    if (rc == SIL_OK &&
        (hipMemcpyAsync(m.row_ptr, rp, sizeof(int32_t) * (n + 1), hipMemcpyDefault, s) != hipSuccess ||
         (nnz > 0 && hipMemcpyAsync(m.col, col, sizeof(int32_t) * nnz, hipMemcpyDefault, s) != hipSuccess) ||
         (nnz > 0 && hipMemcpyAsync(m.val, val, sizeof(double) * nnz, hipMemcpyDefault, s) != hipSuccess) ||
         hipMemsetAsync(flag, 0, sizeof(int), s) != hipSuccess)) {
        set_error("csr_create: copy failed");
        rc = SIL_ERR_HIP;
    }
    int h = 0;
    if (rc == SIL_OK) {
        hipLaunchKernelGGL(k_csr_check, dim3(grid_for(n, 256)), dim3(256), 0, s, n, nnz, m.row_ptr, m.col, flag);
        rc = read_flag(flag, &h, s);
    }
    if (flag) (void)hipFree(flag);
    if (rc == SIL_OK && h) {
        set_error("csr_create: %s", (h & 1) ? "row pointers are not monotone from 0 to nnz"
                                            : ((h & 2) ? "a column index is out of range"
                                                       : "columns are not strictly ascending within a row"));
        rc = SIL_ERR_INVALID;
    }
    if (rc != SIL_OK) {
        csr_free(m);
        return rc;
    }
    choose_lanes(m);
    out = m;
    return SIL_OK;
}

int csr_spmv(const CsrData& a, const double* x, double* y, const double* w, double* partials, const int* done,
             int* nb_out, hipStream_t s)
{
    const int rows = kSpmvBlock / a.lanes * a.groups;
    // a multiple of the XCD count (the kernel splits the rows per XCD)
    const int64_t cap = std::max<int64_t>(kXcds, std::min<int64_t>(a.max_blocks, max_dot_blocks()) / kXcds * kXcds);
    const unsigned g = static_cast<unsigned>(
        std::min<int64_t>((static_cast<int64_t>(grid_for(a.n, rows, cap)) + kXcds - 1) / kXcds * kXcds, cap));
    const CsrView v = view(a);
#define SIL_SPMV(L, K, R)                                                                                        \
    if (a.lanes == L && a.unroll == K && a.groups == R) {                                                        \
        hipLaunchKernelGGL((k_csr_spmv<L, K, R>), dim3(g), dim3(kSpmvBlock), 0, s, v, x, y, w, partials, done); \
        SIL_LAUNCH_CHECK();                                                                                      \
        if (nb_out) *nb_out = static_cast<int>(g);                                                               \
        return SIL_OK;                                                                                           \
    }
    SIL_SPMV(2, 1, 1) SIL_SPMV(2, 2, 1) SIL_SPMV(2, 4, 1)
    SIL_SPMV(4, 1, 1) SIL_SPMV(4, 2, 1) SIL_SPMV(4, 2, 2) SIL_SPMV(4, 2, 4) SIL_SPMV(4, 4, 1) SIL_SPMV(4, 4, 2)
    SIL_SPMV(8, 1, 1) SIL_SPMV(8, 1, 2) SIL_SPMV(8, 2, 1) SIL_SPMV(8, 2, 2) SIL_SPMV(8, 2, 4)
    SIL_SPMV(16, 1, 1) SIL_SPMV(16, 1, 2) SIL_SPMV(16, 1, 4)
#undef SIL_SPMV
    set_error("sparse product: no kernel for %d lanes x %d loads x %d row groups", a.lanes, a.unroll, a.groups);
    return SIL_ERR_UNSUPPORTED;
}

int csr_apply64(const CsrData& a, const double* in, double* out, DotEpilogue dot, const PUpdate<double>* pu,
                int* nb_out, hipStream_t s)
{
    const double* x = in;
    const int* done = nullptr;
    if (pu) {
        int rc = direction_pass64(*pu, a.n, s);
        if (rc) return rc;
        x = pu->p_new;
        done = pu->done;
    }
    return csr_spmv(a, x, out, dot.partials ? x : nullptr, dot.partials, done, nb_out, s);
}

int csr_diagonal(const CsrData& a, double* d, hipStream_t s)
{
    hipLaunchKernelGGL(k_csr_diag, dim3(grid_for(a.n, 256)), dim3(256), 0, s, view(a), d);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int spd_isai_generate(const CsrData& a, CsrData& g_out, CsrData& gt_out, hipStream_t s)
{
    const int64_t n = a.n;
    int64_t* cnt = nullptr;
    int* flag = nullptr; // [0] flags, [1] max lower count
    SIL_CHECK_HIP(hipMalloc(&cnt, sizeof(int64_t) * n));
    CsrData g{}, gt{};
    int rc = SIL_OK;
    auto fail = [&](const char* what) {
        set_error("spd_isai: %s", what);
        rc = SIL_ERR_HIP;
    };
    if (hipMalloc(&flag, 2 * sizeof(int)) != hipSuccess || hipMemsetAsync(flag, 0, 2 * sizeof(int), s) != hipSuccess)
        fail("allocation failed");
    int h[2] = {0, 0};
    if (rc == SIL_OK) {
        hipLaunchKernelGGL(k_fsai_count, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, view(a), cnt, flag, flag + 1);
        if (hipMemcpyAsync(h, flag, 2 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            fail("count kernel failed");
    }
    if (rc == SIL_OK && (h[0] & 1)) {
        set_error("spd_isai: a row has no stored diagonal entry");
        rc = SIL_ERR_INVALID;
    }
    if (rc == SIL_OK && h[1] > kFsaiMaxRow) {
        set_error("spd_isai: a row has %d stored entries in its lower triangle (at most %d)", h[1], kFsaiMaxRow);
        rc = SIL_ERR_INVALID;
    }
    int64_t gnnz = 0;
    if (rc == SIL_OK && hipMalloc(&g.row_ptr, sizeof(int32_t) * (n + 1)) != hipSuccess) fail("allocation failed");
    if (rc == SIL_OK) rc = rowptr_from_counts(cnt, n, g.row_ptr, &gnnz, s);
    if (rc == SIL_OK) {
        g.n = n;
        g.nnz = gnnz;
        if (hipMalloc(&g.col, sizeof(int32_t) * std::max<int64_t>(gnnz, 1)) != hipSuccess ||
            hipMalloc(&g.val, sizeof(double) * std::max<int64_t>(gnnz, 1)) != hipSuccess)
            fail("allocation failed");
    }
    if (rc == SIL_OK) {
        const dim3 grid(static_cast<unsigned>((n + 63) / 64));
        if (h[1] <= 8)
            hipLaunchKernelGGL((k_fsai_generate<8>), grid, dim3(64), 0, s, view(a), g.row_ptr, g.col, g.val, flag);
        else if (h[1] <= 16)
            hipLaunchKernelGGL((k_fsai_generate<16>), grid, dim3(64), 0, s, view(a), g.row_ptr, g.col, g.val, flag);
        else
            hipLaunchKernelGGL((k_fsai_generate<32>), grid, dim3(64), 0, s, view(a), g.row_ptr, g.col, g.val, flag);
        if (hipMemcpyAsync(h, flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            fail("generation kernel failed");
    }
    if (rc == SIL_OK && (h[0] & 2)) {
        set_error("spd_isai: a local system A[J,J] is not positive definite");
        rc = SIL_ERR_INVALID;
    }
    // Gᵀ
    int32_t* cursor = nullptr;
    if (rc == SIL_OK && hipMemsetAsync(cnt, 0, sizeof(int64_t) * n, s) != hipSuccess) fail("memset failed");
    if (rc == SIL_OK) {
        hipLaunchKernelGGL(k_t_count, dim3(grid_for(n, 256)), dim3(256), 0, s, view(g), cnt);
        if (hipMalloc(&gt.row_ptr, sizeof(int32_t) * (n + 1)) != hipSuccess) fail("allocation failed");
    }
    int64_t tnnz = 0;
    if (rc == SIL_OK) rc = rowptr_from_counts(cnt, n, gt.row_ptr, &tnnz, s);
    if (rc == SIL_OK) {
        gt.n = n;
        gt.nnz = tnnz;
        if (hipMalloc(&gt.col, sizeof(int32_t) * std::max<int64_t>(tnnz, 1)) != hipSuccess ||
            hipMalloc(&gt.val, sizeof(double) * std::max<int64_t>(tnnz, 1)) != hipSuccess ||
            hipMalloc(&cursor, sizeof(int32_t) * (n + 1)) != hipSuccess ||
            hipMemcpyAsync(cursor, gt.row_ptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToDevice, s) != hipSuccess)
            fail("allocation failed");
    }
    if (rc == SIL_OK) {
        hipLaunchKernelGGL(k_t_fill, dim3(grid_for(n, 256)), dim3(256), 0, s, view(g), cursor, gt.col, gt.val);
        hipLaunchKernelGGL(k_sort_rows, dim3(grid_for(n, 256)), dim3(256), 0, s, n, gt.row_ptr, gt.col, gt.val);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) fail("transpose failed");
    }
    if (cursor) (void)hipFree(cursor);
    (void)hipFree(cnt);
    if (flag) (void)hipFree(flag);
    if (rc != SIL_OK) {
        csr_free(g);
        csr_free(gt);
        return rc;
    }
    choose_lanes(g);
    choose_lanes(gt);
    g_out = g;
    gt_out = gt;
    return SIL_OK;
}

}  // namespace sil
