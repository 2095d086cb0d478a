// Cell post-processing of a magnetostatics solution (3D grids, and 2D grids
// as one slice of cells).
//
// The reference post-processes on the host, cell by cell
// (onelab_interface/magnetostatics_onelab.hpp:3984-4046): the induction
// B_m = moment_m / (area of the cell face normal to m) from the potential's
// forward_value stencil (:3826-3864), the field H from the cell's material
// (h(|B|)/|B|·B on nonlinear cells, B/μ otherwise, unit Hodge factors), the
// Maxwell stress T = B⊗H − ½(B·H)·I in the order xx, yy, zz, xy, xz, yz;
// then, xy slice by xy slice (:4215-4254), the force density as −δ of three
// in-plane one-forms of the stress, (T_xx, T_xy), (T_xy, T_yy), (T_xz, T_yz),
// on the grid of cell centres with an identity metric (:4048-4168), δ being
// the staged codifferential (src/similie/exterior/codifferential.hpp
// :321-498): ⋆ on the one-form, dᵀ with zeros outside the slice, ⋆ back.
//
// On that rectilinear centre grid with spacings hx(i), hy(j) (forward
// differences of the centres, backward at the last one) the three stages
// reduce to
//   η_x = −(hx/hy)·ω_y,  η_y = (hy/hx)·ω_x           (zero where |hx| or |hy| < 1e-14)
//   G   = η_y(i,j) − η_y(i−1,j) − η_x(i,j) + η_x(i,j−1)   (η = 0 outside)
//   f   = G / (hx·hy)                                 (zero where |hx·hy| < 1e-14)
// (the two codifferential signs cancel against the reference's negation).
//
// One fused kernel does everything: a workgroup owns a 64 × 8 tile of cells
// of one slice, evaluates B, H and T on the tile plus the row and column
// below it (the dᵀ neighbours, 14% recomputation) into LDS, writes the
// tile's B, H and T, then the force from LDS.  Per cell it reads the
// potential at the cell's four sampled nodes (mostly cache hits: 24 B of
// HBM per node), μ and the material flag, and writes 15 doubles.
#include <cstdlib>

#include "bh_curve.hpp"
#include "operators.hpp"

#pragma clang fp contract(off)

namespace sil {
namespace {

constexpr int kTx = 64;
constexpr int kTy = 8;
constexpr int kHx = kTx + 1;
constexpr int kHy = kTy + 1;
constexpr int kPostBlock = 256;
constexpr int kXcdCount = 8;

struct PostArgs {
    int64_t nx, ny, nz;     // nodes per axis
    const double* coord[3]; // node coordinates
    const double* a;        // potential [node][3]
    const double* mu;       // [cell]
    const double* nonlinear; // [cell] or null
    const BhCurve* curve;   // device, or null
    double* induction;      // [cell][3] or null
    double* field;          // [cell][3] or null
    double* stress;         // [cell][6] or null
    double* force;          // [cell][3] or null
    int64_t tiles_x, tiles_y, tiles, per_xcd;
};

// spacing of the cell-centre grid at centre i of m centres
__device__ inline double centre_spacing(const double* c, int64_t m, int64_t i)
{
    if (i + 1 < m) return 0.5 * (c[i + 1] + c[i + 2]) - 0.5 * (c[i] + c[i + 1]);
    if (i > 0) return 0.5 * (c[i] + c[i + 1]) - 0.5 * (c[i - 1] + c[i]);
    return 0.0;
}

// TWO_D: one slice of a 2D grid whose potential is A_z per node; the cell
// induction is the reconstructed in-plane moment pair
// (fill_post_process_fields_on_cell_domain, :3866-3982): B_x = (A_z(s+e_y)
// − A_z(s))/h_y, B_y = (A_z(s) − A_z(s+e_x))/h_x, B_z = 0
// STAGED: the tile's outputs go through LDS and leave as whole rows (each
// wavefront store covers 512 contiguous bytes) instead of one 24- or 48-byte
// record per lane
template <bool TWO_D, bool STAGED>
__global__ void __launch_bounds__(kPostBlock) k_ms3d_post(PostArgs p)
{
    static_assert(kTx * kTy == 2 * kPostBlock, "two tile cells per thread");
    __shared__ double s_t[6][kHy * kHx];
    __shared__ double s_o[6][STAGED ? kTx * kTy : 1];

    // XCD-contiguous tiles: workgroup b runs on XCD b % 8, which takes a
    // contiguous range of tiles so vertical neighbours share its L2
    const int64_t b = blockIdx.x;
    const int64_t tile = (b % kXcdCount) * p.per_xcd + b / kXcdCount;
    if (tile >= p.tiles) return;
    const int64_t mx = p.nx - 1, my = p.ny - 1;
    const int64_t tx = tile % p.tiles_x;
    const int64_t ty = (tile / p.tiles_x) % p.tiles_y;
    const int64_t k = tile / (p.tiles_x * p.tiles_y);
    const int64_t i0 = tx * kTx, j0 = ty * kTy;
    const double* xc = p.coord[0];
    const double* yc = p.coord[1];
    const int64_t sx = TWO_D ? 1 : 3, sy = sx * p.nx, sz = sy * p.ny;
    const double wz = TWO_D ? 1.0 : p.coord[2][k + 1] - p.coord[2][k];

    for (int t = threadIdx.x; t < kHx * kHy; t += kPostBlock) {
        const int lx = t % kHx, ly = t / kHx;
        const int64_t i = i0 - 1 + lx, j = j0 - 1 + ly;
        double T[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (i >= 0 && i < mx && j >= 0 && j < my) {
            const int64_t cell = i + mx * (j + my * k);
            const double* a0 = p.a + (i * sx + j * sy + k * sz);
            const double wx = xc[i + 1] - xc[i], wy = yc[j + 1] - yc[j];
            double B[3];
            if constexpr (TWO_D) {
                const double mx_ = -a0[0] + a0[sy];  // moment of B_x, the stencil's order
                const double my_ = a0[0] - a0[sx];   // moment of B_y
                const double rx = fabs(wx) < 1e-14 ? 0.0 : 1.0 / wx;
                const double ry = fabs(wy) < 1e-14 ? 0.0 : 1.0 / wy;
                B[0] = ry * mx_;
                B[1] = rx * my_;
                B[2] = 0.0;
            } else {
                // forward_value moments, the taps in the stencil's order
                double m0 = a0[sy + 2];
                m0 = m0 - a0[2];
                m0 = m0 - a0[sz + 1];
                m0 = m0 + a0[1];
                double m1 = -a0[sx + 2];
                m1 = m1 + a0[2];
                m1 = m1 + a0[sz + 0];
                m1 = m1 - a0[0];
                double m2 = a0[sx + 1];
                m2 = m2 - a0[1];
                m2 = m2 - a0[sy + 0];
                m2 = m2 + a0[0];
                B[0] = m0 / (wy * wz);
                B[1] = m1 / (wx * wz);
                B[2] = m2 / (wx * wy);
            }
            double H[3] = {0.0, 0.0, 0.0};
            if (p.nonlinear && p.nonlinear[cell] != 0.0) {
                const double q = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
                const double bn = sqrt(q);
                if (bn != 0.0) {
                    const double scale = h_from_b(*p.curve, bn) / bn;
                    for (int m = 0; m < 3; ++m) H[m] = scale * B[m];
                }
            } else {
                const double mu = p.mu[cell];
                for (int m = 0; m < 3; ++m) H[m] = B[m] / mu;
            }
            const double half = 0.5 * (B[0] * H[0] + B[1] * H[1] + B[2] * H[2]);
            T[0] = B[0] * H[0] - half;
            T[1] = B[1] * H[1] - half;
            T[2] = B[2] * H[2] - half;
            T[3] = B[0] * H[1];
            T[4] = B[0] * H[2];
            T[5] = B[1] * H[2];
            if (lx > 0 && ly > 0) {
                if constexpr (STAGED) {
                    const int o = (ly - 1) * kTx + (lx - 1);
                    for (int m = 0; m < 3; ++m) {
                        s_o[m][o] = B[m];
                        s_o[3 + m][o] = H[m];
                    }
                } else {
                    if (p.induction)
                        for (int m = 0; m < 3; ++m) p.induction[3 * cell + m] = B[m];
                    if (p.field)
                        for (int m = 0; m < 3; ++m) p.field[3 * cell + m] = H[m];
                    if (p.stress)
                        for (int m = 0; m < 6; ++m) p.stress[6 * cell + m] = T[m];
                }
            }
        }
        for (int m = 0; m < 6; ++m) s_t[m][t] = T[m];
    }
    const int64_t nv = mx - i0 < kTx ? mx - i0 : kTx;  // valid cells of each tile row
    if constexpr (STAGED) {
        __syncthreads();
        if (p.induction || p.field)
            for (int e = threadIdx.x; e < kTy * 3 * kTx; e += kPostBlock) {
                const int r = e / (3 * kTx), q = e % (3 * kTx), cl = q / 3, m = q % 3;
                const int64_t j = j0 + r;
                if (j >= my || cl >= nv) continue;
                const int64_t base = 3 * (i0 + mx * (j + my * k));
                if (p.induction) __builtin_nontemporal_store(s_o[m][r * kTx + cl], p.induction + base + q);
                if (p.field) __builtin_nontemporal_store(s_o[3 + m][r * kTx + cl], p.field + base + q);
            }
        if (p.stress)
            for (int e = threadIdx.x; e < kTy * 6 * kTx; e += kPostBlock) {
                const int r = e / (6 * kTx), q = e % (6 * kTx), cl = q / 6, m = q % 6;
                const int64_t j = j0 + r;
                if (j >= my || cl >= nv) continue;
                __builtin_nontemporal_store(s_t[m][(r + 1) * kHx + cl + 1], p.stress + 6 * (i0 + mx * (j + my * k)) + q);
            }
        if (!p.force) return;
    } else {
        if (!p.force) return;
        __syncthreads();
    }

    double fo[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    int slot = 0;

    for (int t = threadIdx.x; t < kTx * kTy; t += kPostBlock, ++slot) {
        const int lx = 1 + t % kTx, ly = 1 + t / kTx;
        const int64_t i = i0 - 1 + lx, j = j0 - 1 + ly;
        if (i >= mx || j >= my) continue;
        const int c = ly * kHx + lx;
        const int cw = c - 1, cs = c - kHx;  // (i−1, j) and (i, j−1)
        const double hx = centre_spacing(xc, mx, i), hy = centre_spacing(yc, my, j);
        const double hxw = i > 0 ? centre_spacing(xc, mx, i - 1) : 0.0;
        const double hys = j > 0 ? centre_spacing(yc, my, j - 1) : 0.0;
        // ⋆ coefficients of the one-form at a centre (zero when degenerate)
        auto ok1 = [](double u, double v) { return fabs(u) >= 1e-14 && fabs(v) >= 1e-14; };
        const bool o = ok1(hx, hy), ow = ok1(hxw, hy), os = ok1(hx, hys);
        const double cxy = o ? fabs(hy) / fabs(hx) : 0.0, cyx = o ? -fabs(hx) / fabs(hy) : 0.0;
        const double cxy_w = ow ? fabs(hy) / fabs(hxw) : 0.0;
        const double cyx_s = os ? -fabs(hx) / fabs(hys) : 0.0;
        const double c2 = fabs(hx * hy) >= 1e-14 ? 1.0 / (fabs(hx) * fabs(hy)) : 0.0;
        // one-forms (ω_x, ω_y): (T_xx, T_xy), (T_xy, T_yy), (T_xz, T_yz)
        constexpr int kPick[3][2] = {{0, 3}, {3, 1}, {4, 5}};
        const int64_t cell = i + mx * (j + my * k);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const double eta_y = cxy * s_t[kPick[f][0]][c];
            const double eta_x = cyx * s_t[kPick[f][1]][c];
            const double eta_y_w = cxy_w * s_t[kPick[f][0]][cw];  // zero outside: T = 0 there
            const double eta_x_s = cyx_s * s_t[kPick[f][1]][cs];
            double acc = -eta_y;
            acc = acc + eta_x;
            acc = acc + eta_y_w;
            acc = acc - eta_x_s;
            if constexpr (STAGED)
                fo[slot][f] = c2 * -acc;
            else
                p.force[3 * cell + f] = c2 * -acc;
        }
    }
    if constexpr (STAGED) {
        __syncthreads();  // every read of s_o and s_t is done
        slot = 0;
        for (int t = threadIdx.x; t < kTx * kTy; t += kPostBlock, ++slot)
            for (int f = 0; f < 3; ++f) s_o[f][t] = fo[slot][f];
        __syncthreads();
        for (int e = threadIdx.x; e < kTy * 3 * kTx; e += kPostBlock) {
            const int r = e / (3 * kTx), q = e % (3 * kTx), cl = q / 3, m = q % 3;
            const int64_t j = j0 + r;
            if (j >= my || cl >= nv) continue;
            __builtin_nontemporal_store(s_o[m][r * kTx + cl], p.force + 3 * (i0 + mx * (j + my * k)) + q);
        }
    }
}

int launch_post(PostArgs p, bool two_d, hipStream_t s)
{
    const int64_t mx = p.nx - 1, my = p.ny - 1, mz = p.nz - 1;
    p.tiles_x = (mx + kTx - 1) / kTx;
    p.tiles_y = (my + kTy - 1) / kTy;
    p.tiles = p.tiles_x * p.tiles_y * mz;
    p.per_xcd = (p.tiles + kXcdCount - 1) / kXcdCount;
    const int64_t blocks = p.per_xcd * kXcdCount;
    SIL_REQUIRE(blocks <= INT32_MAX, SIL_ERR_UNSUPPORTED, "magnetostatics post_process: grid too large");
    static const bool staged = [] {
        const char* e = std::getenv("SIL_POST_STAGED");
        return !(e && e[0] == '0');
    }();
    const dim3 g((unsigned)blocks), b(kPostBlock);
    if (two_d && staged)
        hipLaunchKernelGGL((k_ms3d_post<true, true>), g, b, 0, s, p);
    else if (two_d)
        hipLaunchKernelGGL((k_ms3d_post<true, false>), g, b, 0, s, p);
    else if (staged)
        hipLaunchKernelGGL((k_ms3d_post<false, true>), g, b, 0, s, p);
    else
        hipLaunchKernelGGL((k_ms3d_post<false, false>), g, b, 0, s, p);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

}  // namespace

int ms2d_post_process(const Ms2dData& d, const BhCurve* curve, const double* potential, const double* cell_mu,
                      const double* cell_nonlinear, double* induction, double* field, double* stress,
                      double* force, hipStream_t s)
{
    PostArgs p{};
    p.nx = d.nx;
    p.ny = d.ny;
    p.nz = 2;  // one slice of cells
    p.coord[0] = d.coord[0];
    p.coord[1] = d.coord[1];
    p.a = potential;
    p.mu = cell_mu;
    p.nonlinear = cell_nonlinear;
    p.curve = curve;
    p.induction = induction;
    p.field = field;
    p.stress = stress;
    p.force = force;
    return launch_post(p, true, s);
}

int ms3d_post_process(const Ms3dData& d, const BhCurve* curve, const double* potential, const double* cell_mu,
                      const double* cell_nonlinear, double* induction, double* field, double* stress,
                      double* force, hipStream_t s)
{
    PostArgs p{};
    p.nx = d.nx;
    p.ny = d.ny;
    p.nz = d.nz;
    for (int a = 0; a < 3; ++a) p.coord[a] = d.coord[a];
    p.a = potential;
    p.mu = cell_mu;
    p.nonlinear = cell_nonlinear;
    p.curve = curve;
    p.induction = induction;
    p.field = field;
    p.stress = stress;
    p.force = force;
    return launch_post(p, false, s);
}

}  // namespace sil
