// Air-gap diagnostics of a magnetostatics solution
// (onelab_interface/magnetostatics_onelab.hpp:4666-4720 in 2D, 5246-5307 in
// 3D, 460-519 for the traction and the mirror projection).
//
// From the cell fields of the post-processing pass (B and H per cell) and a
// per-cell mask (bit 0: the diagnostic region, bit 1: a positively oriented
// conductor):
//   * an upper air-gap face is the +y face of a region cell whose +y
//     neighbour is not in the region (or that has none); its measure is
//     dx·dz (3D) or dx·length_z (2D);
//   * flux Φ = Σ measure·B_y, traction T·n = (B_x H_y, B_y H_y − ½ B·H,
//     B_z H_y) on that face, its integral and the integral of its magnitude;
//   * current I = Σ dx·dy·J_z over conductor cells (3D: the cell layer
//     k = ncz/2 only);
//   * when the faces are mirror-symmetric in x (pairs at ±x with one z and
//     one measure, within 1e-12), the traction integral is replaced by the
//     pairwise sum of its y and z parts (x cancels).
// The device flags the faces and compacts them (stable, so in the
// reference's loop order), computes each face's terms, and the host sums them
// in that order and applies the projection.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "operators.hpp"

namespace sil {
namespace {

constexpr int kDiagBlock = 256;

struct FaceGrid {
    int64_t ncx, ncy, ncz;  // cells per axis (ncz = 1 in 2D)
    const double* x;        // node coordinates (device)
    const double* y;
    const double* z;        // null in 2D
    double length_z;        // 2D: the extrusion length
    double z_value;         // 2D: z of every face
};

__global__ void k_face_flags(FaceGrid g, const uint8_t* __restrict__ mask, uint8_t* __restrict__ face)
{
    const int64_t nc = g.ncx * g.ncy * g.ncz;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = (c / g.ncx) % g.ncy;
        const bool in = (mask[c] & 1u) != 0;
        const bool upper_in = j + 1 < g.ncy && (mask[c + g.ncx] & 1u) != 0;
        face[c] = (in && !upper_in) ? 1 : 0;
    }
}

// per face: measure, Φ term, traction·measure (3), |traction|·measure,
// cell-centre x and z, traction y and z
constexpr int kFaceRecord = 10;

__global__ void k_face_terms(FaceGrid g, const int64_t* __restrict__ cells, const int64_t* __restrict__ count,
                             const double* __restrict__ cell_b, const double* __restrict__ cell_h,
                             double* __restrict__ rec)
{
    const int64_t nf = *count;
    for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nf; f += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = cells[f];
        const int64_t i = c % g.ncx, k = c / (g.ncx * g.ncy);
        const double dx = g.x[i + 1] - g.x[i];
        const double measure = g.z ? dx * (g.z[k + 1] - g.z[k]) : dx * g.length_z;
        const double b0 = cell_b[3 * c], b1 = cell_b[3 * c + 1], b2 = cell_b[3 * c + 2];
        const double h0 = cell_h[3 * c], h1 = cell_h[3 * c + 1], h2 = cell_h[3 * c + 2];
        const double half_trace = 0.5 * (b0 * h0 + b1 * h1 + b2 * h2);
        const double t0 = b0 * h1, t1 = b1 * h1 - half_trace, t2 = b2 * h1;
        double* r = rec + kFaceRecord * f;
        r[0] = measure;
        r[1] = measure * b1;
        r[2] = measure * t0;
        r[3] = measure * t1;
        r[4] = measure * t2;
        r[5] = measure * std::sqrt(t0 * t0 + t1 * t1 + t2 * t2);
        r[6] = 0.5 * (g.x[i] + g.x[i + 1]);
        r[7] = g.z ? 0.5 * (g.z[k] + g.z[k + 1]) : g.z_value;
        r[8] = t1;
        r[9] = t2;
    }
}

// conductor cells of a layer: area·J_z (0 elsewhere), and a 0/1 count
__global__ void k_current_terms(FaceGrid g, int64_t layer, const uint8_t* __restrict__ mask,
                                const double* __restrict__ jz, int64_t jz_stride, double* __restrict__ term,
                                uint8_t* __restrict__ flag)
{
    const int64_t n2 = g.ncx * g.ncy;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n2; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = layer * n2 + q;
        const int64_t i = q % g.ncx, j = q / g.ncx;
        const bool on = (mask[c] & 2u) != 0;
        const double area = (g.x[i + 1] - g.x[i]) * (g.y[j + 1] - g.y[j]);
        term[q] = on ? area * jz[jz_stride * c] : 0.0;
        flag[q] = on ? 1 : 0;
    }
}

unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kDiagBlock - 1) / kDiagBlock, 16384)); }

struct FaceSample {
    double x, z, measure, t1, t2;
};

// apply_x_mirror_symmetry_projection_to_traction_integral (:475-519)
void mirror_projection(std::vector<FaceSample> samples, sil_air_gap_result& r)
{
    if (samples.empty()) return;
    constexpr double tol = 1.0e-12;
    std::stable_sort(samples.begin(), samples.end(), [](const FaceSample& a, const FaceSample& b) {
        if (std::abs(a.z - b.z) > tol) return a.z < b.z;
        return a.x < b.x;
    });
    double proj[3] = {0.0, 0.0, 0.0};
    size_t left = 0, right = samples.size();
    while (left < right) {
        --right;
        const FaceSample& lf = samples[left];
        const FaceSample& rf = samples[right];
        if (left == right) {
            if (std::abs(lf.x) > tol) return;
            proj[1] += lf.measure * lf.t1;
            proj[2] += lf.measure * lf.t2;
            break;
        }
        if (std::abs(lf.z - rf.z) > tol || std::abs(lf.x + rf.x) > tol || std::abs(lf.measure - rf.measure) > tol)
            return;
        proj[1] += lf.measure * (lf.t1 + rf.t1);
        proj[2] += lf.measure * (lf.t2 + rf.t2);
        ++left;
    }
    for (int a = 0; a < 3; ++a) r.traction[a] = proj[a];
}

int diagnostics(const FaceGrid& g, const double* cell_b, const double* cell_h, const uint8_t* mask,
                const double* jz, int64_t jz_stride, int64_t layer, sil_air_gap_result& out, hipStream_t s)
{
    const int64_t nc = g.ncx * g.ncy * g.ncz, n2 = g.ncx * g.ncy;
    out = sil_air_gap_result{};
    uint8_t* face = nullptr;
    int64_t *cells = nullptr, *count = nullptr;
    double *rec = nullptr, *term = nullptr;
    uint8_t* cflag = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    int rc = SIL_OK;
    auto fail = [&](const char* what) {
        set_error("air_gap_diagnostics: %s", what);
        rc = SIL_ERR_HIP;
    };
    if (hipMalloc(&face, std::max<int64_t>(nc, 1)) != hipSuccess ||
        hipMalloc(&cells, sizeof(int64_t) * std::max<int64_t>(nc, 1)) != hipSuccess ||
        hipMalloc(&count, sizeof(int64_t)) != hipSuccess)
        fail("allocation failed");
    hipcub::CountingInputIterator<int64_t> idx(0);
    if (rc == SIL_OK) {
        hipLaunchKernelGGL(k_face_flags, dim3(grid_of(nc)), dim3(kDiagBlock), 0, s, g, mask, face);
        if (hipcub::DeviceSelect::Flagged(nullptr, tmp_bytes, idx, face, cells, count, (int)nc, s) != hipSuccess ||
            hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)) != hipSuccess ||
            hipcub::DeviceSelect::Flagged(tmp, tmp_bytes, idx, face, cells, count, (int)nc, s) != hipSuccess)
            fail("face compaction failed");
    }
    int64_t nf = 0;
    if (rc == SIL_OK && (hipMemcpyAsync(&nf, count, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
                         hipStreamSynchronize(s) != hipSuccess))
        fail("reading the face count failed");
    std::vector<double> host_rec;
    if (rc == SIL_OK && nf > 0) {
        if (hipMalloc(&rec, sizeof(double) * kFaceRecord * nf) != hipSuccess) {
            fail("allocation failed");
        } else {
            hipLaunchKernelGGL(k_face_terms, dim3(grid_of(nf)), dim3(kDiagBlock), 0, s, g, cells, count, cell_b,
                               cell_h, rec);
            host_rec.resize((size_t)(kFaceRecord * nf));
            if (hipMemcpyAsync(host_rec.data(), rec, sizeof(double) * kFaceRecord * nf, hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                fail("reading the face terms failed");
        }
    }
    std::vector<double> host_term;
    std::vector<uint8_t> host_flag;
    if (rc == SIL_OK && n2 > 0) {
        if (hipMalloc(&term, sizeof(double) * n2) != hipSuccess || hipMalloc(&cflag, n2) != hipSuccess) {
            fail("allocation failed");
        } else {
            hipLaunchKernelGGL(k_current_terms, dim3(grid_of(n2)), dim3(kDiagBlock), 0, s, g, layer, mask, jz,
                               jz_stride, term, cflag);
            host_term.resize((size_t)n2);
            host_flag.resize((size_t)n2);
            if (hipMemcpyAsync(host_term.data(), term, sizeof(double) * n2, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipMemcpyAsync(host_flag.data(), cflag, n2, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                fail("reading the current terms failed");
        }
    }
    for (void* p : {(void*)face, (void*)cells, (void*)count, (void*)rec, (void*)term, (void*)cflag, tmp})
        if (p) (void)hipFree(p);
    if (rc) return rc;
    // sums in the reference's loop order (cell index ascending)
    for (int64_t q = 0; q < n2; ++q)
        if (host_flag[q]) {
            out.current += host_term[q];
            ++out.current_cells;
        }
    std::vector<FaceSample> samples;
    samples.reserve((size_t)nf);
    for (int64_t f = 0; f < nf; ++f) {
        const double* r = &host_rec[(size_t)(kFaceRecord * f)];
        out.flux += r[1];
        out.surface_measure += r[0];
        out.traction[0] += r[2];
        out.traction[1] += r[3];
        out.traction[2] += r[4];
        out.traction_magnitude += r[5];
        samples.push_back(FaceSample{r[6], r[7], r[0], r[8], r[9]});
        ++out.faces;
    }
    mirror_projection(std::move(samples), out);
    return SIL_OK;
}

}  // namespace
}  // namespace sil

using namespace sil;

extern "C" {

int sil_magnetostatics3d_air_gap_diagnostics(sil_operator op, const double* cell_b, const double* cell_h,
                                             const uint8_t* cell_mask, const double* cell_current,
                                             sil_air_gap_result* out, void* stream)
{
    SIL_REQUIRE(op && cell_b && cell_h && cell_mask && cell_current && out, SIL_ERR_INVALID,
                "magnetostatics3d_air_gap_diagnostics: null argument");
    SIL_REQUIRE(op->kind == OP_MAGNETOSTATICS3D || op->kind == OP_MS3D_NONLINEAR, SIL_ERR_INVALID,
                "magnetostatics3d_air_gap_diagnostics: not a 3D magnetostatics operator");
    const Ms3dData& d = op->ms;
    SIL_REQUIRE(d.nx >= 2 && d.ny >= 2 && d.nz >= 2, SIL_ERR_INVALID,
                "magnetostatics3d_air_gap_diagnostics: the grid has no cells");
    FaceGrid g{d.nx - 1, d.ny - 1, d.nz - 1, d.coord[0], d.coord[1], d.coord[2], 0.0, 0.0};
    return diagnostics(g, cell_b, cell_h, cell_mask, cell_current + 2, 3, g.ncz / 2, *out,
                       reinterpret_cast<hipStream_t>(stream));
}

int sil_magnetostatics2d_air_gap_diagnostics(sil_operator op, const double* cell_b, const double* cell_h,
                                             const uint8_t* cell_mask, const double* cell_current_z,
                                             double length_z, double z_value, sil_air_gap_result* out,
                                             void* stream)
{
    SIL_REQUIRE(op && cell_b && cell_h && cell_mask && cell_current_z && out, SIL_ERR_INVALID,
                "magnetostatics2d_air_gap_diagnostics: null argument");
    SIL_REQUIRE(op->kind == OP_MAGNETOSTATICS2D || op->kind == OP_MS2D_NONLINEAR, SIL_ERR_INVALID,
                "magnetostatics2d_air_gap_diagnostics: not a 2D magnetostatics operator");
    const Ms2dData& d = op->m2;
    FaceGrid g{d.nx - 1, d.ny - 1, 1, d.coord[0], d.coord[1], nullptr, length_z, z_value};
    return diagnostics(g, cell_b, cell_h, cell_mask, cell_current_z, 1, 0, *out,
                       reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
