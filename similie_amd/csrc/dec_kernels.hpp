// Internal launchers of the DEC kernels (host side).
#pragma once

#include "common.hpp"

namespace sil {

int launch_coboundary(int k, bool transposed, const double* in, const sil_view& vin, double* out,
                      const sil_view& vout, hipStream_t stream);
int launch_fill_hodge(const DevGeom& g, int k, double* out, hipStream_t stream);
int launch_hodge_apply(const DevGeom& g, int k, const double* in, const sil_view& vin, double* out,
                       const sil_view& vout, hipStream_t stream);
int launch_codifferential(const DevGeom& g, int k, const double* in, const sil_view& vin, double* out,
                          const sil_view& vout, hipStream_t stream);
int launch_laplacian_generic(const DevGeom& g, int k, const double* in, const sil_view& vin, double* out,
                             const sil_view& vout, hipStream_t stream);

// Fused single-pass Laplacians on rectilinear grids with identity or diagonal
// metric, index-last dense arrays.  `spacing[a]` holds the per-node edge
// length along axis a (forward, backward at the last node).
struct FusedGeom {
    int64_t ext[3];
    const double* spacing[3];
    const double* metric; // diagonal [node][ndim] or null for the identity
};

// FUSED_MARCH: x-march, guarded loads, one plane prefetched;
// FUSED_MARCH_DEEP / _CLAMPED: clamped loads, two / one planes prefetched
enum FusedVariant { FUSED_DIRECT = 0, FUSED_MARCH = 1, FUSED_MARCH_DEEP = 2, FUSED_MARCH_CLAMPED = 3, FUSED_MARCH_V0 = 4, FUSED_MARCH_V5 = 5,
                   FUSED_MARCH_V0_OCC8 = 6, FUSED_MARCH_OCC6 = 7 };

struct FusedTile {
    int variant;  // FusedVariant
    int tz, ty;   // tile in (z, y)
    int xchunk;   // x planes per workgroup (march)
};

int launch_laplacian_2d_0form(const FusedGeom& g, const double* in, double* out, int64_t ox, int64_t oy,
                              hipStream_t stream);
int launch_laplacian_3d_1form(const FusedGeom& g, const double* in, double* out, const int64_t* oext,
                              const FusedTile& tile, hipStream_t stream);
int launch_coboundary_3d_0form(const double* in, double* out, const int64_t* ext, hipStream_t stream);
int launch_stream_copy(const double* in, double* out, int64_t n, hipStream_t stream);

}  // namespace sil
