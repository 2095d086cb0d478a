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

// FUSED_DIRECT: one thread per output node, neighbours through the caches;
// FUSED_MARCH: x-march with guarded global loads, one plane prefetched;
// FUSED_PIPE: x-march with branch-free buffer accesses, two planes in flight
// FUSED_STAGE: FUSED_PIPE with z neighbours by DPP and staged contiguous stores
enum FusedVariant { FUSED_DIRECT = 0, FUSED_MARCH = 1, FUSED_PIPE = 2, FUSED_STAGE = 3 };

struct FusedTile {
    int variant;  // FusedVariant
    int tz, ty;   // tile in (z, y)
    int xchunk;   // x planes per workgroup (march)
    int nt = 0;   // pipelined 64 × 12 tile: non-temporal output stores
};

int launch_laplacian_2d_0form(const FusedGeom& g, const double* in, double* out, int64_t ox, int64_t oy,
                              hipStream_t stream);
int launch_laplacian_3d_1form(const FusedGeom& g, const double* in, double* out, const int64_t* oext,
                              const FusedTile& tile, hipStream_t stream);
int launch_coboundary_3d_0form(const double* in, double* out, const int64_t* ext, hipStream_t stream);
int launch_stream_copy(const double* in, double* out, int64_t n, hipStream_t stream);

}  // namespace sil
