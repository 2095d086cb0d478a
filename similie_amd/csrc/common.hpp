// Shared host/device helpers of libsil_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "sil_hip.h"

namespace sil {

// ---------------------------------------------------------------- errors ---

void set_error(const char* fmt, ...);

#define SIL_CHECK_HIP(expr)                                                             \
    do {                                                                                \
        hipError_t sil_e_ = (expr);                                                     \
        if (sil_e_ != hipSuccess) {                                                     \
            ::sil::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(sil_e_), \
                             __FILE__, __LINE__);                                       \
            return SIL_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

#define SIL_REQUIRE(cond, code, ...)       \
    do {                                   \
        if (!(cond)) {                     \
            ::sil::set_error(__VA_ARGS__); \
            return (code);                 \
        }                                  \
    } while (0)

#define SIL_LAUNCH_CHECK()                                                                 \
    do {                                                                                   \
        hipError_t sil_e_ = hipGetLastError();                                             \
        if (sil_e_ != hipSuccess) {                                                        \
            ::sil::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(sil_e_), \
                             __FILE__, __LINE__);                                          \
            return SIL_ERR_HIP;                                                            \
        }                                                                                  \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// --------------------------------------------------------- combinatorics ---
//
// k-subsets of {0..n-1} as bitmasks, listed in lexicographic order of their
// sorted index lists: the antisymmetric mem_id order and the tangent-basis
// order of the reference (the sequence std::prev_permutation walks from
// [1,..,1,0,..,0]).  Tables cover n ≤ 4 and live in constant memory.

struct ComboTables {
    int8_t count[5][5];    // C(n, k)
    uint8_t mask[5][5][6]; // i-th k-subset of n axes
    int8_t index[5][16];   // position of a mask in its (n, popcount) list
};

constexpr ComboTables build_combo_tables()
{
    ComboTables t{};
    for (int n = 0; n <= 4; ++n) {
        for (int k = 0; k <= n; ++k) {
            int cnt = 0;
            // enumerate sorted index lists lexicographically
            int ids[4] = {0, 1, 2, 3};
            if (k == 0) {
                t.mask[n][0][0] = 0;
                t.count[n][0] = 1;
                t.index[n][0] = 0;
                continue;
            }
            for (;;) {
                unsigned m = 0;
                for (int i = 0; i < k; ++i) m |= 1u << ids[i];
                t.mask[n][k][cnt] = static_cast<uint8_t>(m);
                t.index[n][m] = static_cast<int8_t>(cnt);
                ++cnt;
                int i = k - 1;
                while (i >= 0 && ids[i] == n - k + i) --i;
                if (i < 0) break;
                ++ids[i];
                for (int j = i + 1; j < k; ++j) ids[j] = ids[j - 1] + 1;
            }
            t.count[n][k] = static_cast<int8_t>(cnt);
        }
    }
    return t;
}

static constexpr ComboTables kCombos = build_combo_tables();
static __constant__ ComboTables c_combos = build_combo_tables();

__host__ __device__ inline int popcount4(unsigned m)
{
    return static_cast<int>((m & 1u) + ((m >> 1) & 1u) + ((m >> 2) & 1u) + ((m >> 3) & 1u));
}

// parity of the concatenated index list (src ‖ dst): +1 even, -1 odd
__host__ __device__ inline int concat_sign(unsigned src, unsigned dst, int n)
{
    int ids[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int m = 0;
    for (int a = 0; a < n; ++a)
        if (src & (1u << a)) ids[m++] = a;
    for (int a = 0; a < n; ++a)
        if (dst & (1u << a)) ids[m++] = a;
    int odd = 0;
    for (int i = 0; i < m; ++i)
        for (int j = i + 1; j < m; ++j)
            if (ids[i] > ids[j]) odd ^= 1;
    return odd ? -1 : 1;
}

__host__ __device__ inline int64_t binom4(int n, int k)
{
    if (k < 0 || k > n) return 0;
    int64_t r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

// ------------------------------------------------------------------ views ---

struct DevView {
    int ndim, nbatch;
    int64_t ext[SIL_MAX_DIM];
    int64_t str[SIL_MAX_DIM];
    int64_t cstr;
    int64_t bext[SIL_MAX_DIM];
    int64_t bstr[SIL_MAX_DIM];
};

inline DevView to_dev(const sil_view& v)
{
    DevView d{};
    d.ndim = v.ndim;
    d.nbatch = v.nbatch;
    for (int a = 0; a < SIL_MAX_DIM; ++a) {
        d.ext[a] = a < v.ndim ? v.extent[a] : 1;
        d.str[a] = a < v.ndim ? v.stride[a] : 0;
        d.bext[a] = a < v.nbatch ? v.batch_extent[a] : 1;
        d.bstr[a] = a < v.nbatch ? v.batch_stride[a] : 0;
    }
    d.cstr = v.comp_stride;
    return d;
}

int validate_view(const sil_view* v, const char* what);
int64_t view_nodes(const sil_view& v);
int64_t view_batches(const sil_view& v);
// index-last dense layout: component fastest, last axis next, first axis slowest
bool view_is_index_last_dense(const sil_view& v, int ncomp);

// --------------------------------------------------------------- geometry ---

struct DevGeom {
    int ndim;
    int rectilinear; // 1: axis coordinates, 0: nodal positions
    int64_t ext[SIL_MAX_DIM];
    const double* axis[SIL_MAX_DIM]; // device coordinates (rectilinear)
    const double* nodal;             // device [node][ndim] (general)
    int metric_kind;
    const double* metric;
};

}  // namespace sil

struct sil_grid_s {
    sil::DevGeom geom;
    double* axis_storage[SIL_MAX_DIM]; // device copies of the axis coordinates
    double* spacing[SIL_MAX_DIM];      // device per-node edge lengths (rectilinear)
};
