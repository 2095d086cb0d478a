// Raw buffer resources for branch-free masked memory access on CDNA.
//
// A resource covers [base, base + bytes) (stride 0, gfx9 data format word):
// a load at an offset at or beyond its byte range returns zero and a store
// there is dropped, so lanes and planes outside a grid are masked without a
// branch between a load and its consumer — the compiler's wait counts then
// follow the data and several planes stay in flight.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sil {

typedef unsigned int U32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int U32x3 __attribute__((ext_vector_type(3)));
typedef unsigned int U32x4 __attribute__((ext_vector_type(4)));

// an offset no resource range reaches
constexpr uint32_t kOutside = 0x80000000u;
constexpr int64_t kMaxBufferBytes = 0x7FFFFFFFLL;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* base, int64_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000);
}

__device__ __forceinline__ double as_double(unsigned lo, unsigned hi)
{
    return __builtin_bit_cast(double, static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32));
}

// Lanes of a loaded vector are copied into scalars before they are
// reinterpreted: this compiler folds __builtin_bit_cast(float, v[i]) on an
// element of an ext_vector into element 0 (and then loads one dword only).
__device__ __forceinline__ float f32_of(unsigned u) { return __builtin_bit_cast(float, u); }

// one value of T at byte offset off
template <typename T, int AUX = 0>
__device__ __forceinline__ T buffer_load1(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    if constexpr (sizeof(T) == 8) {
        const U32x2 a = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX);
        return as_double(a[0], a[1]);
    } else {
        return f32_of(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, AUX));
    }
}

// three consecutive values of T at byte offset off (FP64: a 16-byte and an
// 8-byte request; FP32: one 12-byte request)
template <typename T, int AUX = 0>
__device__ __forceinline__ void buffer_load3(__amdgpu_buffer_rsrc_t r, uint32_t off, T v[3])
{
    if constexpr (sizeof(T) == 8) {
        const U32x4 a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX);
        const U32x2 b = __builtin_amdgcn_raw_buffer_load_b64(r, off + 16, 0, AUX);
        v[0] = as_double(a[0], a[1]);
        v[1] = as_double(a[2], a[3]);
        v[2] = as_double(b[0], b[1]);
    } else {
        const U32x3 a = __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, AUX);
        const unsigned a0 = a.x, a1 = a.y, a2 = a.z;
        v[0] = f32_of(a0);
        v[1] = f32_of(a1);
        v[2] = f32_of(a2);
    }
}

// four consecutive values of T (a 16-byte aligned record)
template <typename T, int AUX = 0>
__device__ __forceinline__ void buffer_load4(__amdgpu_buffer_rsrc_t r, uint32_t off, T v[4])
{
    if constexpr (sizeof(T) == 8) {
        const U32x4 a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX);
        const U32x4 b = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, 0, AUX);
        v[0] = as_double(a[0], a[1]);
        v[1] = as_double(a[2], a[3]);
        v[2] = as_double(b[0], b[1]);
        v[3] = as_double(b[2], b[3]);
    } else {
        const U32x4 a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX);
        const unsigned a0 = a.x, a1 = a.y, a2 = a.z, a3 = a.w;
        v[0] = f32_of(a0);
        v[1] = f32_of(a1);
        v[2] = f32_of(a2);
        v[3] = f32_of(a3);
    }
}

// three consecutive values of T stored at byte offset off
template <typename T, int AUX = 0>
__device__ __forceinline__ void buffer_store3(__amdgpu_buffer_rsrc_t r, uint32_t off, const T v[3])
{
    if constexpr (sizeof(T) == 8) {
        const uint64_t ua = __builtin_bit_cast(uint64_t, v[0]), ub = __builtin_bit_cast(uint64_t, v[1]),
                       uc = __builtin_bit_cast(uint64_t, v[2]);
        const U32x4 lo = {static_cast<unsigned>(ua), static_cast<unsigned>(ua >> 32), static_cast<unsigned>(ub),
                          static_cast<unsigned>(ub >> 32)};
        const U32x2 hi = {static_cast<unsigned>(uc), static_cast<unsigned>(uc >> 32)};
        __builtin_amdgcn_raw_buffer_store_b128(lo, r, off, 0, AUX);
        __builtin_amdgcn_raw_buffer_store_b64(hi, r, off + 16, 0, AUX);
    } else {
        const U32x3 a = {__builtin_bit_cast(unsigned, v[0]), __builtin_bit_cast(unsigned, v[1]),
                         __builtin_bit_cast(unsigned, v[2])};
        __builtin_amdgcn_raw_buffer_store_b96(a, r, off, 0, AUX);
    }
}

}  // namespace sil
