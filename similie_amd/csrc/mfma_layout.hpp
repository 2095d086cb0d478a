// Lane layouts of the gfx950 matrix-core instructions used by the
// learned preconditioner's convolutions (FP32 accumulate).
//
// v_mfma_f32_32x32x16_bf16  D(32×32) += A(32×16) · B(16×32)
//   A: lane l holds row m = l % 32, columns k = 8·(l / 32) + i, i = 0..7
//   B: lane l holds column n = l % 32, rows k = 8·(l / 32) + i
//   D: lane l holds column n = l % 32, rows m = 8·g + 4·(l / 32) + j in
//      acc[4·g + j], g, j = 0..3
// v_mfma_f32_16x16x32_bf16  D(16×16) += A(16×32) · B(32×16)
//   A: lane l holds row m = l % 16, columns k = 8·(l / 16) + i
//   B: lane l holds column n = l % 16, rows k = 8·(l / 16) + i
//   D: lane l holds column n = l % 16, rows m = 4·(l / 16) + j in acc[j]
// These mappings are checked against a host matrix product by
// tests/test_gpu_mfma_layout.py before any kernel relies on them.
#pragma once

#include <hip/hip_runtime.h>

namespace sil {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct Mfma32 {
    __device__ static int a_row(int lane) { return lane & 31; }
    __device__ static int a_k0(int lane) { return 8 * (lane >> 5); }
    __device__ static int b_col(int lane) { return lane & 31; }
    __device__ static int b_k0(int lane) { return 8 * (lane >> 5); }
    __device__ static int d_col(int lane) { return lane & 31; }
    __device__ static int d_row(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
    __device__ static f32x16 mma(bf16x8 a, bf16x8 b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

struct Mfma16 {
    __device__ static int a_row(int lane) { return lane & 15; }
    __device__ static int a_k0(int lane) { return 8 * (lane >> 4); }
    __device__ static int b_col(int lane) { return lane & 15; }
    __device__ static int b_k0(int lane) { return 8 * (lane >> 4); }
    __device__ static int d_col(int lane) { return lane & 15; }
    __device__ static int d_row(int lane, int r) { return 4 * (lane >> 4) + r; }
    __device__ static f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ __bf16 to_bf16(float v) { return static_cast<__bf16>(v); }
__device__ __forceinline__ float from_bf16(__bf16 v) { return static_cast<float>(v); }

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
// a and b rounded to BF16 as to_bf16 does, packed with a in the low half: one
// v_cvt_pk_bf16_f32 (two scalar conversions compile to two of them plus a
// shift and an OR)
__device__ __forceinline__ unsigned pack_bf16x2(float a, float b)
{
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{a, b}, bf16x2));
}
typedef short i16x2 __attribute__((ext_vector_type(2)));
// ReLU of a packed BF16 pair: a value with its sign bit set (a negative
// number, −0 or a negative NaN) becomes +0, as a signed 16-bit max with 0
// (one v_pk_max_i16).  ReLU and rounding to BF16 commute, so this equals
// the rounded max(v, 0) of the FP32 value except that a positive NaN
// stays NaN.
__device__ __forceinline__ unsigned relu_bf16x2(unsigned w)
{
    return __builtin_bit_cast(unsigned, __builtin_elementwise_max(__builtin_bit_cast(i16x2, w), i16x2{0, 0}));
}
// the BF16 values in the low and high half of a packed word, as FP32
__device__ __forceinline__ float bf16_lo(unsigned u) { return __builtin_bit_cast(float, u << 16); }
__device__ __forceinline__ float bf16_hi(unsigned u) { return __builtin_bit_cast(float, u & 0xFFFF0000u); }

}  // namespace sil
