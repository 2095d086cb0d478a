// Lane-layout check of the matrix-core instructions (one wavefront,
// in-bounds loads from host-provided A (M×K) and B (K×N) row-major FP32
// matrices, rounded to BF16 for the BF16 instructions; D (M×N) row-major
// FP32).  Built by
// `python -m similie_amd.build` into tests/cpp/libmfma_probe.so.
#include <hip/hip_runtime.h>

#include "mfma_layout.hpp"

using namespace sil;

template <class L, int M, int N, int K, class Acc, int R>
__global__ void k_probe(const float* A, const float* B, float* D)
{
    const int lane = threadIdx.x;
    bf16x8 a, b;
    for (int i = 0; i < 8; ++i) {
        a[i] = to_bf16(A[L::a_row(lane) * K + L::a_k0(lane) + i]);
        b[i] = to_bf16(B[(L::b_k0(lane) + i) * N + L::b_col(lane)]);
    }
    Acc acc;
    for (int r = 0; r < R; ++r) acc[r] = 0.f;
    acc = L::mma(a, b, acc);
    for (int r = 0; r < R; ++r) D[L::d_row(lane, r) * N + L::d_col(lane)] = acc[r];
}

extern "C" int probe_mfma32(const float* A, const float* B, float* D)
{
    hipLaunchKernelGGL((k_probe<Mfma32, 32, 32, 16, f32x16, 16>), dim3(1), dim3(64), 0, 0, A, B, D);
    return hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}

extern "C" int probe_mfma16(const float* A, const float* B, float* D)
{
    hipLaunchKernelGGL((k_probe<Mfma16, 16, 16, 32, f32x4, 4>), dim3(1), dim3(64), 0, 0, A, B, D);
    return hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}
