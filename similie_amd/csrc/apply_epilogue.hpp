// Device helpers shared by the matrix-free operator applications: the CG
// direction formed on the fly and the per-block partial sums of ⟨in, out⟩.
#pragma once

#include <hipcub/hipcub.hpp>

#include "operators.hpp"

namespace sil {

constexpr int kApplyBlock = 256;

template <typename T>
__device__ __forceinline__ T p_value(const PUpdate<T>* pu, const T* x, int64_t idx)
{
    // value of the direction at row idx: either x[idx] or z + β p_old
    if (pu == nullptr) return x[idx];
    const double beta = *pu->beta;
    const double z = pu->inv_diag ? (double)pu->r[idx] * (double)pu->inv_diag[idx] : (double)pu->r[idx];
    return (T)(z + beta * (double)pu->p_old[idx]);
}

// one partial per workgroup of kApplyBlock threads
template <typename T>
__device__ __forceinline__ void block_dot_store(double v, double* partials)
{
    using BR = hipcub::BlockReduce<double, kApplyBlock>;
    __shared__ typename BR::TempStorage tmp;
    double s = BR(tmp).Sum(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

}  // namespace sil
