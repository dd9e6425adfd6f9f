// Reductions of per-pair gradient rows: the reference's blocked weighted sum
// (batch_engine.hpp:697-718, 1788-1827), the same additions in the same order.
#pragma once

#include "common.hpp"

namespace nwg {

// partial[blk * gs + k] = sum over pairs i of block blk, in pair order, of
// (w ? grads[i][k] * w[i] : grads[i][k]).  The product is rounded before the add: the
// library is compiled with -ffp-contract=off, as the reference forms the product out of
// line (scale_rows_) so no FMA fuses it.
__global__ __launch_bounds__(64) void block_sum_kernel(const double* __restrict__ grads,
                                                       const double* __restrict__ w, int64_t N, int gs,
                                                       int64_t block, double* __restrict__ partial) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nblocks = (N + block - 1) / block;
    if (t >= nblocks * gs) return;
    const int64_t blk = t / gs;
    const int k = (int)(t - blk * gs);
    const int64_t i0 = blk * block, i1 = (i0 + block < N) ? i0 + block : N;
    double acc = 0.0;
    if (w) {
        for (int64_t i = i0; i < i1; ++i) {
            const double prod = grads[i * gs + k] * w[i];
            acc += prod;
        }
    } else {
        for (int64_t i = i0; i < i1; ++i) acc += grads[i * gs + k];
    }
    partial[blk * gs + k] = acc;
}

// out[k] = sum of the block partials in block order.
__global__ __launch_bounds__(64) void block_total_kernel(const double* __restrict__ partial, int64_t nblocks,
                                                         int gs, double* __restrict__ out) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= gs) return;
    double t = 0.0;
    for (int64_t b = 0; b < nblocks; ++b) t += partial[b * gs + k];
    out[k] = t;
}

}  // namespace nwg
