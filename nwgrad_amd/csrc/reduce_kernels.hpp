// Reductions of per-pair gradient rows: the reference's blocked weighted sum
// (batch_engine.hpp:697-718, 1788-1827), the same additions in the same order.
#pragma once

#include "common.hpp"

namespace nwg {

// partial[blk * gs + k] = sum over pairs i of block blk, in pair order, of
// (w ? grads[i][k] * w[i] : grads[i][k]).  The product is rounded before the add: the
// library is compiled with -ffp-contract=off, as the reference forms the product out of
// line (scale_rows_) so no FMA fuses it.
// ok (soft batches, may be null): a pair whose flag is 0 -- a float32 refusal not yet rerun in
// double, as in an asynchronous or captured training step -- adds nothing.  After a finished
// pass every flag is 1, and the sum is the reference's.
__global__ __launch_bounds__(64) void block_sum_kernel(const double* __restrict__ grads,
                                                       const double* __restrict__ w, const uint8_t* __restrict__ ok,
                                                       int64_t N, int gs, int64_t block,
                                                       double* __restrict__ partial) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nblocks = (N + block - 1) / block;
    if (t >= nblocks * gs) return;
    const int64_t blk = t / gs;
    const int k = (int)(t - blk * gs);
    const int64_t i0 = blk * block, i1 = (i0 + block < N) ? i0 + block : N;
    double acc = 0.0;
    if (w) {
        for (int64_t i = i0; i < i1; ++i) {
            if (ok && !ok[i]) continue;
            const double prod = grads[i * gs + k] * w[i];
            acc += prod;
        }
    } else {
        for (int64_t i = i0; i < i1; ++i)
            if (!ok || ok[i]) acc += grads[i * gs + k];
    }
    partial[blk * gs + k] = acc;
}

// The refusal record of a soft batch's main passes: rec[0] pair-passes the float32 pass refused,
// rec[1] the first pass (1-based) with a refusal, 0 for none, rec[2] passes run.  Per pass,
// soft_refusal_step_kernel advances rec[2], then this scan (the next kernel on the stream)
// counts the refused flags and names the pass in rec[1] if it is the first with any.
__global__ __launch_bounds__(256) void soft_refusal_scan_kernel(const uint8_t* __restrict__ ok, int64_t N,
                                                                unsigned long long* __restrict__ rec) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int mine = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (int64_t)gridDim.x * 256)
        mine += ok[i] ? 0 : 1;
    if (mine) atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0 && cnt > 0) {
        atomicAdd(&rec[0], (unsigned long long)cnt);
        atomicCAS(&rec[1], 0ull, rec[2]);
    }
}

// Advances the pass counter rec[2] (one thread; launched before the scan of the same pass).
__global__ void soft_refusal_step_kernel(unsigned long long* __restrict__ rec) {
    if (threadIdx.x == 0) rec[2] += 1ull;
}

// p[0 .. n) = 0: counters and progress words cleared by a kernel rather than a memset, so the
// pass can be captured in a graph without memset nodes.
__global__ __launch_bounds__(256) void clear_words_kernel(int32_t* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

// The batch total without a pass in pair order, where no partial sum can round: the scores'
// sum and sum of magnitudes (in any order), the lowest set-bit exponent over the nonzero
// scores (each score is then a multiple of 2^qmin) and whether any score is NaN or infinite.
// The host takes `sum` as the pair-order total only when mag < 2^(53 + qmin) with margin:
// every partial sum, in any order, is then a multiple of 2^qmin below 2^(53 + qmin) in
// magnitude, a double, so no addition rounds and all orders give the same total.
struct ScoreStats {
    double sum, mag;
    int32_t qmin, bad;
    int32_t pad[2];
};
static_assert(sizeof(ScoreStats) == 32, "ScoreStats is 32 bytes");
constexpr int kStatsThreads = 256;
constexpr int32_t kNoBit = 0x7fffffff;   // qmin of an all-zero set

// Exponent of the lowest set bit of a finite nonzero double: s = odd * 2^result.
__device__ __forceinline__ int lowbit_exponent(double s) {
    const uint64_t bits = (uint64_t)__double_as_longlong(s);
    const uint64_t frac = bits & ((1ull << 52) - 1);
    const int e = (int)((bits >> 52) & 0x7ff);
    if (e == 0) return __builtin_ctzll(frac) - 1074;   // subnormal: frac * 2^-1074
    return __builtin_ctzll(frac | (1ull << 52)) + e - 1075;
}

__device__ __forceinline__ ScoreStats stats_merge(ScoreStats a, const ScoreStats& b) {
    a.sum += b.sum;
    a.mag += b.mag;
    a.qmin = a.qmin < b.qmin ? a.qmin : b.qmin;
    a.bad |= b.bad;
    return a;
}

__device__ __forceinline__ ScoreStats stats_block_reduce(ScoreStats v) {
    __shared__ ScoreStats waves[kStatsThreads / kWave];
    for (int off = kWave / 2; off > 0; off >>= 1) {
        ScoreStats o;
        o.sum = __shfl_xor(v.sum, off, kWave);
        o.mag = __shfl_xor(v.mag, off, kWave);
        o.qmin = __shfl_xor(v.qmin, off, kWave);
        o.bad = __shfl_xor(v.bad, off, kWave);
        v = stats_merge(v, o);
    }
    const int w = (int)threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) waves[w] = v;
    __syncthreads();
    ScoreStats r = waves[0];
    for (int k = 1; k < kStatsThreads / kWave; ++k) r = stats_merge(r, waves[k]);
    return r;
}

// Stage 1: part[blockIdx.x] over a grid-stride share of the scores.
__global__ __launch_bounds__(kStatsThreads) void score_stats_kernel(const double* __restrict__ scores, int64_t N,
                                                                    ScoreStats* __restrict__ part) {
    ScoreStats v{0.0, 0.0, kNoBit, 0, {0, 0}};
    for (int64_t i = (int64_t)blockIdx.x * kStatsThreads + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kStatsThreads) {
        const double s = scores[i];
        const double a = __builtin_fabs(s);
        if (!(a <= 1.7976931348623157e308)) {
            v.bad = 1;
        } else if (s != 0.0) {
            const int q = lowbit_exponent(s);
            v.qmin = q < v.qmin ? q : v.qmin;
            v.sum += s;
            v.mag += a;
        }
    }
    const ScoreStats r = stats_block_reduce(v);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// Stage 2, one block: out[0] over the n stage-1 parts.
__global__ __launch_bounds__(kStatsThreads) void score_stats_final_kernel(const ScoreStats* __restrict__ part, int n,
                                                                          ScoreStats* __restrict__ out) {
    ScoreStats v{0.0, 0.0, kNoBit, 0, {0, 0}};
    for (int k = (int)threadIdx.x; k < n; k += kStatsThreads) v = stats_merge(v, part[k]);
    const ScoreStats r = stats_block_reduce(v);
    if (threadIdx.x == 0) out[0] = r;
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
