// Shared device-side definitions of the gfx950 alignment kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nwg {

constexpr int kWave = 64;

// Problem-type codes, matching include/nwgrad_hip.h.
enum : int { kLinear = 0, kAffine = 1 };
enum : int { kGlobal = 0, kLocal = 1, kGlocal = 2 };

// One tile = the pairs one wavefront fills together.  G lanes per pair (a power of two
// dividing 64), npairs <= 64 / G pairs, their indices at order[first .. first + npairs).
// steps = max_m + G - 1 row steps of the skewed lane-group pipeline.
struct TileDesc {
    int32_t first;
    int16_t G;
    int16_t npairs;
    int32_t max_m;
    int32_t max_n;
};
static_assert(sizeof(TileDesc) == 16, "TileDesc is 16 bytes");

template <typename T>
__device__ __forceinline__ T ninf();
template <>
__device__ __forceinline__ float ninf<float>() { return -__builtin_huge_valf(); }
template <>
__device__ __forceinline__ double ninf<double>() { return -__builtin_huge_val(); }

// IEEE max of two finite-or-(-inf) values.  It equals std::max on every value the DP can
// hold, and differs from it only in the sign bit of a +-0 tie, which no comparison sees.
__device__ __forceinline__ float dmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double dmax(double a, double b) { return __builtin_fmax(a, b); }

// Lane-shift by one towards higher lanes across the whole wavefront (lane 0 keeps its own
// value, which the caller overwrites).  ds_bpermute based: works for any group width.
__device__ __forceinline__ float shift_up(float v) { return __shfl_up(v, 1, kWave); }
__device__ __forceinline__ double shift_up(double v) { return __shfl_up(v, 1, kWave); }
__device__ __forceinline__ int shift_up(int v) { return __shfl_up(v, 1, kWave); }
__device__ __forceinline__ float shift_down(float v) { return __shfl_down(v, 1, kWave); }
__device__ __forceinline__ double shift_down(double v) { return __shfl_down(v, 1, kWave); }
__device__ __forceinline__ int shift_down(int v) { return __shfl_down(v, 1, kWave); }

// The value of the wave's first active lane as a wave-uniform (scalar) value: branches on it
// are uniform to the compiler.  Persistent loops fetch their next tile with it, from an
// atomicAdd that every lane issues (lane 0 adding 1, the others 0) rather than one issued
// under `if (lane == 0)`: with a lane-dependent branch in the loop the compiler has split the
// loop into an outer loop for lane 0 (the fetch) and an inner one that the other lanes never
// leave, each lane then running different tiles and the wave never reconverging.
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float wave_uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double wave_uniform(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Wavefront-local memory ordering for LDS handed between lanes of one wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Waits (the whole wave) until *word >= need, then an agent-scope acquire fence.  Polls with
// relaxed loads (an acquire per poll would invalidate the L2 each time) and sleeps between
// polls.  Bounded in real time: past `ticks` s_memrealtime ticks (100 MHz), or once any wave
// has raised *err, it raises `bit` in *err and returns false; nothing waits after that, and
// the host fails the call.
__device__ __forceinline__ bool bounded_wait(const int32_t* word, int need, int32_t* err, uint64_t ticks, int lane,
                                             int bit) {
    bool ok = true;
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    for (;;) {
        const int have = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (have >= need) break;
        const int flagged = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (flagged != 0 || __builtin_amdgcn_s_memrealtime() - t0 > ticks) {
            if (lane == 0) __hip_atomic_fetch_or(err, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = false;
            break;
        }
        __builtin_amdgcn_s_sleep(8);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return ok;
}

// Makes the wave's earlier global writes visible at agent scope, then stores v to *word
// (release): a bounded_wait on the word that sees v sees the writes.
__device__ __forceinline__ void publish(int32_t* word, int v, int lane) {
    __threadfence();
    if (lane == 0) __hip_atomic_store(word, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace nwg
