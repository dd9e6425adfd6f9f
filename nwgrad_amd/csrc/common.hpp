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

// The same maximum of two finite-or-(-inf) float32 values as one median, med3(a, b, FLT_MAX)
// (no value here exceeds the largest finite float): fmax's operands must be NaN-quieted first
// wherever the compiler cannot prove them results of arithmetic (values carried around a
// loop, lane shifts), one more instruction each; the median takes them as they are.  (With
// +inf as the third operand the compiler folds the median back into fmax.)  float64 has no
// median instruction and keeps fmax.
__device__ __forceinline__ float vmax(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 3.40282347e+38f); }
__device__ __forceinline__ double vmax(double a, double b) { return dmax(a, b); }

// One-lane wave shifts on the DPP path (a VALU move, no LDS round trip): up = from the lane
// below (wave_shr:1), down = from the lane above (wave_shl:1).  Lane 0 (up) and lane 63
// (down) receive 0; every caller replaces or ignores that value.
__device__ __forceinline__ int dpp_up(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_down(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ float wshift_up(float v) { return __int_as_float(dpp_up(__float_as_int(v))); }
__device__ __forceinline__ float wshift_down(float v) { return __int_as_float(dpp_down(__float_as_int(v))); }
__device__ __forceinline__ int wshift_up(int v) { return dpp_up(v); }
__device__ __forceinline__ int wshift_down(int v) { return dpp_down(v); }
__device__ __forceinline__ double wshift_up(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_up((int)(b & 0xffffffff)), hi = dpp_up((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double wshift_down(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_down((int)(b & 0xffffffff)), hi = dpp_down((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Lane-shift by one towards higher lanes across the whole wavefront, as the DP's lane
// pipelines use it: every lane of the wave active, lane 0 receiving 0, which the caller
// overwrites (its column 0 or its panel boundary).
__device__ __forceinline__ float shift_up(float v) { return wshift_up(v); }
__device__ __forceinline__ double shift_up(double v) { return wshift_up(v); }
__device__ __forceinline__ int shift_up(int v) { return wshift_up(v); }

// A lane's value for the whole wave, the lane given by a wave-uniform index (a lane read into
// a scalar register, not a permute through LDS).
__device__ __forceinline__ float bcast(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double bcast(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Loaded values made ordinary register values: waits here for the loads that produce them (the
// empty asm reads and writes each register), so that later uses -- in later trips of a loop,
// where the compiler can no longer count the loads in flight and waits for every outstanding
// vector memory operation, stores included -- need no wait at all.
template <typename V>
__device__ __forceinline__ void settle(V& v) {
    asm volatile("" : "+v"(v));
}
template <typename V, typename... Vs>
__device__ __forceinline__ void settle(V& v, Vs&... vs) {
    settle(v);
    settle(vs...);
}

// The value of the wave's first active lane as a wave-uniform (scalar) value: branches on it
// are uniform to the compiler.
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

// The next index of a persistent queue, for the whole wave: one atomic add of 1, issued by
// lane 0 alone, whose old value comes back wave-uniform.  The narrowing to lane 0 happens
// inside one asm block (save exec, exec = lane 0, the returning atomic, wait for it, restore
// exec), so the source holds no lane-dependent branch: a loop that fetches with it stays one
// uniform loop with a scalar exit test.  Both simpler forms cost something:
//  - `if (lane == 0) old = atomicAdd(...)` has let the compiler restructure the persistent
//    loop so that its exit became an exec-mask exit lane 0 alone could take (the other lanes
//    never left the loop);
//  - atomicAdd(counter, lane == 0 ? 1 : 0) from every lane is safe, but its per-lane operand
//    makes the compiler's atomic optimizer scan the 64 operands one lane at a time before its
//    single atomic, on every fetch.
// The caller reaches it with lane 0 active (the persistent loops run with all 64 lanes).
// The s_nop 4 after each exec write keeps the next vector instruction clear of it.
__device__ __forceinline__ int wave_fetch(int32_t* counter) {
    int old;
    uint64_t saved;
    asm volatile(
        "s_mov_b64 %1, exec\n\t"
        "s_mov_b64 exec, 1\n\t"
        "s_nop 4\n\t"
        "global_atomic_add %0, %2, %3, %4 sc0\n\t"
        "s_waitcnt vmcnt(0)\n\t"
        "s_mov_b64 exec, %1\n\t"
        "s_nop 4"
        : "=&v"(old), "=&s"(saved)
        : "v"(0), "v"(1), "s"(counter)
        : "memory");
    return __builtin_amdgcn_readfirstlane(old);
}

// The walk's outputs besides the gradient, written by the one lane that walks a pair:
//  - the guide (the reference's guide_j: gj[i] = the column after row i of A,
//    aligner.hpp:3155-3160 affine, :2761-2773 linear): gj[0] = 0, gj[m] = n, then the end cell's
//    gj[best_i] = best_j; each M move out of (i, j) sets gj[i-1] = j-1, each move in B (X, up)
//    gj[i-1] = j; a Global / Glocal column-0 run sets its rows to 0.  Rows a Local path does not
//    cross are interpolated as fill_guide_gaps does (aligner.hpp:2734-2749): between gj[0] and
//    the start row, and between the end row and gj[m];
//  - the stored path (keep_paths): run-length ops (type << 30) | length, type 0 both advance,
//    1 a gap in B (A advances), 2 a gap in A, written backward from the end of the pair's
//    buffer of m + n entries as the walk meets the moves last first (batch_engine.hpp:2056-2092,
//    OpsSink), so they read forward in path order.
struct WalkOut {
    int32_t* gj = nullptr;    // the pair's m + 1 guide entries (null: no guide)
    uint32_t* w = nullptr;    // the next op slot, one past it (null: no stored path)
    uint32_t t = 3u, len = 0u;
    int low = -1, lowv = 0;   // the lowest row a move's guide entry wrote, and its value

    __device__ __forceinline__ void begin(int32_t* g, uint32_t* ops_end, int m, int n, int bi, int bj) {
        gj = g;
        w = ops_end;
        if (gj) {
            gj[0] = 0;
            gj[m] = n;
            gj[bi] = bj;
        }
    }
    __device__ __forceinline__ void op(uint32_t k, uint32_t cnt) {
        if (!w) return;
        if (k == t) {
            len += cnt;
            return;
        }
        if (len) *--w = (t << 30) | len;
        t = k;
        len = cnt;
    }
    // a move out of cell (i, j): k 0 both advance, 1 a gap in B, 2 a gap in A
    __device__ __forceinline__ void move(int k, int i, int j) {
        if (gj && k != 2) {
            const int v = k == 0 ? j - 1 : j;
            gj[i - 1] = v;
            low = i - 1;
            lowv = v;
        }
        op((uint32_t)k, 1u);
    }
    // the Global / Glocal run down column 0 from row i, and along row 0 from column j
    __device__ __forceinline__ void column0(int i) {
        if (gj)
            for (int r = 0; r < i; ++r) gj[r] = 0;
        if (i > 0) {
            low = 0;
            lowv = 0;
        }
        op(1u, (uint32_t)i);
    }
    __device__ __forceinline__ void row0(int j) { op(2u, (uint32_t)j); }
    // After the walk: Local's interpolated rows (i0 the row the walk stopped in); the number of
    // ops written, counted from ops_end.
    // Each interpolated entry is the exact floor of base + span * k / rows, stepped as
    // q = span / rows and a remainder carried against rows: two divisions a run, none a row.
    __device__ __forceinline__ void end_local(int i0, int m, int n, int bi, int bj) {
        if (!gj) return;
        const int gi0 = (low == i0) ? lowv : (i0 == bi ? bj : 0);
        ramp(1, i0, 0, gi0, i0);                 // rows 1 .. i0-1: gi0 * k / i0
        ramp(bi + 1, m, bj, n - bj, m - bi);     // rows bi+1 .. m-1: bj + (n - bj)(k - bi) / (m - bi)
    }
    // gj[k] = base + floor(span * (k - k0 + 1) / rows) for k in [k0, k1); span >= 0, rows >= 1
    __device__ __forceinline__ void ramp(int k0, int k1, int base, int span, int rows) {
        if (k0 >= k1) return;
        const int q = span / rows, r = span - q * rows;
        int v = base, e = 0;
        for (int k = k0; k < k1; ++k) {
            v += q;
            e += r;
            if (e >= rows) {
                e -= rows;
                ++v;
            }
            gj[k] = v;
        }
    }
    __device__ __forceinline__ int32_t finish(const uint32_t* ops_end) {
        if (!w) return 0;
        if (len) *--w = (t << 30) | len;
        len = 0;
        return (int32_t)(ops_end - w);
    }
};

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
