// Hard and score-only fills of pairs wider than one wavefront's columns (len(B) > 64 S):
// column panels pipelined across waves, then a traceback walk per pair.
//
// Reference behaviour: the same cells, codes and end cells as hard_kernels.hpp (relative to
// the reference's src/nwgrad/cpp/nwgrad/: aligner.hpp:3955-4071, 3756-3865, 5975-5984,
// 6121-6134), over the full pointer table the reference's own long-pair default keeps
// (traceback="auto": Grid, aligner.hpp:89-100 and 5267-5742, returns the Pointers fill's
// exact path); the walk is aligner.hpp:2986-3056 / 2826-2855.
//
// Fill.  A pair of len(B) = n is cut into ceil(n / (64 S)) panels of 64 S columns.  Each
// (pair, panel) is one item of a device queue, filled by one wave exactly as hard_tile fills
// a one-pair tile, except that lane 0 takes its left neighbour from the previous panel's
// right boundary column instead of the matrix border.  The panel's last lane writes its
// right column, row by row, to HBM and publishes every RB rows (and row m) through the
// item's counter with a release store; the next panel's wave waits for a block's rows with
// acquire loads before it starts the block.  A pair's items are queued in panel order, so
// every awaited block belongs to an item a running wave took earlier: the waits end.  Each
// wait is still bounded in real time; past it the wave flags an error (the host raises)
// and goes on, and no wave waits once the flag is up.  Pointer codes go to a
// per-item table in the hard fill's layout.
//
// Walk.  One block per pair combines the panels' end-cell candidates (larger value, ties to
// the smaller (i, j): the reference's row-major first rule), then walks the codes back.
// The walk only ever moves to the same, the previous or the one-before-previous step of a
// panel's table, so it reads the table through an LDS window of 48 KiB of steps that all
// threads of the block refill together.  Within a panel the step only decreases, so the next
// window is known: it is loaded into registers while the walk runs through the current one,
// and a refill is a store to LDS.  The window also stages the codes of B in the panel's
// columns and those of A (tracked: the row sets too) in the window's rows, so a walk step
// reads LDS only.
#pragma once

#include "hard_kernels.hpp"

namespace nwg {

constexpr int kWalkWindowBytes = 48 << 10;   // the walk's LDS window of table steps (S * 64 bytes each)
constexpr int kWalkThreads = 256;  // threads of a walk block (they refill the window together)
constexpr int kWalkRows = 160;     // rows of A staged per window: at least its steps + kWave - 1
template <int S>
__host__ __device__ constexpr int walk_steps() { return kWalkWindowBytes / (S * kWave); }

struct PanelItem {
    int32_t pair;
    int32_t panel;
    int64_t table_off;   // dwords: this panel's pointer codes (hard)
    int64_t bnd_in;      // T elements: the previous panel's right column (3 per row); -1 for the first panel
    int64_t bnd_out;     // T elements: this panel's right column; -1 for the last panel
};
static_assert(sizeof(PanelItem) == 32, "PanelItem is 32 bytes");

struct LongPair {
    int32_t pair;
    int32_t item;     // its first item (panel 0)
    int32_t panels;
    int32_t pad;
};

template <typename T>
struct CandRec {
    T v;
    int32_t i, j, st;
};

template <typename T>
struct LongArgs {
    const uint8_t* codes_a;
    const uint8_t* codes_b;
    const int64_t* off_a;
    const int64_t* off_b;
    const PanelItem* items;
    int32_t n_items;
    int32_t* queue;
    const LongPair* lpairs;
    const T* sub;         // K matrices cast to T (K = 1 untracked)
    int32_t na;
    T go_a, ge_a, go_b, ge_b;
    TrackArgs<T> tr;      // tracked batches (the TR kernels) only
    int32_t* progress;    // per item: rows of its right column published
    T* bnd;               // right boundary columns
    CandRec<T>* cand;     // per item: the panel's end-cell candidate
    uint32_t* table;      // pointer codes (hard)
    double* scores;
    double* grads;
    int32_t* err;         // bit 0: a wait passed its bound (or the walk met an invalid end cell)
    uint64_t wait_ticks;  // bound of one wait, in s_memrealtime ticks
    // hard only, each may be null: guides, stored paths, op counts and end cells, as HardArgs
    int32_t* guide;
    uint32_t* ops;
    int32_t* ops_n;
    int32_t* ends;
    int32_t band;         // the BAND kernels: the half-width around each pair's guide (> 0)
};

template <typename T, int GM, int AM, int S, bool HARD, int RB, bool TR, bool BAND>
__device__ void long_panel(const LongArgs<T>& A, const T* __restrict__ ssub, uint8_t* __restrict__ acodes, int it,
                           int lane) {
    static_assert(RB >= 1 && RB <= kWave, "a row block fits the lanes of a wave");
    const PanelItem item = A.items[it];
    const int p = item.pair, na = A.na;
    const int64_t a0 = A.off_a[p], b0 = A.off_b[p];
    const int m = (int)(A.off_a[p + 1] - a0), n = (int)(A.off_b[p + 1] - b0);
    const uint8_t* ap = A.codes_a + a0;
    const uint8_t* bp = A.codes_b + b0;
    const int g = lane;
    const int j0 = item.panel * (kWave * S) + g * S;
    const int ncols = n - j0 < 0 ? 0 : (n - j0 > S ? S : n - j0);
    T go_a = A.go_a, ge_a = A.ge_a, go_b = A.go_b, ge_b = A.ge_b;
    const int nn = na * na, K = TR ? A.tr.K : 1;
    const T* __restrict__ gT = ssub + K * nn;   // tracked: the K x 4 gap table in LDS
    const uint16_t* tk = A.tr.trk + (TR ? a0 : 0);
    const int64_t r0 = TR ? a0 + p : 0;
    if constexpr (TR) {   // row 0's gaps in A
        const int s0 = A.tr.gtrk[r0];
        go_a = gT[4 * s0]; ge_a = gT[4 * s0 + 1];
    }
    const T NI = ninf<T>();
    const bool first = item.bnd_in < 0, last = item.bnd_out < 0;
    // banded: the pair's guide, the band clamped to n, row 0's last column (as in hard_tile)
    const int32_t* bgj = nullptr;
    int bw = 0, bj0 = 0;
    if constexpr (BAND) {
        bgj = A.guide + a0 + p;
        bw = A.band < n ? A.band : n;
        bj0 = band_hi<AM>(bgj, 0, m, n, bw);
    }

    int coff[S];
#pragma unroll
    for (int k = 0; k < S; ++k) coff[k] = (k < ncols) ? (int)bp[j0 + k] : 0;

    // Row 0 of the strip and the diagonal of row 1's first cell, as in hard_tile.
    T M[S], X[S], Y[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const T jt = (T)(j0 + k + 1);
        if constexpr (GM == kAffine) {
            if (AM == kGlobal) { M[k] = NI; X[k] = NI; Y[k] = -(go_a + jt * ge_a); }
            else { M[k] = T(0); X[k] = NI; Y[k] = NI; }
        } else {
            M[k] = (AM == kGlobal) ? (-jt) * ge_a : T(0);
            X[k] = NI; Y[k] = NI;
        }
        if constexpr (BAND && AM == kGlobal) {
            if (j0 + k + 1 > bj0) { if (GM == kAffine) Y[k] = NI; else M[k] = NI; }
        }
    }
    T pM, pX, pY;
    {
        const T jt = (T)j0;
        if constexpr (GM == kAffine) {
            if (AM == kGlobal) {
                pM = (j0 == 0) ? T(0) : NI; pX = NI; pY = (j0 == 0) ? NI : -(go_a + jt * ge_a);
                if (BAND && j0 > bj0) pY = NI;
            } else { pM = T(0); pX = NI; pY = NI; }
        } else {
            pM = (AM == kGlobal) ? (-jt) * ge_a : T(0); pX = NI; pY = NI;
            if (BAND && AM == kGlobal && j0 > bj0) pM = NI;
        }
    }
    T oM = M[S - 1], oX = X[S - 1], oY = Y[S - 1];

    Cand<T> best;
    best.v = T(0); best.i = (AM == kLocal) ? 0 : -1; best.j = 0; best.st = 0;

    const T* __restrict__ bin = A.bnd + (first ? 0 : item.bnd_in);
    T* __restrict__ bout = A.bnd + (last ? 0 : item.bnd_out);
    int32_t* prog_in = A.progress + (first ? it : it - 1);
    int32_t* prog_out = A.progress + it;
    uint32_t* tb = A.table + item.table_off;
    T bM = NI, bX = NI, bY = NI;   // lane r: the left boundary of row t0 + 1 + r of the current block

    const int steps = m + kWave - 1;
    // The codes of A pass through a 256-row ring in LDS: at step t0 = 64 k the wave stores rows
    // t0 + 1 .. t0 + 64, loaded one block earlier, and loads the next block; lane g reads row
    // t - g + 1 from LDS.  A load in the step itself would make the step wait for every vector
    // memory operation in flight (the loop's trips cannot be counted apart), the previous
    // step's code stores included.
    auto arow = [&](int r) -> int { return (r >= 1 && r <= m) ? (int)ap[r - 1] : 0; };
    int anext = arow(lane + 1);
    for (int t = 0; t < steps; ++t) {
        const int i = t - g + 1;
        if (t % kWave == 0) {
            acodes[(t + 1 + lane) & 255] = (uint8_t)anext;
            anext = arow(t + kWave + 1 + lane);
            wave_sync();
        }
        int rs = 0;   // tracked: the row's set (matrix, gaps in B)
        if constexpr (TR) {
            if (i >= 1 && i <= m) {
                rs = tk[i - 1];
                const int gsi = A.tr.gtrk[r0 + i];
                go_a = gT[4 * gsi]; ge_a = gT[4 * gsi + 1];
                go_b = gT[4 * rs + 2]; ge_b = gT[4 * rs + 3];
            }
        }
        if (!first && t < m && t % RB == 0) {
            // Rows t+1 .. min(t+RB, m) of the previous panel's right column.
            const int need = t + RB < m ? t + RB : m;
            // Poll with relaxed loads (an acquire per poll would invalidate the L2 each time),
            // then one acquire fence.  The wait is bounded in real time (s_memrealtime ticks);
            // past the bound, or once any wave has flagged, nothing waits and the call fails.
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            for (;;) {
                const int have = __builtin_amdgcn_readfirstlane(
                    __hip_atomic_load(prog_in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                if (have >= need) break;
                const int flagged = __builtin_amdgcn_readfirstlane(
                    __hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                if (flagged != 0 || __builtin_amdgcn_s_memrealtime() - t0 > A.wait_ticks) {
                    if (lane == 0) __hip_atomic_fetch_or(A.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(8);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const int r = t + 1 + lane;
            if (lane < RB && r <= m) {
                bM = bin[3 * (int64_t)r];
                if constexpr (GM == kAffine) { bX = bin[3 * (int64_t)r + 1]; bY = bin[3 * (int64_t)r + 2]; }
            }
            settle(bM, bX, bY);   // waited for here, once a block, not at every step's use
        }
        T iM = shift_up(oM);
        T iX, iY;
        if constexpr (GM == kAffine) { iX = shift_up(oX); iY = shift_up(oY); }
        else { iX = NI; iY = NI; }
        if (!first) {
            const int sl = t % RB;   // wave-uniform: a lane read, not a permute
            const T lM = bcast(bM, sl);
            T lX = NI, lY = NI;
            if constexpr (GM == kAffine) { lX = bcast(bX, sl); lY = bcast(bY, sl); }
            if (g == 0) { iM = lM; iX = lX; iY = lY; }
        } else if (g == 0) {   // column 0 of row i (aligner.hpp:4009-4018, 3823-3826)
            T border = NI;
            if constexpr (AM != kLocal) {
                if (!TR || (i >= 1 && i <= m)) border = col0_border<T, GM, TR>(A.tr, r0, i, go_b, ge_b);
                if constexpr (BAND) {
                    if (!(i >= 1 && i <= m && bgj[i] <= bw)) border = NI;
                }
            }
            if constexpr (GM == kAffine) {
                if (AM == kLocal) { iM = T(0); iX = NI; }
                else { iM = NI; iX = border; }
                iY = NI;
            } else {
                iM = (AM == kLocal) ? T(0) : border;
            }
        }
        const bool active = i >= 1 && i <= m && ncols > 0;
        if (active) {
            const T* __restrict__ srow = ssub + ((TR ? rs * na : 0) + (int)acodes[i & 255]) * na;
            uint32_t cw[S / 4];
            T rowmax;
            int klo = 0, khi = S;
            if constexpr (BAND) {
                klo = band_lo(bgj, i, bw) - j0 - 1;
                khi = band_hi<AM>(bgj, i, m, n, bw) - j0 - 1;
            }
            strip_row<T, GM, AM, S, HARD, BAND>(M, X, Y, iM, iX, iY, pM, pX, pY, srow, coff, ncols, go_a, ge_a, go_b,
                                                ge_b, cw, rowmax, klo, khi);
            oM = M[S - 1];
            if constexpr (GM == kAffine) { oX = X[S - 1]; oY = Y[S - 1]; }
            if constexpr (HARD) {
#pragma unroll
                for (int k = 0; k < S / 4; ++k)
                    tb[((int64_t)t * (S / 4) + k) * kWave + lane] = cw[k];
            }
            if constexpr (AM == kLocal) local_update<T, GM, S>(best, M, X, Y, rowmax, i, j0, ncols);
            if (!last && g == kWave - 1) {
                bout[3 * (int64_t)i] = M[S - 1];
                if constexpr (GM == kAffine) { bout[3 * (int64_t)i + 1] = X[S - 1]; bout[3 * (int64_t)i + 2] = Y[S - 1]; }
                if (i % RB == 0 || i == m)
                    __hip_atomic_store(prog_out, i, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        pM = iM; pX = iX; pY = iY;
    }

    T col0v = (m == 0 || AM == kLocal) ? T(0) : col0_border<T, GM, TR>(A.tr, r0, m, A.go_b, A.ge_b);
    if constexpr (BAND) {
        if (AM != kLocal && m > 0 && bgj[m] > bw) col0v = NI;
    }
    Cand<T> c = end_candidate<T, GM, AM, S>(M, X, Y, best, m, n, j0, ncols, item.panel == 0 && g == 0, col0v);
    c = reduce_slot(c, kWave);
    if (lane == 0) {
        CandRec<T> r;
        r.v = c.v; r.i = c.i; r.j = c.j; r.st = c.st;
        A.cand[it] = r;
    }
}

// One wavefront per block, pulling (pair, panel) items until the queue is drained; every
// wave reaches the exit test once the counter passes n_items.
template <typename T, int GM, int AM, int S, bool HARD, int RB, bool TR, bool BAND = false>
__global__ __launch_bounds__(64) void long_fill_kernel(LongArgs<T> A) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int nn = A.na * A.na, K = TR ? A.tr.K : 1;
    T* ssub = reinterpret_cast<T*>(lds);
    uint8_t* acodes = lds + hard_lds_sub_bytes<T>(A.na, K, TR);   // 256 bytes: the ring of A's codes
    const int lane = threadIdx.x;
    for (int k = lane; k < K * nn; k += kWave) ssub[k] = A.sub[k];
    if constexpr (TR)
        for (int k = lane; k < 4 * K; k += kWave) ssub[K * nn + k] = A.tr.gaps[k];
    __syncthreads();
    for (;;) {
        const int it = wave_fetch(A.queue);
        if (it >= A.n_items) break;
        long_panel<T, GM, AM, S, HARD, RB, TR, BAND>(A, ssub, acodes, it, lane);
    }
}

// LDS of the long-pair walk: K * n_alpha^2 uint32 match counts and (tracked) K x 4 int gap
// counters, then the table window, the panel's codes of B, and the window rows' codes of A
// (tracked: their sets and gap sets, uint16).
__host__ __device__ constexpr int long_walk_count_bytes(int na, int K, bool tr) {
    return ((K * na * na + (tr ? 4 * K : 0)) * 4 + 15) & ~15;
}
__host__ __device__ constexpr int long_walk_lds_bytes(int na, int K, bool tr, int S) {
    return long_walk_count_bytes(na, K, tr) + kWalkWindowBytes + kWave * S + kWalkRows +
           (tr ? 2 * kWalkRows * 2 : 0);
}

// The walk's prefetched window: a thread's share of the window's table steps, 12 vectors of 4
// dwords (48 KiB / 16 B / 256 threads, whatever the panel width).  Separate fields of a plain
// vector type, reached through compile-time indices, so the compiler keeps them in registers
// while they are carried across walk steps (an array of uint4 there stayed in scratch).
typedef uint32_t walk_u4 __attribute__((ext_vector_type(4)));   // a plain vector: promoted to registers
struct WalkBuf {
    walk_u4 v0, v1, v2, v3, v4, v5, v6, v7, v8, v9, v10, v11;
};
template <int R>
__device__ __forceinline__ walk_u4& walk_field(WalkBuf& b) {
    static_assert(R >= 0 && R < 12, "WalkBuf holds 12 vectors");
    if constexpr (R == 0) return b.v0;
    else if constexpr (R == 1) return b.v1;
    else if constexpr (R == 2) return b.v2;
    else if constexpr (R == 3) return b.v3;
    else if constexpr (R == 4) return b.v4;
    else if constexpr (R == 5) return b.v5;
    else if constexpr (R == 6) return b.v6;
    else if constexpr (R == 7) return b.v7;
    else if constexpr (R == 8) return b.v8;
    else if constexpr (R == 9) return b.v9;
    else if constexpr (R == 10) return b.v10;
    else return b.v11;
}
// Thread tid's share of n4 vectors from src (the last one repeated past n4: every load in flight
// at once), and back out to dst below n4.
template <int R>
__device__ __forceinline__ void walk_load(WalkBuf& b, const walk_u4* __restrict__ src, int n4, int tid) {
    if constexpr (R < 12) {
        const int q = tid + R * kWalkThreads;
        walk_field<R>(b) = src[q < n4 ? q : n4 - 1];
        walk_load<R + 1>(b, src, n4, tid);
    }
}
template <int R>
__device__ __forceinline__ void walk_store(WalkBuf& b, walk_u4* dst, int n4, int tid) {
    if constexpr (R < 12) {
        const int q = tid + R * kWalkThreads;
        if (q < n4) dst[q] = walk_field<R>(b);
        walk_store<R + 1>(b, dst, n4, tid);
    }
}

// Window [lo, hi] of a panel's table (its first dword at `panel`) into registers: the thread's
// share of the steps, and for thread tid < kWalkRows the staged row alo + tid of A, alo =
// max(0, lo - 63) (a step t of the window is row t - go + 1 for a lane go in 0 .. 63): its
// code and, tracked, its set and gap set (gt: the pair's gap sets, one per DP row).
template <bool TR, int SD>
__device__ __forceinline__ void walk_fetch(WalkBuf& buf, int& ra, int& rtk, int& rgt, const uint32_t* panel,
                                           int lo, int hi, int tid, int m, const uint8_t* ap, const uint16_t* tk,
                                           const uint16_t* gt) {
    walk_load<0>(buf, reinterpret_cast<const walk_u4*>(panel + (int64_t)lo * SD), (hi - lo + 1) * (SD / 4), tid);
    const int alo = lo - (kWave - 1) > 0 ? lo - (kWave - 1) : 0;
    const int x = alo + tid;
    if (tid < kWalkRows && x <= hi && x < m) {
        ra = ap[x];
        if constexpr (TR) {
            rtk = tk[x];
            rgt = gt[x + 1];
        }
    }
}
// The fetched window [lo, hi] into LDS.
template <bool TR, int SD>
__device__ __forceinline__ void walk_put(WalkBuf& buf, int ra, int rtk, int rgt, uint32_t* win, uint8_t* sa,
                                         uint16_t* stk, uint16_t* sgt, int lo, int hi, int tid) {
    walk_store<0>(buf, reinterpret_cast<walk_u4*>(win), (hi - lo + 1) * (SD / 4), tid);
    if (tid < kWalkRows) {
        sa[tid] = (uint8_t)ra;
        if constexpr (TR) {
            stk[tid] = (uint16_t)rtk;
            sgt[tid] = (uint16_t)rgt;
        }
    }
}

// Block b: pair lpairs[b].  Combines its panels' candidates into the end cell, writes the
// score, and (hard) walks the codes back, counting matches in an LDS histogram and gap moves
// in registers, or (tracked) every count in its row's set in LDS.  Every thread runs the walk
// with the same values; thread 0 writes the counts.
template <typename T, int GM, int AM, int S, bool HARD, bool TR>
__global__ __launch_bounds__(kWalkThreads) void long_walk_kernel(LongArgs<T> A) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    const int na = A.na, nn = na * na, K = TR ? A.tr.K : 1, knn = K * nn, gs = K * (nn + 4);
    const LongPair lp = A.lpairs[blockIdx.x];
    const int p = lp.pair;

    Cand<T> c;
    c.v = T(0); c.i = -1; c.j = 0; c.st = 0;
    for (int base = 0; base < lp.panels; base += kWave) {
        Cand<T> o;
        o.v = T(0); o.i = -1; o.j = 0; o.st = 0;
        if (base + lane < lp.panels) {
            const CandRec<T> r = A.cand[lp.item + base + lane];
            o.v = r.v; o.i = r.i; o.j = r.j; o.st = r.st;
        }
        o = reduce_slot(o, kWave);
        if (cand_better(o, c)) c = o;
    }
    c.v = wave_uniform(c.v);   // lane 0 holds the combined candidate
    c.i = wave_uniform(c.i);
    c.j = wave_uniform(c.j);
    c.st = wave_uniform(c.st);
    if (tid == 0) A.scores[p] = (double)c.v;
    if constexpr (HARD) {
        uint32_t* hist = reinterpret_cast<uint32_t*>(lds);
        int* gcnt = reinterpret_cast<int*>(hist + knn);   // tracked: K x 4
        uint32_t* win = reinterpret_cast<uint32_t*>(lds + long_walk_count_bytes(na, K, TR));
        for (int k = tid; k < knn + (TR ? 4 * K : 0); k += kWalkThreads) hist[k] = 0u;
        __syncthreads();
        const int64_t a0 = A.off_a[p], b0 = A.off_b[p];
        const int m = (int)(A.off_a[p + 1] - a0), n = (int)(A.off_b[p + 1] - b0);
        const uint8_t* ap = A.codes_a + a0;
        const uint8_t* bp = A.codes_b + b0;
        const uint16_t* tk = A.tr.trk + (TR ? a0 : 0);
        const int64_t r0 = TR ? a0 + p : 0;
        constexpr int PW = kWave * S;
        constexpr int SD = (S / 4) * kWave;   // dwords per table step
        int c_goa = 0, c_gea = 0, c_gob = 0, c_geb = 0;
        // count a gap move: f 0 go_a, 1 ge_a (set gtrk of row `row`), 2 go_b, 3 ge_b (set trk)
        auto gap = [&](int f, int row, int v) {
            if constexpr (TR) {
                const int sset = f < 2 ? A.tr.gtrk[r0 + row] : tk[row - 1];
                if (tid == 0)
                    __hip_atomic_fetch_add(&gcnt[4 * sset + f], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
                if (f == 0) c_goa += v;
                else if (f == 1) c_gea += v;
                else if (f == 2) c_gob += v;
                else c_geb += v;
            }
        };
        int i = c.i, j = c.j, st = c.st;
        if (i < 0 || i > m || j < 0 || j > n || st < 0 || st > 2 ||
            __hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            // a failed fill (the call raises): no walk
            if (tid == 0) __hip_atomic_fetch_or(A.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            i = 0; j = 0;
        }
        // thread 0 writes the guide and the stored path (every thread walks the same moves)
        uint32_t* const oend = (A.ops && tid == 0) ? A.ops + A.off_a[p + 1] + A.off_b[p + 1] : nullptr;
        WalkOut wo;
        const int bi = i, bj = j;
        wo.begin((A.guide && tid == 0) ? A.guide + a0 + p : nullptr, oend, m, n, bi, bj);
        constexpr int WS = walk_steps<S>();   // table steps of the window
        uint8_t* sb = reinterpret_cast<uint8_t*>(win + WS * SD);   // B codes of the window's panel
        uint8_t* sa = sb + PW;                                               // A codes of rows alo ..
        uint16_t* stk = reinterpret_cast<uint16_t*>(sa + kWalkRows);          // tracked: their sets
        uint16_t* sgt = stk + kWalkRows;                                     // and gap sets (of row x + 1)
        static_assert(12 * kWalkThreads == WS * (SD / 4), "the window is 12 vectors a thread");
        static_assert(kWalkRows >= WS + kWave - 1 && kWalkRows <= kWalkThreads, "staged rows");
        // Window [lo, hi] of panel P into registers: the table steps, and per thread tid <
        // kWalkRows the staged row alo + tid (alo = max(0, lo - 63): a step t of the window is
        // row t - go + 1 for a lane go in 0 .. 63).
        WalkBuf buf;
        int ra = 0, rtk = 0, rgt = 0;
        int wpanel = -1;
        int wlo = 1, whi = 0;   // the window holds steps [wlo, whi] of panel wpanel
        int alo = 0;            // and rows alo .. of A
        bool ahead = false;     // registers hold steps [wlo - WS, wlo - 1] (clamped at 0)
        for (;;) {
            if (i <= 0 && j <= 0) break;
            if (i == 0 || j == 0) {   // borders: runs (Global / Glocal column 0) or a stop
                if (AM == kLocal) break;
                if (i == 0) {
                    if (AM == kGlocal) break;
                    gap(1, 0, j);
                    if (GM == kAffine) gap(0, 0, 1);
                    wo.row0(j);
                    j = 0;
                    break;
                }
                if constexpr (TR) {
                    for (int r = i; r >= 1; --r) gap(3, r, 1);
                } else {
                    gap(3, i, i);
                }
                if (GM == kAffine) gap(2, 1, 1);
                wo.column0(i);
                i = 0;
                break;
            }
            const int jj = j - 1;
            const int P = jj / PW, jr = jj - P * PW;
            const int go = jr / S, k = jr - go * S;
            const int t = i - 1 + go;
            if (P != wpanel || t < wlo || t > whi) {
                __syncthreads();   // every read of the old window is done
                const int nlo = wlo - WS > 0 ? wlo - WS : 0;
                if (P == wpanel && ahead && t >= nlo && t < wlo) {
                    whi = wlo - 1;   // the prefetched window: t is at most 2 steps below wlo
                    wlo = nlo;
                } else {
                    whi = t;
                    wlo = t - (WS - 1) > 0 ? t - (WS - 1) : 0;
                    walk_fetch<TR, SD>(buf, ra, rtk, rgt, A.table + A.items[lp.item + P].table_off, wlo, whi, tid, m,
                                       ap, tk, A.tr.gtrk + r0);
                    if (P != wpanel) {
                        const int c0 = P * PW, cn = n - c0 < PW ? n - c0 : PW;
                        for (int x = tid; x < cn; x += kWalkThreads) sb[x] = bp[c0 + x];
                    }
                    wpanel = P;
                }
                alo = wlo - (kWave - 1) > 0 ? wlo - (kWave - 1) : 0;
                walk_put<TR, SD>(buf, ra, rtk, rgt, win, sa, stk, sgt, wlo, whi, tid);
                __syncthreads();
                ahead = wlo > 0;
                if (ahead)
                    walk_fetch<TR, SD>(buf, ra, rtk, rgt, A.table + A.items[lp.item + P].table_off,
                                       wlo - WS > 0 ? wlo - WS : 0, wlo - 1, tid, m, ap, tk,
                                       A.tr.gtrk + r0);
            }
            // every thread reads the same word: a wave-uniform value keeps the walk's control
            // flow scalar
            const uint32_t w = (uint32_t)wave_uniform((int)win[(t - wlo) * SD + (k >> 2) * kWave + go]);
            const uint32_t cc = (w >> (8 * (k & 3))) & 0xffu;
            const int ax = i - 1 - alo;   // row i's staged entries
            const int mrow = ((TR ? (int)stk[ax] * na : 0) + (int)sa[ax]) * na + (int)sb[jr];
            // a gap move of row i: f 0 go_a, 1 ge_a (its gap set), 2 go_b, 3 ge_b (its set)
            auto gapw = [&](int f) {
                if constexpr (TR) {
                    const int sset = f < 2 ? (int)sgt[ax] : (int)stk[ax];
                    if (tid == 0)
                        __hip_atomic_fetch_add(&gcnt[4 * sset + f], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    gap(f, i, 1);
                }
            };
            if constexpr (GM == kAffine) {
                if (st == 0) {
                    if (AM == kLocal && (cc & 3u) == 3u) break;
                    const uint32_t kk = cc & 3u;
                    if (tid == 0)
                        __hip_atomic_fetch_add(&hist[mrow], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    wo.move(0, i, j);
                    st = kk == 1u ? 1 : (kk == 2u ? 2 : 0);
                    --i; --j;
                } else if (st == 1) {
                    const uint32_t kk = (cc >> 2) & 3u;
                    const int pv = kk == 1u ? 1 : (kk == 2u ? 2 : 0);
                    gapw(3);
                    if (pv != 1) gapw(2);
                    wo.move(1, i, j);
                    st = pv;
                    --i;
                } else {
                    const uint32_t kk = cc >> 4;
                    const int pv = kk == 1u ? 1 : (kk == 2u ? 2 : 0);
                    gapw(1);
                    if (pv != 2) gapw(0);
                    wo.move(2, i, j);
                    st = pv;
                    --j;
                }
            } else {
                if (cc == 3u) break;
                if (cc == 0u) {
                    if (tid == 0)
                        __hip_atomic_fetch_add(&hist[mrow], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    wo.move(0, i, j);
                    --i; --j;
                } else if (cc == 1u) { gapw(3); wo.move(1, i, j); --i; }
                else { gapw(1); wo.move(2, i, j); --j; }
            }
        }
        if constexpr (AM == kLocal) wo.end_local(i, m, n, bi, bj);
        if (oend) {
            A.ops_n[p] = wo.finish(oend);
            A.ends[2 * p] = bi;
            A.ends[2 * p + 1] = bj;
        }
        __syncthreads();
        double* gr = A.grads + (int64_t)p * gs;
        for (int k = tid; k < knn; k += kWalkThreads) gr[k] = (double)hist[k];
        if constexpr (TR) {
            for (int k = tid; k < 4 * K; k += kWalkThreads) gr[knn + k] = 0.0 - (double)gcnt[k];
        } else if (tid == 0) {
            gr[nn + 0] = 0.0 - (double)c_goa;
            gr[nn + 1] = 0.0 - (double)c_gea;
            gr[nn + 2] = 0.0 - (double)c_gob;
            gr[nn + 3] = 0.0 - (double)c_geb;
        }
    }
}

}  // namespace nwg
