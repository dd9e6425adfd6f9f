// Hirschberg (linear-memory) traceback of Global affine hard pairs on gfx950: the recursion's
// half sweeps, joins, base cases and the replay of the joined path.
//
// Reference behaviour restated (relative to the reference's src/nwgrad/cpp/nwgrad/):
//   which pairs split (more than hb_cutoff rows)     aligner.hpp:136-141, 3295-3330
//   half sweeps, both carries                         aligner.hpp:4150-4255 (hb_fwd / hb_rev)
//   base case: the Pointers fill of a block            aligner.hpp:4313-4400 (hb_base)
//   the join (first strict maximum, span refund)       aligner.hpp:4456-4518 (hb_solve)
//   the score replayed along the path, in T            aligner.hpp:4525-4560
//   counts and guide of the held path                  aligner.hpp:5191-5207, 5233-5250
//
// Mapping.  A block of the recursion is rows (i0, i1] x columns (j0, j1] of one pair.  A half
// sweep is a score-only fill of rows 0..H of a block, keeping only its last row: the
// forward half reads A and B in order, the reverse half reads both backwards, which is the
// reference's mirrored sweep (hb_rev computes the same recurrences over the reversed
// suffixes), so both run the same kernel with a signed stride.  Row 0 is computed as a DP
// row whose upper neighbours are all -inf: that gives the reference's row-0 chain (its Y
// carried left to right, one rounding per column) with the cell code of every other row.
// Column 0 is lane 0's own chain down the rows (in_x: the first row continues the run).
//
// A sweep is cut into column panels of 64 lanes x S columns; lanes run skewed as in the
// other fills (lane g computes row t - g at step t).  Panels are pipelined across waves:
// panel k's last lane publishes its right column every RB rows, and panel k+1's wave waits
// for each block of rows (the long-pair protocol: items are queued in panel order, so every
// wait is on an item a running wave took earlier; each wait is bounded in real time).
// With the prefix-max carry (hirschberg_pmax) a lane also passes the running maximum P of
// its row to the right, and Y is P minus the ramp (c - 1) * ge_a.
//
// The base case fills a block of at most hb_cutoff rows with 1-byte codes (the reference's
// M | X<<2 | Y<<4) in the hard fill's layout, one wave per block, panels in sequence, then
// lane 0 walks the codes back from (H, W) -- X if out_x, else the best of the three states
// -- and writes the block's moves backward into its region.  The replay walks a pair's
// blocks in row order, scoring the path in T one move at a time (the reference's own
// order), counting matches and gaps, and writing the guide and, when kept, the stored path.
#pragma once

#include "hard_kernels.hpp"

namespace nwg {

// Columns per lane: the sweeps 16 (1,024-column panels: half the panels of 8, so half the
// pipeline's hand-offs along a level-0 sweep, and about one wave a SIMD, each step's work
// twice as long); the base cases 8 (their codes and walk; 16 would spill in float64).
constexpr int kHbSweepS = 16;
constexpr int kHbSweepPanel = kHbSweepS * kWave;   // columns per sweep panel
constexpr int kHbS = 8;
constexpr int kHbPanel = kHbS * kWave;             // columns per base-case panel
constexpr int kHbRB = 64;                  // rows a panel publishes at a time

// One panel of one half sweep.
struct HbItem {
    int32_t node;       // the level's node
    int32_t panel;
    int32_t H, W;       // the half's rows, the node's columns
    int32_t flag;       // in_x (forward half) / out_x (reverse half)
    int32_t a_step, b_step;
    int32_t pad;
    int64_t a_base, b_base;   // code index of row 1 / column 1
    int64_t bnd_in, bnd_out;  // T elements: the previous / this panel's right column (NB per row), -1: none
    int64_t out;              // T elements: the half's last row, 3 per column (M, X, Y), W + 1 columns
};

// One node's join: the two halves' last rows.
struct HbJoin {
    int64_t fwd, rev;   // T elements, 3 * (W + 1) each
    int32_t W, pad;
};

// One base-case block.
struct HbBase {
    int32_t H, W;
    int32_t in_x, out_x;
    int64_t a_base, b_base;   // code index of row 1 / column 1
    int64_t table;            // bytes: the block's codes, panels x (H + 64) x 512
    int64_t bnd;              // T elements: right columns of all but the last panel, 3 per row
    int64_t hops;             // bytes: the block's moves region, H + W bytes, filled from its end
};

// One split pair of the replay.
struct HbPair {
    int32_t pair;             // batch index
    int32_t first, count;     // its base blocks, in row order
    int32_t m, n;
    int64_t a0, b0;           // code index of A[0], B[0]
};

template <typename T>
struct HbSweepArgs {
    const uint8_t* codes_a;
    const uint8_t* codes_b;
    const T* sub;
    int32_t na;
    T go_a, ge_a, go_b, ge_b;
    const HbItem* items;
    int32_t n_items;
    int32_t* queue;
    int32_t* progress;   // per item: rows of its right column published
    T* bnd;
    T* out;
    int32_t* err;        // bit 0: a wait passed its bound
    uint64_t wait_ticks;
};

template <typename T>
struct HbJoinArgs {
    const HbJoin* nodes;
    const T* out;
    T go_b;
    int32_t* res;        // per node: cstar, span
};

template <typename T>
struct HbBaseArgs {
    const uint8_t* codes_a;
    const uint8_t* codes_b;
    const T* sub;
    int32_t na;
    T go_a, ge_a, go_b, ge_b;
    const HbBase* blocks;
    int32_t n_blocks;
    uint8_t* table;
    T* bnd;
    uint8_t* hops;
    int32_t* hcount;     // per block: moves written
};

template <typename T>
struct HbReplayArgs {
    const uint8_t* codes_a;
    const uint8_t* codes_b;
    const T* sub;
    int32_t na;
    T go_a, ge_a, go_b, ge_b;
    const HbPair* pairs;
    const HbBase* blocks;
    const uint8_t* hops;
    const int32_t* hcount;
    double* scores;
    double* grads;       // rows of na*na + 4
    int32_t* guide;      // pair p's m + 1 entries at a0 + p (null: none)
    uint32_t* ops;       // stored paths (null: none): pair p's region ends at a0 + b0 + m + n
    int32_t* ops_n;
    int32_t* ends;
};

template <typename T>
__device__ __forceinline__ T max3(T a, T b, T c) { return dmax(dmax(a, b), c); }

// One row of a lane's strip, cells c0 + 1 .. c0 + S, updated in place: the reference's affine
// Global cell (aligner.hpp:4174-4176, 4366-4374) in its order, as strip_row computes it, with the
// row's substitution values sv already loaded.  (lM, lX, lY) is the left neighbour, (dM, dX, dY)
// the diagonal; HARD: the cells' codes M | X<<2 | Y<<4 (the >= chains of the reference's pick)
// into cw, four to a dword.
//
// Score only, X and Y take a shorter form of the same values: x -> (x - go) - ge and x -> x - ge
// are non-decreasing in x (every rounding is), so the largest of the three candidates is
// ((max(uM, uY) - go_b) max uX) - ge_b: a maximum of images under a non-decreasing map is the
// image of the maximum.  4 operations a state instead of 6; the values are the reference's
// (at most the sign of a zero differs, which no comparison sees).
template <typename T, int S, bool HARD>
__device__ __forceinline__ void hb_strip(T (&M)[S], T (&X)[S], T (&Y)[S], T lM, T lX, T lY, T dM, T dX, T dY,
                                         const T (&sv)[S], T go_a, T ge_a, T go_b, T ge_b, uint32_t (&cw)[S / 4]) {
#pragma unroll
    for (int k = 0; k < S / 4; ++k) cw[k] = 0u;
    if constexpr (!HARD) {
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const T uM = M[k], uX = X[k], uY = Y[k];
            const T mv = dmax(dM, dmax(dX, dY)) + sv[k];
            const T xv = dmax(dmax(uM, uY) - go_b, uX) - ge_b;
            const T yv = dmax(dmax(lM, lX) - go_a, lY) - ge_a;
            M[k] = mv; X[k] = xv; Y[k] = yv;
            dM = uM; dX = uX; dY = uY;
            lM = mv; lX = xv; lY = yv;
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const T uM = M[k], uX = X[k], uY = Y[k];
        const T tXY = dmax(dX, dY);
        const T mv = dmax(dM, tXY) + sv[k];
        const T ux = (uM - go_b) - ge_b, vx = uX - ge_b, wx = (uY - go_b) - ge_b;
        const T tx = dmax(vx, wx);
        const T xv = dmax(ux, tx);
        const T uy = (lM - go_a) - ge_a, vy = (lX - go_a) - ge_a, wy = lY - ge_a;
        const T ty = dmax(vy, wy);
        const T yv = dmax(uy, ty);
        if constexpr (HARD) {
            const uint32_t km = (dM >= tXY) ? 0u : ((dX >= dY) ? 1u : 2u);
            const uint32_t kx = (ux >= tx) ? 0u : ((vx >= wx) ? 1u : 2u);
            const uint32_t ky = (uy >= ty) ? 0u : ((vy >= wy) ? 1u : 2u);
            cw[k >> 2] |= (km | (kx << 2) | (ky << 4)) << (8 * (k & 3));
        }
        M[k] = mv; X[k] = xv; Y[k] = yv;
        dM = uM; dX = uX; dY = uY;
        lM = mv; lX = xv; lY = yv;
    }
}

// The same row under the prefix-max carry (aligner.hpp:4176-4190): M and X as the exact cell,
// Y from the running maximum P of g[k] + k * ge_a along the row.  c0 is the block column of
// the strip's first cell minus one.  Score only.
template <typename T, int S>
__device__ __forceinline__ void hb_strip_pmax(T (&M)[S], T (&X)[S], T (&Y)[S], T lM, T lX, T& P, T dM, T dX,
                                              T dY, const T (&sv)[S], int c0, T go_a, T ge_a, T go_b, T ge_b) {
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const T uM = M[k], uX = X[k], uY = Y[k];
        const T mv = dmax(dM, dmax(dX, dY)) + sv[k];
        const T xv = dmax(dmax(uM, uY) - go_b, uX) - ge_b;   // as hb_strip's score-only form
        const T g = (dmax(lM, lX) - go_a) - ge_a;
        const T ramp = (T)(c0 + k) * ge_a;   // rmp[c - 1], c = c0 + k + 1
        P = dmax(g + ramp, P);
        const T yv = P - ramp;
        M[k] = mv; X[k] = xv; Y[k] = yv;
        dM = uM; dX = uX; dY = uY;
        lM = mv; lX = xv;
    }
}

// Fill of one panel of a block, rows 0..H: the sweep's (OUT: the last row into `out`; WAIT:
// the left column from another wave, published by `prog_in`) and the base case's (HARD: codes
// into `tbl`; the left column written earlier by this same wave).  bin / bout hold NB values
// per row (M, X, Y, and P under PMAX); bout == null for the last panel.
template <typename T, int S, bool PMAX, bool HARD, bool WAIT>
__device__ void hb_panel(const T* __restrict__ ssub, int na, const uint8_t* __restrict__ codes_a,
                         const uint8_t* __restrict__ codes_b, int64_t a_base, int a_step, int64_t b_base, int b_step,
                         int H, int W, int panel, int flag, T go_a, T ge_a, T go_b, T ge_b, const T* bin, T* bout,
                         const int32_t* prog_in, int32_t* prog_out, int32_t* err, uint64_t ticks, T* out,
                         uint32_t* tbl, T* fin, uint8_t* acodes, int lane) {
    constexpr int NB = PMAX ? 4 : 3;
    const T NI = ninf<T>();
    const int g = lane;
    const int jl = panel * (kWave * S) + g * S;   // the lane's columns are jl + 1 .. jl + S
    const int ncols = W - jl < 0 ? 0 : (W - jl > S ? S : W - jl);
    const bool first = panel == 0, last = bout == nullptr;
    int coff[S];
#pragma unroll
    for (int k = 0; k < S; ++k) coff[k] = (k < ncols) ? (int)codes_b[b_base + (int64_t)b_step * (jl + k)] : 0;
    T M[S], X[S], Y[S];
#pragma unroll
    for (int k = 0; k < S; ++k) { M[k] = NI; X[k] = NI; Y[k] = NI; }
    T pM = NI, pX = NI, pY = NI;              // the diagonal: the left column of the previous row
    T oM = NI, oX = NI, oY = NI, oP = NI;     // the lane's last column, passed right
    T bM = NI, bX = NI, bY = NI, bP = NI;     // lane r: the left column of row t0 + r of the block
    const int rows = H + 1;
    const int steps = rows + kWave - 1;
    // The codes of A pass through a 256-row ring in LDS (acodes): at step t0 = 64 k the wave
    // stores rows t0 + 1 .. t0 + 64, loaded one block earlier, and loads the next block; lane g
    // then reads row t - g from LDS.  A step never waits on a global load.
    auto arow = [&](int rr) -> int {
        return (rr >= 1 && rr <= H) ? (int)codes_a[a_base + (int64_t)a_step * (rr - 1)] : 0;
    };
    int anext = arow(lane + 1);
    // The substitution values of the lane's next row, loaded one step ahead (the row's code
    // from the ring, then its S values from the matrix in LDS), so no step waits on LDS.
    T sv[S];
#pragma unroll
    for (int k = 0; k < S; ++k) sv[k] = ssub[coff[k]];
    for (int t = 0; t < steps; ++t) {
        const int r = t - g;
        if (t % kWave == 0) {
            acodes[(t + 1 + lane) & 255] = (uint8_t)anext;
            anext = arow(t + kWave + 1 + lane);
            wave_sync();
        }
        // (the block's loads are waited for here, once, by settle(): a value still pending from
        // a load issued in an earlier trip of the loop makes the compiler wait for every vector
        // memory operation outstanding at each of its uses -- this wave's boundary and code
        // stores of the step before included, a store's full round trip every step)
        if (!first && t < rows && t % kHbRB == 0) {
            const int need = t + kHbRB < rows ? t + kHbRB : rows;
            if constexpr (WAIT) (void)bounded_wait(prog_in, need, err, ticks, lane, 1);
            const int rr = t + lane;
            if (lane < kHbRB && rr < rows) {
                bM = bin[(int64_t)NB * rr];
                bX = bin[(int64_t)NB * rr + 1];
                bY = bin[(int64_t)NB * rr + 2];
                if constexpr (PMAX) bP = bin[(int64_t)NB * rr + 3];
            }
            settle(bM, bX, bY, bP);
        }
        T iM = shift_up(oM), iX = shift_up(oX), iY = shift_up(oY);
        T iP = PMAX ? shift_up(oP) : NI;
        if (!first) {
            const int sl = t % kHbRB;
            const T lM = bcast(bM, sl), lX = bcast(bX, sl), lY = bcast(bY, sl);
            T lP = NI;
            if constexpr (PMAX) lP = bcast(bP, sl);
            if (g == 0) { iM = lM; iX = lX; iY = lY; iP = lP; }
        } else if (g == 0) {   // column 0 of row t (aligner.hpp:4161-4165, 4169-4171)
            if (t == 0) {
                iM = flag ? NI : T(0);
                iX = NI;
            } else {
                iM = NI;
                iX = max3((pM - go_b) - ge_b, pX - ge_b, (pY - go_b) - ge_b);
                if (flag && t == 1) iX = T(0) - ge_b;
            }
            iY = NI;
            iP = NI;   // the carry starts at cY[0] - ge_a = -inf
        }
        // Every lane computes every step, which keeps the loop free of divergent branches: a lane
        // before its row 0 (r < 0) holds -inf in every state and receives -inf from its left
        // neighbour (which has not started either), so it stays at -inf; after its row H, or
        // past the block's last column, what it computes is never stored nor passed on to a
        // lane that stores.
        {
            T P = iP;
            uint32_t cw[S / 4];
            if (PMAX && r >= 1) {
                hb_strip_pmax<T, S>(M, X, Y, iM, iX, P, pM, pX, pY, sv, jl, go_a, ge_a, go_b, ge_b);
            } else {
                hb_strip<T, S, HARD>(M, X, Y, iM, iX, iY, pM, pX, pY, sv, go_a, ge_a, go_b, ge_b, cw);
                P = NI;
            }
            oM = M[S - 1]; oX = X[S - 1]; oY = Y[S - 1];
            if constexpr (PMAX) oP = P;
          if (r >= 0 && r <= H && ncols > 0) {
            if constexpr (HARD) {
#pragma unroll
                for (int k = 0; k < S / 4; ++k) tbl[((int64_t)t * (S / 4) + k) * kWave + lane] = cw[k];
            }
            if (!last && g == kWave - 1) {
                bout[(int64_t)NB * r] = M[S - 1];
                bout[(int64_t)NB * r + 1] = X[S - 1];
                bout[(int64_t)NB * r + 2] = Y[S - 1];
                if constexpr (PMAX) bout[(int64_t)NB * r + 3] = P;
                if constexpr (WAIT) {
                    if ((r + 1) % kHbRB == 0 || r == H)
                        __hip_atomic_store(prog_out, r + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            if (r == H) {
                if (out) {
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        if (k < ncols) {
                            T* o = out + 3 * (int64_t)(jl + k + 1);
                            o[0] = M[k]; o[1] = X[k]; o[2] = Y[k];
                        }
                    }
                }
                if (fin && jl + ncols == W) {   // the block's last cell (H, W)
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        if (k == ncols - 1) { fin[0] = M[k]; fin[1] = X[k]; fin[2] = Y[k]; }
                    }
                }
            }
          }
        }
        if (first && g == 0 && t == H) {   // column 0 of the last row
            if (out) { out[0] = iM; out[1] = iX; out[2] = iY; }
            if (fin && W == 0) { fin[0] = iM; fin[1] = iX; fin[2] = iY; }
        }
        pM = iM; pX = iX; pY = iY;
        {   // the next step's row, r + 1: stored in the ring by now (rows up to t + 1 are)
            const int rn = r + 1;
            const int ac = (rn >= 1 && rn <= H) ? (int)acodes[rn & 255] : 0;
            const T* __restrict__ srow = ssub + ac * na;
#pragma unroll
            for (int k = 0; k < S; ++k) sv[k] = srow[coff[k]];
        }
    }
}

// Half sweeps of one level: one wave per block, pulling panel items until the queue is
// drained; every wave reaches the exit test once the counter passes n_items.
template <typename T, bool PMAX>
__global__ __launch_bounds__(64) void hb_sweep_kernel(HbSweepArgs<T> A) {
    extern __shared__ __align__(16) unsigned char lds[];
    uint8_t* acodes = lds;   // 256 bytes: the ring of A's codes
    T* ssub = reinterpret_cast<T*>(lds + 256);
    const int lane = threadIdx.x;
    for (int k = lane; k < A.na * A.na; k += kWave) ssub[k] = A.sub[k];
    __syncthreads();
    for (;;) {
        const int it = wave_fetch(A.queue);
        if (it >= A.n_items) break;
        const HbItem item = A.items[it];
        const bool first = item.bnd_in < 0, last = item.bnd_out < 0;
        hb_panel<T, kHbSweepS, PMAX, false, true>(
            ssub, A.na, A.codes_a, A.codes_b, item.a_base, item.a_step, item.b_base, item.b_step, item.H, item.W,
            item.panel, item.flag, A.go_a, A.ge_a, A.go_b, A.ge_b, first ? nullptr : A.bnd + item.bnd_in,
            last ? nullptr : A.bnd + item.bnd_out, A.progress + (first ? it : it - 1), A.progress + it, A.err,
            A.wait_ticks, A.out + item.out, nullptr, nullptr, acodes, lane);
    }
}

// The join of one node (aligner.hpp:4481-4512): the first column c of the strict maximum of
// max(plain, joined), plain = max(F[c]) + max(R[W - c]), joined = FX[c] + RX[W - c] + go_b
// (an X run across the cut pays one open, not two); span where joined is the larger.  A
// 256-thread block per node: each thread keeps its first strict maximum over its columns in
// order, and the block reduces (larger value, then smaller column).
template <typename T>
__global__ __launch_bounds__(256) void hb_join_kernel(HbJoinArgs<T> A) {
    __shared__ T sv[256];
    __shared__ int sc[256], ss[256];
    const HbJoin nd = A.nodes[blockIdx.x];
    const T* F = A.out + nd.fwd;
    const T* R = A.out + nd.rev;
    const T NI = ninf<T>();
    T best = NI;
    int bc = 0x7fffffff, bs = 0;
    for (int c = threadIdx.x; c <= nd.W; c += blockDim.x) {
        const int d = nd.W - c;
        const T fM = F[3 * (int64_t)c], fX = F[3 * (int64_t)c + 1], fY = F[3 * (int64_t)c + 2];
        const T rM = R[3 * (int64_t)d], rX = R[3 * (int64_t)d + 1], rY = R[3 * (int64_t)d + 2];
        const T plain = max3(fM, fX, fY) + max3(rM, rX, rY);
        T v = plain;
        int sp = 0;
        if (fX > NI && rX > NI) {
            const T joined = (fX + rX) + A.go_b;
            if (joined > plain) { v = joined; sp = 1; }
        }
        if (v > best) { best = v; bc = c; bs = sp; }
    }
    sv[threadIdx.x] = best;
    sc[threadIdx.x] = bc;
    ss[threadIdx.x] = bs;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            const T ov = sv[threadIdx.x + off];
            const int oc = sc[threadIdx.x + off];
            if (ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oc < sc[threadIdx.x])) {
                sv[threadIdx.x] = ov;
                sc[threadIdx.x] = oc;
                ss[threadIdx.x] = ss[threadIdx.x + off];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const bool any = sv[0] > NI;
        A.res[2 * blockIdx.x] = any ? sc[0] : 0;
        A.res[2 * blockIdx.x + 1] = any ? ss[0] : 0;
    }
}

// Byte of cell (r, c), r, c >= 1, in a base block's code table: panels of 512 columns, each
// (H + 64) steps of 512 bytes; lane l of step t holds its 8 cells as two dwords at
// ((t * 2 + k / 4) * 64 + l) * 4 + k % 4.
__device__ __forceinline__ int64_t hb_code_at(int H, int r, int c) {
    const int q = c - 1;
    const int pp = q / kHbPanel, w = q % kHbPanel, l = w / kHbS, k = w % kHbS;
    const int64_t panel_bytes = (int64_t)(H + kWave) * kHbPanel;
    return pp * panel_bytes + (((int64_t)(r + l) * (kHbS / 4) + (k >> 2)) * kWave + l) * 4 + (k & 3);
}

// Base cases: one wave per block fills its panels in order, then lane 0 walks the codes back
// (aligner.hpp:4384-4400) and writes the block's moves backward from its region's end.
template <typename T>
__global__ __launch_bounds__(64) void hb_base_kernel(HbBaseArgs<T> A) {
    extern __shared__ __align__(16) unsigned char lds[];
    uint8_t* acodes = lds;   // 256 bytes: the ring of A's codes
    T* ssub = reinterpret_cast<T*>(lds + 256);
    T* fin = ssub + A.na * A.na;
    const int lane = threadIdx.x;
    for (int k = lane; k < A.na * A.na; k += kWave) ssub[k] = A.sub[k];
    if (lane < 3) fin[lane] = ninf<T>();
    __syncthreads();
    const HbBase B = A.blocks[blockIdx.x];
    const int P = B.W > 0 ? (B.W + kHbPanel - 1) / kHbPanel : 0;
    const int64_t panel_dwords = (int64_t)(B.H + kWave) * (kHbPanel / 4);
    T* bnd = A.bnd + B.bnd;
    uint32_t* tb = reinterpret_cast<uint32_t*>(A.table + B.table);
    for (int pp = 0; pp < (P > 0 ? P : 1); ++pp) {
        const T* bin = pp == 0 ? nullptr : bnd + 3 * (int64_t)(pp - 1) * (B.H + 1);
        T* bout = pp + 1 < P ? bnd + 3 * (int64_t)pp * (B.H + 1) : nullptr;
        hb_panel<T, kHbS, false, true, false>(ssub, A.na, A.codes_a, A.codes_b, B.a_base, 1, B.b_base, 1, B.H, B.W,
                                              pp, B.in_x, A.go_a, A.ge_a, A.go_b, A.ge_b, bin, bout, nullptr,
                                              nullptr, nullptr, 0, nullptr, tb + pp * panel_dwords, fin, acodes,
                                              lane);
        // this wave's right column (and, at the end, its codes) are read back by this wave
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    if (lane != 0) return;
    const uint8_t* cb = A.table + B.table;
    int r = B.H, c = B.W;
    int tbl = B.out_x ? 1 : state_of(fin[0], fin[1], fin[2]);
    uint8_t* const end = A.hops + B.hops + B.H + B.W;
    uint8_t* hp = end;
    while (r > 0 && c > 0) {
        const unsigned p = cb[hb_code_at(B.H, r, c)];
        unsigned code;
        if (tbl == 0) { code = p & 3u; *--hp = 0; --r; --c; }
        else if (tbl == 1) { code = (p >> 2) & 3u; *--hp = 1; --r; }
        else { code = p >> 4; *--hp = 2; --c; }
        tbl = code == 1u ? 1 : (code == 2u ? 2 : 0);
    }
    // Column 0 holds a finite X only, row 0 a finite Y only (both -inf elsewhere), so the
    // path's remainder is a straight run to the origin.
    for (; r > 0; --r) *--hp = 1;
    for (; c > 0; --c) *--hp = 2;
    A.hcount[blockIdx.x] = (int32_t)(end - hp);
}

// Inclusive prefix sum over the wave.
__device__ __forceinline__ int wave_incl_sum(int v, int lane) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int o = __shfl_up(v, off, kWave);
        if (lane >= off) v += o;
    }
    return v;
}

// LDS of the replay: the matrix (T), then match counts (uint32).
template <typename T>
__host__ __device__ constexpr int hb_replay_lds_bytes(int na) {
    return ((na * na * (int)sizeof(T) + 15) & ~15) + ((na * na * 4 + 15) & ~15);
}

// A lane's value, read by the whole wave (q a compile-time lane).
template <int Q>
__device__ __forceinline__ float lane_value(float v) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), Q));
}
template <int Q>
__device__ __forceinline__ double lane_value(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), Q);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), Q);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// s = (s + t1[q]) + t2[q] for the wave's lanes q = Q .. 63, in lane order.
template <typename T, int Q>
__device__ __forceinline__ T add_in_lane_order(T s, T t1, T t2) {
    if constexpr (Q == kWave) {
        return s;
    } else {
        s = (s + lane_value<Q>(t1)) + lane_value<Q>(t2);
        return add_in_lane_order<T, Q + 1>(s, t1, t2);
    }
}

// Global memory handed between lanes of one wave.
__device__ __forceinline__ void wg_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// The replay of one split pair: its blocks' moves in order, 64 at a time.  Positions come
// from prefix sums; every lane counts its move and forms its terms of the score -- a match's
// substitution value, or a gap's -go (at a run's first move, else 0) and -ge; the wave then
// adds them in path order, (s + t1) + t2 move by move, which is the reference's own sequence
// of roundings (aligner.hpp:4530-4553): s never becomes -0 (it starts at +0, and x + -x is
// +0), so a zero term changes nothing.  A stored path is run-length encoded from the runs'
// first moves (their type and position), then converted to lengths at the region's end.
template <typename T>
__global__ __launch_bounds__(64) void hb_replay_kernel(HbReplayArgs<T> A) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int nn = A.na * A.na;
    T* ssub = reinterpret_cast<T*>(lds);
    uint32_t* hist = reinterpret_cast<uint32_t*>(lds + ((nn * (int)sizeof(T) + 15) & ~15));
    const int lane = threadIdx.x;
    for (int k = lane; k < nn; k += kWave) { ssub[k] = A.sub[k]; hist[k] = 0u; }
    __syncthreads();
    const HbPair pr = A.pairs[blockIdx.x];
    const int p = pr.pair;
    int32_t* gj = A.guide ? A.guide + pr.a0 + p : nullptr;
    uint32_t* ops = A.ops ? A.ops + pr.a0 + pr.b0 : nullptr;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    int i = 0, j = 0, prev = 3, done = 0, nruns = 0;
    int open_a = 0, ext_a = 0, open_b = 0, ext_b = 0;
    T s = T(0);
    for (int bk = 0; bk < pr.count; ++bk) {
        const HbBase B = A.blocks[pr.first + bk];
        const int cnt = A.hcount[pr.first + bk];
        const uint8_t* mv = A.hops + B.hops + B.H + B.W - cnt;
        for (int base = 0; base < cnt; base += kWave) {
            const int k = base + lane;
            const bool valid = k < cnt;
            const int op = valid ? (int)mv[k] : 3;
            const int ii = i + wave_incl_sum(valid && op != 2 ? 1 : 0, lane);
            const int jj = j + wave_incl_sum(valid && op != 1 ? 1 : 0, lane);
            int pop = __shfl_up(op, 1, kWave);
            if (lane == 0) pop = prev;
            const bool opn = valid && op != pop;
            T t1 = T(0), t2 = T(0);
            if (valid) {
                if (op == 0) {
                    const int ca = A.codes_a[pr.a0 + ii - 1], cbv = A.codes_b[pr.b0 + jj - 1];
                    t1 = ssub[ca * A.na + cbv];
                    atomicAdd(&hist[ca * A.na + cbv], 1u);
                    if (gj) gj[ii] = jj;
                } else if (op == 1) {
                    ++ext_b;
                    open_b += opn;
                    t1 = opn ? -A.go_b : T(0);
                    t2 = -A.ge_b;
                    if (gj) gj[ii] = jj;
                } else {
                    ++ext_a;
                    open_a += opn;
                    t1 = opn ? -A.go_a : T(0);
                    t2 = -A.ge_a;
                }
            }
            s = add_in_lane_order<T, 0>(s, t1, t2);
            if (ops) {
                const uint64_t starts = __ballot(opn);
                if (opn) ops[nruns + __popcll(starts & below)] = ((uint32_t)op << 30) | (uint32_t)(done + k);
                nruns += __popcll(starts);
            }
            const int lastl = (cnt - base < kWave ? cnt - base : kWave) - 1;
            i = __shfl(ii, kWave - 1, kWave);
            j = __shfl(jj, kWave - 1, kWave);
            prev = __shfl(op, lastl, kWave);
        }
        done += cnt;
    }
    // counts: matches from LDS, gaps summed over the wave (aligner.hpp:5191-5207)
    for (int off = kWave / 2; off > 0; off >>= 1) {
        open_a += __shfl_down(open_a, off, kWave);
        ext_a += __shfl_down(ext_a, off, kWave);
        open_b += __shfl_down(open_b, off, kWave);
        ext_b += __shfl_down(ext_b, off, kWave);
    }
    __syncthreads();
    double* g = A.grads + (int64_t)p * (nn + 4);
    for (int k = lane; k < nn; k += kWave) g[k] = (double)hist[k];
    if (lane == 0) {
        g[nn + 0] = -(double)open_a;
        g[nn + 1] = -(double)ext_a;
        g[nn + 2] = -(double)open_b;
        g[nn + 3] = -(double)ext_b;
        A.scores[p] = (double)s;
        if (gj) { gj[0] = 0; gj[pr.m] = pr.n; }
        if (ops) {
            A.ops_n[p] = nruns;
            A.ends[2 * p] = pr.m;
            A.ends[2 * p + 1] = pr.n;
        }
    }
    if (ops) {
        // run r starts at move pos[r] (low 30 bits) and ends where run r + 1 starts (the last
        // at the path's end); its op (type << 30 | length) goes to the region's end, which is
        // at or above its record, so chunks are converted from the top down, each chunk's
        // records read before it is written, the next chunk's first position carried down
        wg_sync();
        const int cap = pr.m + pr.n;
        uint32_t carry = (uint32_t)done;
        for (int hi = nruns; hi > 0; hi -= kWave) {
            const int r = hi - kWave + lane;
            const uint32_t v = r >= 0 ? ops[r] : 0u;
            uint32_t nx = (uint32_t)__shfl_down((int)v, 1, kWave);
            if (lane == kWave - 1) nx = carry;
            carry = (uint32_t)__shfl((int)v, 0, kWave);
            wg_sync();
            if (r >= 0) ops[cap - nruns + r] = (v & 0xc0000000u) | ((nx & 0x3fffffffu) - (v & 0x3fffffffu));
            wg_sync();
        }
    }
}

}  // namespace nwg
