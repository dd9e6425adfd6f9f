// Hard (Viterbi) and score-only fills on gfx950: the lane-group inter-pair fill.
//
// Reference behaviour restated (relative to the reference's src/nwgrad/cpp/nwgrad/):
//   affine cell + codes         aligner.hpp:3955-4071 (and inter_kernel_impl.inl:790-801)
//   linear cell + codes         aligner.hpp:3756-3865 (and inter_kernel_impl.inl:689-695)
//   local end (row-major first) aligner.hpp:5975-5984
//   glocal end (row m)          aligner.hpp:6121-6134
//   walks + counts              aligner.hpp:2986-3056, 2826-2855, 3177-3197, 3886-3925
//
// Mapping onto a wavefront.  A tile holds 64/G pairs; pair slot q owns lanes
// [q*G, q*G + G).  Lane g of the slot owns the S columns j = g*S+1 .. g*S+S of its pair
// and keeps that strip of the current DP row in VGPRs (M, X, Y -- or H for linear gaps),
// updating it in place row by row.  Lanes of a slot run skewed: lane g computes row
// i = t - g + 1 at step t, and receives the value of column g*S of that row (its left
// neighbour) from lane g-1 through a one-lane shift at the top of the step; the value of
// the previous step is the diagonal.  Every lane executes the reference's per-cell
// operations in the reference's order, so every value and every predecessor code is the
// reference's, bit for bit (up to the sign of a +-0 tie, see dmax).
//
// Hard mode stores one byte of predecessor codes per cell (affine M | X<<2 | Y<<4,
// linear 0 diag / 1 up / 2 left / 3 stop, the reference's encoding) into a per-wave
// slot of HBM scratch, four cells to a dword (code_word gives the layout: at 8 columns a lane
// a lane's two dwords of a step are contiguous, one vector store a lane; at 16 the 64 lanes'
// k-th dwords of a step are, four stores a lane -- each the faster of the two at its width).  After the fill, lane 0 of each slot walks its pair's codes
// back from the end cell, counting matches into a per-slot LDS histogram and gap moves
// in registers, and the wave writes the per-pair gradient rows.
#pragma once

#include "common.hpp"

namespace nwg {

// Tracks: K parameter sets chosen per DP row (aligner.hpp:1800-1826).  Row i (1..m) of pair p
// takes its matrix and its gaps in B from set trk[off_a[p] + i - 1] and its gaps in A from set
// gtrk[off_a[p] + p + i] (row 0 included); bx0[off_a[p] + p + i] is its column-0 border, the
// reference's run-wise sum (setup_tracks_, aligner.hpp:1981-2007), formed on the host.
template <typename T>
struct TrackArgs {
    const uint16_t* trk;
    const uint16_t* gtrk;
    const T* bx0;
    const T* gaps;    // K x 4: go_a, ge_a, go_b, ge_b, each static_cast<T>(double)
    int32_t K;
};

template <typename T>
struct HardArgs {
    const uint8_t* codes_a;
    const uint8_t* codes_b;
    const int64_t* off_a;
    const int64_t* off_b;
    const int32_t* order;     // pair indices, tile-major
    const TileDesc* tiles;
    int32_t n_tiles;
    int32_t tile_base;        // plain launch: block b fills tile tile_base + b (below n_tiles)
    int32_t* queue;           // persistent launch: next-tile counter (nullptr: plain)
    const T* sub;             // K * n_alpha * n_alpha, the matrices cast to T (K = 1 untracked)
    int32_t na;
    int32_t hist_slots;       // LDS histogram slots per wave
    T go_a, ge_a, go_b, ge_b;
    TrackArgs<T> tr;          // tracked batches (the TR kernels) only
    double* scores;           // per pair
    double* grads;            // per pair, K * (na*na + 4) (hard only)
    uint32_t* table;          // pointer-code scratch (hard only)
    int64_t slot_dwords;      // dwords per table slot
    // hard only, each may be null: guides (pair p's m + 1 entries at off_a[p] + p), stored paths
    // (pair p's ops end at off_a[p+1] + off_b[p+1], m + n entries of room), their op counts and
    // end cells (i, j)
    int32_t* guide;
    uint32_t* ops;
    int32_t* ops_n;
    int32_t* ends;
    int32_t band;             // the BAND kernels: the half-width around each pair's guide (> 0)
};

// The guide band of banded re-alignment (the reference's GuideBanded fill, aligner.hpp:2644-2688):
// row i (1..m) computes columns [band_lo, band_hi]; Global's row m reaches n.  gj is the pair's
// guide (m + 1 entries, non-decreasing in [0, n]) and bw the band clamped to n.  Column 0 of row
// i >= 1 is in the band (its Global / Glocal border value, else -inf) when gj[i] <= bw: the rows
// border_rows() counts, as a guide never decreases; row 0 up to column band_hi(0).
__device__ __forceinline__ int band_lo(const int32_t* gj, int i, int bw) {
    const int v = gj[i] - bw;
    return v > 1 ? v : 1;
}
template <int AM>
__device__ __forceinline__ int band_hi(const int32_t* gj, int i, int m, int n, int bw) {
    if (AM == kGlobal && i == m) return n;
    int hi = gj[i] + bw;
    if (i < m) {
        const int h2 = gj[i + 1] - 1 + bw;
        hi = h2 > hi ? h2 : hi;
    }
    return hi < n ? hi : n;
}

// Best-cell candidate, combined across the lanes of a slot: larger value wins, ties go to
// the smaller (i, j) -- the reference's first-in-row-major-order rule.
template <typename T>
struct Cand {
    T v;
    int i, j, st;  // st: 0 M, 1 X, 2 Y; i < 0: no candidate
};

template <typename T>
__device__ __forceinline__ bool cand_better(const Cand<T>& b, const Cand<T>& a) {
    if (b.i < 0) return false;
    if (a.i < 0) return true;
    if (b.v > a.v) return true;
    if (b.v < a.v) return false;
    return b.i < a.i || (b.i == a.i && b.j < a.j);
}

template <typename T>
__device__ __forceinline__ int state_of(T vm, T vx, T vy) {
    return (vm >= vx && vm >= vy) ? 0 : (vx >= vy ? 1 : 2);
}

// Reduce a slot's candidates into its lane 0 (G a power of two, slots aligned).
template <typename T>
__device__ __forceinline__ Cand<T> reduce_slot(Cand<T> c, int G) {
    for (int off = 1; off < G; off <<= 1) {
        Cand<T> o;
        o.v = __shfl_down(c.v, off, G);
        o.i = __shfl_down(c.i, off, G);
        o.j = __shfl_down(c.j, off, G);
        o.st = __shfl_down(c.st, off, G);
        if (cand_better(o, c)) c = o;
    }
    return c;
}

// One row of a lane's strip: cells j0+1 .. j0+S of row i, updated in place.  (lM, lX, lY) is
// the left neighbour (column j0 of row i), (dM, dX, dY) the diagonal (column j0 of row i-1).
// cw receives the predecessor codes, four cells to a dword; rowmax the row's maximum over
// the lane's real columns (Local).  Affine: aligner.hpp:4020-4047; linear: aligner.hpp:3828-3849.
// BAND (banded re-alignment): only the strip's cells k in [klo, khi] are in the row's band; the
// others are -inf in every state, as the reference's banded fill leaves them (aligner.hpp:
// 5991-6098, band_fill then the row's [jlo(i), jhi(i)]), before any maximum or row best sees them.
template <typename T, int GM, int AM, int S, bool HARD, bool BAND = false>
__device__ __forceinline__ void strip_row(T (&M)[S], T (&X)[S], T (&Y)[S], T lM, T lX, T lY, T dM, T dX, T dY,
                                          const T* __restrict__ srow, const int (&coff)[S], int ncols, T go_a,
                                          T ge_a, T go_b, T ge_b, uint32_t (&cw)[S / 4], T& rowmax, int klo = 0,
                                          int khi = S) {
#pragma unroll
    for (int k = 0; k < S / 4; ++k) cw[k] = 0u;
    rowmax = T(0);
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const T s = srow[coff[k]];
        if constexpr (GM == kAffine) {
            const T uM = M[k], uX = X[k], uY = Y[k];
            const T tXY = vmax(dX, dY);
            T mv = vmax(dM, tXY) + s;
            uint32_t km = 0;
            if constexpr (HARD) km = (dM >= tXY) ? 0u : ((dX >= dY) ? 1u : 2u);
            const T ux = (uM - go_b) - ge_b, vx = uX - ge_b, wx = (uY - go_b) - ge_b;
            const T tx = vmax(vx, wx);
            const T xv = vmax(ux, tx);
            const T uy = (lM - go_a) - ge_a, vy = (lX - go_a) - ge_a, wy = lY - ge_a;
            const T ty = vmax(vy, wy);
            const T yv = vmax(uy, ty);
            if constexpr (AM == kLocal) {
                mv = vmax(mv, T(0));
                if constexpr (HARD) km = (mv <= T(0)) ? 3u : km;
            }
            T mo = mv, xo = xv, yo = yv;
            if constexpr (BAND) {
                if (k < klo || k > khi) { mo = ninf<T>(); xo = ninf<T>(); yo = ninf<T>(); }
            }
            if constexpr (AM == kLocal) {
                if (k < ncols) rowmax = dmax(rowmax, dmax(mo, dmax(xo, yo)));
            }
            if constexpr (HARD) {
                const uint32_t kx = (ux >= tx) ? 0u : ((vx >= wx) ? 1u : 2u);
                const uint32_t ky = (uy >= ty) ? 0u : ((vy >= wy) ? 1u : 2u);
                cw[k >> 2] |= (km | (kx << 2) | (ky << 4)) << (8 * (k & 3));
            }
            M[k] = mo; X[k] = xo; Y[k] = yo;
            dM = uM; dX = uX; dY = uY;
            lM = mo; lX = xo; lY = yo;
        } else {
            const T uH = M[k];
            const T d0 = dM + s;
            const T u = uH - ge_b;
            const T tt = dmax(d0, u);
            const T l = lM - ge_a;
            T v = dmax(tt, l);
            uint32_t c = 0;
            if constexpr (HARD) c = (tt < l) ? 2u : ((d0 < u) ? 1u : 0u);
            if constexpr (AM == kLocal) {
                v = dmax(v, T(0));
                if constexpr (HARD) c = (v <= T(0)) ? 3u : c;
            }
            if constexpr (BAND) {
                if (k < klo || k > khi) v = ninf<T>();
            }
            if constexpr (AM == kLocal) {
                if (k < ncols) rowmax = dmax(rowmax, v);
            }
            if constexpr (HARD) cw[k >> 2] |= c << (8 * (k & 3));
            M[k] = v;
            dM = uH;
            lM = v;
        }
    }
}

// local_row_best_ (aligner.hpp:5975-5984): the first column holding the row maximum, when
// the row beats the lane's best so far.
template <typename T, int GM, int S>
__device__ __forceinline__ void local_update(Cand<T>& best, const T (&M)[S], const T (&X)[S], const T (&Y)[S],
                                             T rowmax, int i, int j0, int ncols) {
    if (rowmax > best.v) {
        bool found = false;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            T v3;
            if constexpr (GM == kAffine) v3 = dmax(dmax(M[k], X[k]), Y[k]);
            else v3 = M[k];
            if (!found && k < ncols && v3 == rowmax) {
                found = true;
                best.v = rowmax; best.i = i; best.j = j0 + k + 1;
                if constexpr (GM == kAffine) best.st = state_of(M[k], X[k], Y[k]);
                else best.st = 0;
            }
        }
    }
}

// The lane's end-cell candidate once row m is done (aligner.hpp:4054-4066, 3851-3860; Glocal
// 6121-6134).  col0: this lane also offers column 0 (the first lane of the pair's first panel),
// whose value at row m is col0v (0 when m = 0).
template <typename T, int GM, int AM, int S>
__device__ __forceinline__ Cand<T> end_candidate(const T (&M)[S], const T (&X)[S], const T (&Y)[S],
                                                 const Cand<T>& best, int m, int n, int j0, int ncols, bool col0,
                                                 T col0v) {
    Cand<T> c;
    c.v = T(0); c.i = -1; c.j = 0; c.st = 0;
    if constexpr (AM == kLocal) {
        c = best;
    } else if constexpr (AM == kGlobal) {
        if (n == 0) {
            if (col0) {
                c.i = m; c.j = 0;
                c.v = col0v;
                c.st = (m > 0 && GM == kAffine) ? 1 : 0;
            }
        } else if (n - 1 >= j0 && n - 1 < j0 + S) {
            const int kn = n - 1 - j0;
#pragma unroll
            for (int k = 0; k < S; ++k) {
                if (k == kn) {
                    c.i = m; c.j = n;
                    if constexpr (GM == kAffine) {
                        c.v = dmax(dmax(M[k], X[k]), Y[k]);
                        c.st = state_of(M[k], X[k], Y[k]);
                    } else { c.v = M[k]; c.st = 0; }
                }
            }
        }
    } else {   // Glocal: row m, the first column with the strictly largest value
        if (col0) {   // column 0 first
            c.i = m; c.j = 0;
            c.v = col0v;
            c.st = (m > 0 && GM == kAffine) ? 1 : 0;
        }
#pragma unroll
        for (int k = 0; k < S; ++k) {
            if (k < ncols) {
                T v3;
                int st3 = 0;
                if constexpr (GM == kAffine) { v3 = dmax(dmax(M[k], X[k]), Y[k]); st3 = state_of(M[k], X[k], Y[k]); }
                else v3 = M[k];
                if (c.i < 0 || v3 > c.v) { c.v = v3; c.i = m; c.j = j0 + k + 1; c.st = st3; }
            }
        }
    }
    return c;
}

// The dword of a wave's pointer-code slot that holds step t's codes k*4 .. k*4+3 of lane `lane`.
// Measured: lane-major at 8 columns a lane (protein_global_hard 300 -> 305 GCUPS, dwordx2
// stores) and step-major at 16 (dna_local_short 835 against 812 lane-major).
template <int S>
__device__ __forceinline__ int64_t code_word(int t, int lane, int k) {
    if constexpr (S == 8) return ((int64_t)t * kWave + lane) * (S / 4) + k;
    else return ((int64_t)t * (S / 4) + k) * kWave + lane;
}

// The column-0 cell of row i >= 1 (Global / Glocal; aligner.hpp:4009-4018, 3823-3826): the
// closed form untracked, the row's precomputed border tracked.
template <typename T, int GM, bool TR>
__device__ __forceinline__ T col0_border(const TrackArgs<T>& tr, int64_t r0, int i, T go_b, T ge_b) {
    if constexpr (TR) return tr.bx0[r0 + i];
    else if constexpr (GM == kAffine) return -(go_b + (T)i * ge_b);
    else return (-(T)i) * ge_b;
}

template <typename T, int GM, int AM, int S, bool HARD, bool TR, bool BAND>
__device__ void hard_tile(const HardArgs<T>& A, const T* __restrict__ ssub, uint16_t* __restrict__ hist,
                          int tile, int64_t slot, int lane) {
    static_assert(S % 4 == 0, "S is a multiple of 4");
    const TileDesc td = A.tiles[tile];
    const int G = td.G;
    const int lg = __builtin_ctz(G);
    const int q = lane >> lg, g = lane & (G - 1);
    const int p = q < td.npairs ? A.order[td.first + q] : -1;
    const int na = A.na, nn = na * na;
    const int K = TR ? A.tr.K : 1, knn = K * nn, gs = K * (nn + 4);
    int m = 0, n = 0;
    int64_t r0 = 0;   // tracked: pair p's first entry of the per-row arrays (gtrk, bx0)
    const uint8_t* ap = A.codes_a;
    const uint8_t* bp = A.codes_b;
    const uint16_t* tk = A.tr.trk;
    if (p >= 0) {
        const int64_t a0 = A.off_a[p], b0 = A.off_b[p];
        m = (int)(A.off_a[p + 1] - a0);
        n = (int)(A.off_b[p + 1] - b0);
        ap += a0;
        bp += b0;
        if constexpr (TR) { tk += a0; r0 = a0 + p; }
    }
    const int j0 = g * S;
    const int ncols = n - j0 < 0 ? 0 : (n - j0 > S ? S : n - j0);
    T go_a = A.go_a, ge_a = A.ge_a, go_b = A.go_b, ge_b = A.ge_b;
    // banded: the pair's guide from the previous pass, the band clamped to n, row 0's last column
    const int32_t* bgj = nullptr;
    int bw = 0, bj0 = 0;
    if constexpr (BAND) {
        if (p >= 0) {
            bgj = A.guide + A.off_a[p] + p;
            bw = A.band < n ? A.band : n;
            bj0 = band_hi<AM>(bgj, 0, m, n, bw);
        }
    }
    const T* __restrict__ gT = ssub + knn;   // tracked: the K x 4 gap table in LDS
    if constexpr (TR) {   // row 0's gaps in A: set gtrk of row 0
        if (p >= 0) {
            const int s0 = A.tr.gtrk[r0];
            go_a = gT[4 * s0]; ge_a = gT[4 * s0 + 1];
        }
    }
    const T NI = ninf<T>();

    int coff[S];
#pragma unroll
    for (int k = 0; k < S; ++k) coff[k] = (k < ncols) ? (int)bp[j0 + k] : 0;

    // Row 0 of the strip (aligner.hpp:3983-3999, 3790-3796).
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
        if constexpr (BAND && AM == kGlobal) {   // row 0's border runs up to column band_hi(0)
            if (j0 + k + 1 > bj0) { if (GM == kAffine) Y[k] = NI; else M[k] = NI; }
        }
    }
    // The diagonal of row 1's first cell: row 0 at column j0.
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

    Cand<T> best;   // Local: the lane's best cell so far (the reference starts at 0, (0,0), M)
    best.v = T(0); best.i = (AM == kLocal) ? 0 : -1; best.j = 0; best.st = 0;

    uint32_t* tb = A.table + slot * A.slot_dwords;
    // Score-only fills: the codes of the lane's rows i, i+1, i+2, loaded three rows ahead (a row's
    // code is otherwise loaded in the step that uses it; measured: dna_local_short score-only
    // 1,956 -> 2,019 GCUPS).  The hard fills load it in the step: their three more registers
    // cost dna_local_short a wave per SIMD (96 -> 98 registers), and the hard fills measured
    // slower with them.
    int ac0 = 0, ac1 = 0, ac2 = 0;
    if (!HARD && p >= 0) {
        if (m >= 1) ac0 = ap[0];
        if (m >= 2) ac1 = ap[1];
        if (m >= 3) ac2 = ap[2];
    }
    const int steps = td.max_m + G - 1;
    for (int t = 0; t < steps; ++t) {
        const int i = t - g + 1;
        const bool active = (p >= 0) && i >= 1 && i <= m && ncols > 0;
        int rs = 0;   // tracked: the row's set (matrix, gaps in B)
        if constexpr (TR) {
            if (p >= 0 && i >= 1 && i <= m) {
                rs = tk[i - 1];
                const int gsi = A.tr.gtrk[r0 + i];
                go_a = gT[4 * gsi]; ge_a = gT[4 * gsi + 1];
                go_b = gT[4 * rs + 2]; ge_b = gT[4 * rs + 3];
            }
        }
        T iM = shift_up(oM);
        T iX, iY;
        if constexpr (GM == kAffine) { iX = shift_up(oX); iY = shift_up(oY); }
        else { iX = NI; iY = NI; }
        if (g == 0) {   // column 0 of row i (aligner.hpp:4009-4018, 3823-3826)
            T border = NI;
            if constexpr (AM != kLocal) {
                if (!TR || (p >= 0 && i >= 1 && i <= m)) border = col0_border<T, GM, TR>(A.tr, r0, i, go_b, ge_b);
                if constexpr (BAND) {   // past border_rows(): outside the band
                    if (!(p >= 0 && i >= 1 && i <= m && bgj[i] <= bw)) border = NI;
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
        if (active) {
            const T* __restrict__ srow = ssub + ((TR ? rs * na : 0) + (HARD ? (int)ap[i - 1] : ac0)) * na;
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
                for (int k = 0; k < S / 4; ++k) tb[code_word<S>(t, lane, k)] = cw[k];
            }
            if constexpr (AM == kLocal) local_update<T, GM, S>(best, M, X, Y, rowmax, i, j0, ncols);
        }
        pM = iM; pX = iX; pY = iY;
        if (!HARD && p >= 0 && i >= 1 && i <= m) {   // the next row's code, and one more ahead
            ac0 = ac1;
            ac1 = ac2;
            if (i + 2 < m) ac2 = ap[i + 2];
        }
    }

    Cand<T> c;
    c.v = T(0); c.i = -1; c.j = 0; c.st = 0;
    if (p >= 0) {
        T col0v = (m == 0 || AM == kLocal) ? T(0) : col0_border<T, GM, TR>(A.tr, r0, m, A.go_b, A.ge_b);
        if constexpr (BAND) {
            if (AM != kLocal && m > 0 && bgj[m] > bw) col0v = NI;
        }
        c = end_candidate<T, GM, AM, S>(M, X, Y, best, m, n, j0, ncols, g == 0, col0v);
    }
    c = reduce_slot(c, G);

    // The walk.  Untracked: gap counts in registers.  Tracked: every count in its row's set,
    // K x 4 gap counters per pair slot in LDS after the slots' histograms.
    int c_goa = 0, c_gea = 0, c_gob = 0, c_geb = 0;
    int* const gall = reinterpret_cast<int*>(reinterpret_cast<unsigned char*>(hist) +
                                             ((A.hist_slots * knn * 2 + 15) & ~15));
    int* gcnt = gall + q * 4 * K;
    if (g == 0 && p >= 0) {
        if constexpr (HARD) {
            uint16_t* h = hist + q * knn;
            int i = c.i, j = c.j, st = c.st;
            const uint32_t* tr = tb;
            // The pointer-code word of cell (i, j), i, j >= 1.  Each step loads the next step's
            // word before it stores this step's guide entry and path op (and the first before
            // the score and the guide's fixed entries): a load issued after a store would wait
            // for the store too (one counter covers both).
            auto word_at = [&](int ci, int cj) -> uint32_t {
                const int cg = (cj - 1) / S, ck = (cj - 1) - cg * S;
                return tr[code_word<S>(ci - 1 + cg, q * G + cg, ck >> 2)];
            };
            uint32_t wnext = (i > 0 && j > 0) ? word_at(i, j) : 0u;
            A.scores[p] = (double)c.v;
            uint32_t* const oend = A.ops ? A.ops + A.off_a[p + 1] + A.off_b[p + 1] : nullptr;
            WalkOut wo;
            wo.begin(A.guide ? A.guide + A.off_a[p] + p : nullptr, oend, m, n, c.i, c.j);
            // count a gap move: f 0 go_a, 1 ge_a (set gtrk of row i), 2 go_b, 3 ge_b (set trk)
            auto gap = [&](int f, int row, int v) {
                if constexpr (TR) {
                    const int sset = f < 2 ? A.tr.gtrk[r0 + row] : tk[row - 1];
                    gcnt[4 * sset + f] += v;
                } else {
                    if (f == 0) c_goa += v;
                    else if (f == 1) c_gea += v;
                    else if (f == 2) c_gob += v;
                    else c_geb += v;
                }
            };
            for (;;) {
                if (i == 0 && j == 0) break;
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
                const int k = (j - 1) % S;
                const uint32_t cc = (wnext >> (8 * (k & 3))) & 0xffu;
                // the move out of (i, j): 0 diagonal, 1 up (a gap in B), 2 left (a gap in A),
                // and the state it enters (affine)
                int mv, nst = 0;
                if constexpr (GM == kAffine) {
                    const uint32_t kk = st == 0 ? (cc & 3u) : (st == 1 ? ((cc >> 2) & 3u) : (cc >> 4));
                    if (AM == kLocal && st == 0 && kk == 3u) break;
                    mv = st;
                    nst = kk == 1u ? 1 : (kk == 2u ? 2 : 0);
                } else {
                    if (cc == 3u) break;
                    mv = cc == 0u ? 0 : (cc == 1u ? 1 : 2);
                }
                const int ni = mv == 2 ? i : i - 1, nj = mv == 1 ? j : j - 1;
                if (ni > 0 && nj > 0) wnext = word_at(ni, nj);
                if (mv == 0) {
                    const int mrow = ((TR ? (int)tk[i - 1] * na : 0) + (int)ap[i - 1]) * na + (int)bp[j - 1];
                    h[mrow] += 1;
                } else if (mv == 1) {
                    gap(3, i, 1);
                    if (GM == kAffine && nst != 1) gap(2, i, 1);
                } else {
                    gap(1, i, 1);
                    if (GM == kAffine && nst != 2) gap(0, i, 1);
                }
                wo.move(mv, i, j);
                if constexpr (GM == kAffine) st = nst;
                i = ni;
                j = nj;
            }
            if constexpr (AM == kLocal) wo.end_local(i, m, n, c.i, c.j);
            if (oend) {
                A.ops_n[p] = wo.finish(oend);
                A.ends[2 * p] = c.i;
                A.ends[2 * p + 1] = c.j;
            }
        } else {
            A.scores[p] = (double)c.v;
        }
    }
    if constexpr (HARD) {
        wave_sync();
        const int np = td.npairs;
        for (int idx = lane; idx < np * knn; idx += kWave) {
            const int qq = idx / knn, e = idx - qq * knn;
            const int pp = A.order[td.first + qq];
            A.grads[(int64_t)pp * gs + e] = (double)hist[qq * knn + e];
            hist[qq * knn + e] = 0;
        }
        if constexpr (TR) {
            for (int idx = lane; idx < np * 4 * K; idx += kWave) {
                const int qq = idx / (4 * K), e = idx - qq * 4 * K;
                const int pp = A.order[td.first + qq];
                A.grads[(int64_t)pp * gs + knn + e] = 0.0 - (double)gall[qq * 4 * K + e];
                gall[qq * 4 * K + e] = 0;
            }
        } else if (g == 0 && p >= 0) {
            double* gr = A.grads + (int64_t)p * gs + nn;
            gr[0] = 0.0 - (double)c_goa;
            gr[1] = 0.0 - (double)c_gea;
            gr[2] = 0.0 - (double)c_gob;
            gr[3] = 0.0 - (double)c_geb;
        }
        wave_sync();
    }
}

// LDS of the hard / score-only fills: the K matrices and (tracked) the K x 4 gap table, then
// (hard) per pair slot a histogram of K * n_alpha^2 uint16 counts, then (tracked) K x 4 int
// gap counters per slot.
template <typename T>
__host__ __device__ constexpr int hard_lds_sub_bytes(int na, int K, bool tr) {
    return ((K * na * na + (tr ? 4 * K : 0)) * (int)sizeof(T) + 15) & ~15;
}
__host__ __device__ constexpr int hard_lds_hist_bytes(int na, int K, bool tr, int slots) {
    return ((slots * K * na * na * 2 + 15) & ~15) + (tr ? slots * 4 * K * 4 : 0);
}

// One wavefront per block.
//  - Plain launch (queue == nullptr): block b fills tile tile_base + b into pointer-code slot
//    b.  Over the tiles in their persistent order (most row steps first) this is a
//    longest-first work queue without a counter: the dispatcher hands blocks to the CUs in
//    index order as slots free up.  It needs a slot per tile.
//  - Persistent launch: each block pulls tiles from the queue into its own slot until the
//    queue is drained (every wave reaches the exit test once the counter passes n_tiles),
//    for batches whose codes would not fit a slot per tile.
// The slot histograms start at zero and each tile's flush clears what it used.
template <typename T, int GM, int AM, int S, bool HARD, bool TR, bool BAND = false>
__global__ __launch_bounds__(64) void hard_fill_kernel(HardArgs<T> A) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int na = A.na, nn = na * na;
    const int K = TR ? A.tr.K : 1;
    T* ssub = reinterpret_cast<T*>(lds);
    uint16_t* hist = reinterpret_cast<uint16_t*>(lds + hard_lds_sub_bytes<T>(na, K, TR));
    const int lane = threadIdx.x;
    for (int k = lane; k < K * nn; k += kWave) ssub[k] = A.sub[k];
    if constexpr (TR)
        for (int k = lane; k < 4 * K; k += kWave) ssub[K * nn + k] = A.tr.gaps[k];
    if constexpr (HARD) {
        const int words = hard_lds_hist_bytes(na, K, TR, A.hist_slots) / 2;
        for (int k = lane; k < words; k += kWave) hist[k] = 0;
    }
    __syncthreads();
    if (A.queue == nullptr) {
        const int tile = A.tile_base + (int)blockIdx.x;
        if (tile < A.n_tiles) hard_tile<T, GM, AM, S, HARD, TR, BAND>(A, ssub, hist, tile, (int64_t)blockIdx.x, lane);
        return;
    }
    for (;;) {
        const int tile = wave_fetch(A.queue);
        if (tile >= A.n_tiles) break;
        hard_tile<T, GM, AM, S, HARD, TR, BAND>(A, ssub, hist, tile, (int64_t)blockIdx.x, lane);
    }
}

}  // namespace nwg
