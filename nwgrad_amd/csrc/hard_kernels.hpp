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
// slot of HBM scratch, four cells to a dword, laid out so the 64 lanes of a step write
// 64 consecutive dwords.  After the fill, lane 0 of each slot walks its pair's codes
// back from the end cell, counting matches into a per-slot LDS histogram and gap moves
// in registers, and the wave writes the per-pair gradient rows.
#pragma once

#include "common.hpp"

namespace nwg {

template <typename T>
struct HardArgs {
    const uint8_t* codes_a;
    const uint8_t* codes_b;
    const int64_t* off_a;
    const int64_t* off_b;
    const int32_t* order;     // pair indices, tile-major
    const TileDesc* tiles;
    int32_t n_tiles;
    int32_t tile_base;        // plain launch: the first tile of this launch
    int32_t* queue;           // persistent launch: next-tile counter (nullptr: plain)
    const T* sub;             // n_alpha * n_alpha, the matrix cast to T
    int32_t na;
    int32_t hist_slots;       // LDS histogram slots per wave
    T go_a, ge_a, go_b, ge_b;
    double* scores;           // per pair
    double* grads;            // per pair, na*na + 4 (hard only)
    uint32_t* table;          // pointer-code scratch (hard only)
    int64_t slot_dwords;      // dwords per table slot
};

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
template <typename T, int GM, int AM, int S, bool HARD>
__device__ __forceinline__ void strip_row(T (&M)[S], T (&X)[S], T (&Y)[S], T lM, T lX, T lY, T dM, T dX, T dY,
                                          const T* __restrict__ srow, const int (&coff)[S], int ncols, T go_a,
                                          T ge_a, T go_b, T ge_b, uint32_t (&cw)[S / 4], T& rowmax) {
#pragma unroll
    for (int k = 0; k < S / 4; ++k) cw[k] = 0u;
    rowmax = T(0);
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const T s = srow[coff[k]];
        if constexpr (GM == kAffine) {
            const T uM = M[k], uX = X[k], uY = Y[k];
            const T tXY = dmax(dX, dY);
            T mv = dmax(dM, tXY) + s;
            uint32_t km = 0;
            if constexpr (HARD) km = (dM >= tXY) ? 0u : ((dX >= dY) ? 1u : 2u);
            const T ux = (uM - go_b) - ge_b, vx = uX - ge_b, wx = (uY - go_b) - ge_b;
            const T tx = dmax(vx, wx);
            const T xv = dmax(ux, tx);
            const T uy = (lM - go_a) - ge_a, vy = (lX - go_a) - ge_a, wy = lY - ge_a;
            const T ty = dmax(vy, wy);
            const T yv = dmax(uy, ty);
            if constexpr (AM == kLocal) {
                mv = dmax(mv, T(0));
                if constexpr (HARD) km = (mv <= T(0)) ? 3u : km;
                if (k < ncols) rowmax = dmax(rowmax, dmax(mv, dmax(xv, yv)));
            }
            if constexpr (HARD) {
                const uint32_t kx = (ux >= tx) ? 0u : ((vx >= wx) ? 1u : 2u);
                const uint32_t ky = (uy >= ty) ? 0u : ((vy >= wy) ? 1u : 2u);
                cw[k >> 2] |= (km | (kx << 2) | (ky << 4)) << (8 * (k & 3));
            }
            M[k] = mv; X[k] = xv; Y[k] = yv;
            dM = uM; dX = uX; dY = uY;
            lM = mv; lX = xv; lY = yv;
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
// 6121-6134).  col0: this lane also offers column 0 (the first lane of the pair's first panel).
template <typename T, int GM, int AM, int S>
__device__ __forceinline__ Cand<T> end_candidate(const T (&M)[S], const T (&X)[S], const T (&Y)[S],
                                                 const Cand<T>& best, int m, int n, int j0, int ncols, bool col0,
                                                 T go_b, T ge_b) {
    Cand<T> c;
    c.v = T(0); c.i = -1; c.j = 0; c.st = 0;
    if constexpr (AM == kLocal) {
        c = best;
    } else if constexpr (AM == kGlobal) {
        if (n == 0) {
            if (col0) {
                c.i = m; c.j = 0;
                if (m == 0) { c.v = T(0); c.st = 0; }
                else if constexpr (GM == kAffine) { c.v = -(go_b + (T)m * ge_b); c.st = 1; }
                else { c.v = (-(T)m) * ge_b; c.st = 0; }
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
            if (m == 0) { c.v = T(0); c.st = 0; }
            else if constexpr (GM == kAffine) { c.v = -(go_b + (T)m * ge_b); c.st = 1; }
            else { c.v = (-(T)m) * ge_b; c.st = 0; }
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

template <typename T, int GM, int AM, int S, bool HARD>
__device__ void hard_tile(const HardArgs<T>& A, const T* __restrict__ ssub, uint16_t* __restrict__ hist,
                          int tile, int64_t slot, int lane) {
    static_assert(S % 4 == 0, "S is a multiple of 4");
    const TileDesc td = A.tiles[tile];
    const int G = td.G;
    const int lg = __builtin_ctz(G);
    const int q = lane >> lg, g = lane & (G - 1);
    const int p = q < td.npairs ? A.order[td.first + q] : -1;
    const int na = A.na, nn = na * na, gs = nn + 4;
    int m = 0, n = 0;
    const uint8_t* ap = A.codes_a;
    const uint8_t* bp = A.codes_b;
    if (p >= 0) {
        const int64_t a0 = A.off_a[p], b0 = A.off_b[p];
        m = (int)(A.off_a[p + 1] - a0);
        n = (int)(A.off_b[p + 1] - b0);
        ap += a0;
        bp += b0;
    }
    const int j0 = g * S;
    const int ncols = n - j0 < 0 ? 0 : (n - j0 > S ? S : n - j0);
    const T go_a = A.go_a, ge_a = A.ge_a, go_b = A.go_b, ge_b = A.ge_b;
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
    }
    // The diagonal of row 1's first cell: row 0 at column j0.
    T pM, pX, pY;
    {
        const T jt = (T)j0;
        if constexpr (GM == kAffine) {
            if (AM == kGlobal) {
                pM = (j0 == 0) ? T(0) : NI; pX = NI; pY = (j0 == 0) ? NI : -(go_a + jt * ge_a);
            } else { pM = T(0); pX = NI; pY = NI; }
        } else {
            pM = (AM == kGlobal) ? (-jt) * ge_a : T(0); pX = NI; pY = NI;
        }
    }
    T oM = M[S - 1], oX = X[S - 1], oY = Y[S - 1];

    Cand<T> best;   // Local: the lane's best cell so far (the reference starts at 0, (0,0), M)
    best.v = T(0); best.i = (AM == kLocal) ? 0 : -1; best.j = 0; best.st = 0;

    uint32_t* tb = A.table + slot * A.slot_dwords;
    const int steps = td.max_m + G - 1;
    for (int t = 0; t < steps; ++t) {
        const int i = t - g + 1;
        T iM = shift_up(oM);
        T iX, iY;
        if constexpr (GM == kAffine) { iX = shift_up(oX); iY = shift_up(oY); }
        else { iX = NI; iY = NI; }
        if (g == 0) {   // column 0 of row i (aligner.hpp:4009-4018, 3823-3826)
            const T it = (T)i;
            if constexpr (GM == kAffine) {
                if (AM == kLocal) { iM = T(0); iX = NI; }
                else { iM = NI; iX = -(go_b + it * ge_b); }
                iY = NI;
            } else {
                iM = (AM == kLocal) ? T(0) : (-it) * ge_b;
            }
        }
        const bool active = (p >= 0) && i >= 1 && i <= m && ncols > 0;
        if (active) {
            const int rowoff = (int)ap[i - 1] * na;
            const T* __restrict__ srow = ssub + rowoff;
            uint32_t cw[S / 4];
            T rowmax;
            strip_row<T, GM, AM, S, HARD>(M, X, Y, iM, iX, iY, pM, pX, pY, srow, coff, ncols, go_a, ge_a, go_b, ge_b,
                                          cw, rowmax);
            oM = M[S - 1];
            if constexpr (GM == kAffine) { oX = X[S - 1]; oY = Y[S - 1]; }
            if constexpr (HARD) {
#pragma unroll
                for (int k = 0; k < S / 4; ++k)
                    tb[((int64_t)t * (S / 4) + k) * kWave + lane] = cw[k];
            }
            if constexpr (AM == kLocal) local_update<T, GM, S>(best, M, X, Y, rowmax, i, j0, ncols);
        }
        pM = iM; pX = iX; pY = iY;
    }

    Cand<T> c;
    c.v = T(0); c.i = -1; c.j = 0; c.st = 0;
    if (p >= 0) c = end_candidate<T, GM, AM, S>(M, X, Y, best, m, n, j0, ncols, g == 0, go_b, ge_b);
    c = reduce_slot(c, G);

    int c_goa = 0, c_gea = 0, c_gob = 0, c_geb = 0;
    if (g == 0 && p >= 0) {
        A.scores[p] = (double)c.v;
        if constexpr (HARD) {
            uint16_t* h = hist + q * nn;
            int i = c.i, j = c.j, st = c.st;
            const uint32_t* tr = tb + q * G;
            for (;;) {
                if (i == 0 && j == 0) break;
                if (i == 0 || j == 0) {   // borders: runs (Global / Glocal column 0) or a stop
                    if (AM == kLocal) break;
                    if (i == 0) {
                        if (AM == kGlocal) break;
                        c_gea += j;
                        if (GM == kAffine) c_goa += 1;
                        break;
                    }
                    c_geb += i;
                    if (GM == kAffine) c_gob += 1;
                    break;
                }
                const int jj = j - 1;
                const int go = jj / S, k = jj - go * S;
                const uint32_t w = tr[((int64_t)(i - 1 + go) * (S / 4) + (k >> 2)) * kWave + go];
                const uint32_t cc = (w >> (8 * (k & 3))) & 0xffu;
                if constexpr (GM == kAffine) {
                    if (st == 0) {
                        if (AM == kLocal && (cc & 3u) == 3u) break;
                        const uint32_t kk = cc & 3u;
                        h[(int)ap[i - 1] * na + (int)bp[j - 1]] += 1;
                        st = kk == 1u ? 1 : (kk == 2u ? 2 : 0);
                        --i; --j;
                    } else if (st == 1) {
                        const uint32_t kk = (cc >> 2) & 3u;
                        const int pv = kk == 1u ? 1 : (kk == 2u ? 2 : 0);
                        c_geb += 1;
                        c_gob += (pv != 1);
                        st = pv;
                        --i;
                    } else {
                        const uint32_t kk = cc >> 4;
                        const int pv = kk == 1u ? 1 : (kk == 2u ? 2 : 0);
                        c_gea += 1;
                        c_goa += (pv != 2);
                        st = pv;
                        --j;
                    }
                } else {
                    if (cc == 3u) break;
                    if (cc == 0u) { h[(int)ap[i - 1] * na + (int)bp[j - 1]] += 1; --i; --j; }
                    else if (cc == 1u) { c_geb += 1; --i; }
                    else { c_gea += 1; --j; }
                }
            }
        }
    }
    if constexpr (HARD) {
        wave_sync();
        const int np = td.npairs;
        for (int idx = lane; idx < np * nn; idx += kWave) {
            const int qq = idx / nn, e = idx - qq * nn;
            const int pp = A.order[td.first + qq];
            A.grads[(int64_t)pp * gs + e] = (double)hist[qq * nn + e];
            hist[qq * nn + e] = 0;
        }
        if (g == 0 && p >= 0) {
            double* gr = A.grads + (int64_t)p * gs + nn;
            gr[0] = 0.0 - (double)c_goa;
            gr[1] = 0.0 - (double)c_gea;
            gr[2] = 0.0 - (double)c_gob;
            gr[3] = 0.0 - (double)c_geb;
        }
        wave_sync();
    }
}

// One wavefront per block.  Plain launch: block b fills tile tile_base + b.  Persistent
// launch: each block pulls tiles from the queue until it is drained; every wave reaches
// the exit test once the counter passes n_tiles.
template <typename T, int GM, int AM, int S, bool HARD>
__global__ __launch_bounds__(64) void hard_fill_kernel(HardArgs<T> A) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int na = A.na, nn = na * na;
    T* ssub = reinterpret_cast<T*>(lds);
    uint16_t* hist = reinterpret_cast<uint16_t*>(lds + ((nn * (int)sizeof(T) + 15) & ~15));
    const int lane = threadIdx.x;
    for (int k = lane; k < nn; k += kWave) ssub[k] = A.sub[k];
    if constexpr (HARD)
        for (int k = lane; k < A.hist_slots * nn; k += kWave) hist[k] = 0;
    __syncthreads();
    if (A.queue == nullptr) {
        const int tile = A.tile_base + (int)blockIdx.x;
        if (tile < A.n_tiles) hard_tile<T, GM, AM, S, HARD>(A, ssub, hist, tile, tile, lane);
        return;
    }
    for (;;) {
        const int tile = wave_uniform(atomicAdd(A.queue, lane == 0 ? 1 : 0));
        if (tile >= A.n_tiles) break;
        hard_tile<T, GM, AM, S, HARD>(A, ssub, hist, tile, (int64_t)blockIdx.x, lane);
    }
}

}  // namespace nwg
