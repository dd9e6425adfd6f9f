// Soft (forward-backward) pass on gfx950: expected counts of the alignment posterior.
//
// Reference behaviour restated (relative to the reference's src/nwgrad/cpp/nwgrad/):
//   scaled forward          soft_kernel_impl.inl:643-708 (sp_affine_fwd), 739-760 (row 0)
//   backward + fused sums   soft_kernel_impl.inl:790-977
//   Z and the end rules     soft_kernel_impl.inl:294-321 (sp_log_z)
//   gap-target rows         soft_kernel_impl.inl:736-737 (kmin, kmin0)
//   lost-mass bound         soft_kernel_impl.inl:141-178 (sp_loss), 933-948 (flushed products)
//   score = T log Z, gap fields negated  aligner.hpp:6554-6626, soft_kernel_impl.inl:987
// Linear gaps run as the affine model with open = 0: with oa = ea and ob = eb the affine
// total M + X + Y obeys H(i,j) = H(i-1,j-1) P + H(i-1,j) eb + H(i,j-1) ea, which is the linear
// recurrence (aligner.hpp:6231-6260), and the expected X / Y moves are its gap counts.
//
// Mapping.  A tile is one wavefront holding up to kTilePairs pairs side by side: pair q owns
// lanes lo[q] .. lo[q]+G[q]-1, and lane g of it owns S columns of each of the tile's column
// panels (panel pp: columns pp*G*S + g*S + 1 .. + S).  Lanes run skewed (row t-g+1 at step t)
// and receive the boundary column from the lane below by a DPP wave shift; a panel's last
// column passes to the next panel through HBM.  The backward runs the mirrored skew.
//
// Scaling.  The reference scales each DP row by a power of two chosen from the row maximum;
// lanes of a slot run skewed and never hold a whole row at once, so every (lane, row)
// segment carries its own power-of-two exponent E (true value = stored value * 2^E).  A row
// is computed and stored in one exponent; when its maximum leaves [2^-w, 2^w] the NEXT row is
// computed in new units, by folding the factor 2^-k into that row's weights (the profile
// values and the two vertical gap weights) instead of rescaling the registers.  Every value
// is therefore computed once, in the units it is stored in, and a value that flushes was
// below the format's MIN in its own stored units -- the premise of the lost-mass bound.
// Boundary values from a neighbour are converted by an exact power-of-two product; a segment
// without mass of its own takes units in which the neighbour's boundary values are of order
// one, and one whose neighbour's values would not fit below 2^headroom of its own units moves
// to the neighbour's units (tracked: to the boundary's magnitude) minus the headroom (its old
// values then flush towards zero, as in the reference's row scaling).
//
// Storage.  The forward stores every row segment of every lane -- M, X, Y of its S columns,
// the boundary column it received, the exponent and the segment maximum -- as 16-byte
// vectors laid out so that each vector of the 64 lanes is contiguous.  The backward needs ONE
// forward row per step: the sums of row i+1 that read forward row i (the X-open counts and
// the match posteriors) are formed one step late, from the row-(i+1) backward values still in
// registers.  The next forward row is loaded one step ahead.  Row sums are formed in the
// working precision, as in the reference, then scaled into double by the row's posterior
// factor 2^(E_F + E_B - ze) / zr; match posteriors go into a per-pair double LDS histogram
// (ds_add_f64).
#pragma once

#include "common.hpp"

namespace nwg {

constexpr int kTilePairs = 16;  // pairs per soft tile at most
constexpr int kCK = 4;          // backward bands: one forward row in kCK stored in full
constexpr int kPipeChunk = 64;  // pipelined wide tiles: steps between progress waits / stores
static_assert(kPipeChunk % (2 * kCK) == 0, "a chunk is whole backward loop trips");

// One soft tile: pairs order[first .. first + npairs), pair q on lanes lo[q] .. lo[q]+G[q]-1.
struct SoftTile {
    int32_t first;
    int32_t max_m;
    int32_t max_n;
    int32_t fsteps;    // max over pairs of m + G - 1: row steps of one forward panel
    int16_t npairs;
    int16_t panels;    // column panels (every pair of the tile runs the same count)
    uint8_t G[kTilePairs];
    uint8_t lo[kTilePairs];
    // Persistent wide tiles (the queue's order): the tile's words in SoftArgs::lsync, its
    // region of the wide pool (first row index; panels x (fsteps + 1) rows), the tiles whose
    // regions it reuses (lwait[wait0 .. wait0 + nwait)), and (pipelined) the index of its
    // first panel among the pipelined tiles' panels.
    int32_t sync;
    int64_t rbase;
    int32_t wait0, nwait;
    int32_t poff;
    int32_t pad;
};
static_assert(sizeof(SoftTile) == 80, "SoftTile is 80 bytes");

// One scratch pool: row segments and panel-boundary columns per wave slot.
template <typename SF>
struct SoftView {
    SF* seg;          // row segments: per row index, (NV x 64) vectors
    SF* pbnd;         // panel boundary columns: per row index, pair: M, X, Y, BM, BY
    int32_t* pexp;    // their exponents: forward, backward
    int64_t rows;     // slots: rows of one panel region (at least fsteps + 1 over the pool's tiles)
    int32_t panels;   // slots: panel regions per slot
};

template <typename SF>
struct SoftArgs {
    const uint8_t* codes_a;
    const uint8_t* codes_b;
    const int64_t* off_a;
    const int64_t* off_b;
    // main queue (or, without a queue, the plain launch's tiles from tile_base)
    const int32_t* order;
    const SoftTile* tiles;
    int32_t n_tiles;
    int32_t tile_base;
    int32_t* queue;
    int32_t n_mslots;        // blocks below this own a main-pool slot and drain the main queue
    SoftView<SF> view;
    // Wide tiles (column panels).  Persistent: longest first.  Tiles below n_ptiles have their
    // items in the item queue, each tile in its own region of the circular wide pool: (tile,
    // 2 * panel + direction) one panel's forward (0) or backward (1) of a pipelined tile, (tile,
    // -1) a whole tile (see soft_tile).  The others run whole on one wave, from the wide queue,
    // in the block's wide slot.  Blocks below n_pwaves drain the item queue first, blocks below
    // n_lslots the wide queue.  Plain: whole tiles in block slots.
    const int32_t* lorder;
    const SoftTile* ltiles;
    int32_t n_ltiles;
    int32_t n_ptiles;
    int32_t n_lslots;
    int32_t n_pwaves;
    int32_t* lqueue;         // the item queue's counter, then the wide queue's
    SoftView<SF> lview;
    const int2* litems;
    int32_t n_litems;
    const int32_t* lwait;    // the tiles whose pool regions a tile reuses
    int32_t* lsync;          // [0] error word, [1 + t] tile t done, then each tile's words
    double* lzpart;          // per pipelined panel, pair slot: forward Z part (acc, E) and flag
    double* lpart;           // per pipelined panel, histogram slot: n_alpha^2 counts, 5 sums, flag
    uint64_t wait_ticks;     // bound of one wait, s_memrealtime ticks
    const SF* psub;          // exp(matrix / T), K x n_alpha x n_alpha (LDS adds a row of ones)
    int32_t na;
    int32_t hist_slots;
    const SF* gapw;          // per set: oa, ea, ob, eb: exp(-(go+ge)/T), exp(-ge/T) per sequence
    // Tracks (the TR kernels): K sets; DP row i of pair p scores with set trk[off_a[p] + i - 1]
    // (its matrix and its gaps in B) and charges gaps in A to set gtrk[off_a[p] + p + i].
    const uint16_t* trk;
    const uint16_t* gtrk;
    int32_t K;
    int32_t linear;          // 1: report the open fields as 0
    double temperature;
    double* scores;
    double* grads;
    uint8_t* ok;             // per pair: 1 accepted, 0 refused (range, non-finite or lost mass)
    double tol;              // lost-mass tolerance (2^-20 float32, 2^-45 double)
    double vmin;             // the working format's smallest normal value
    double vtiny;            // its smallest subnormal value
    // Banded passes (the BAND kernels): each pair's guide (m + 1 columns at off_a[p] + p) and
    // the band's half-width.
    const int32_t* guide;
    int32_t band;
};

template <typename SF>
struct SoftRange;
template <>
struct SoftRange<float> {
    static constexpr int headroom = 64;  // boundary values at most 2^64 above a segment's units
    static constexpr int kmax = 120;     // largest renormalisation step
    static constexpr float hi = 65536.0f, lo = 1.0f / 65536.0f;   // renormalise outside [2^-16, 2^16]
    static constexpr float fmax = 3.4028234663852886e38f;
};
template <>
struct SoftRange<double> {
    static constexpr int headroom = 512;
    static constexpr int kmax = 1000;
    static constexpr double hi = 1.157920892373162e77, lo = 8.636168555094445e-78;   // 2^+-256
    static constexpr double fmax = 1.7976931348623157e308;
};

__device__ __forceinline__ float sldexp(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
__device__ __forceinline__ double sldexp(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }
__device__ __forceinline__ int silogb(float x) { return __builtin_amdgcn_frexp_expf(x) - 1; }
__device__ __forceinline__ int silogb(double x) { return __builtin_amdgcn_frexp_exp(x) - 1; }
__device__ __forceinline__ float sfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double sfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
// A row segment's exponent E and a power-of-two bound on its maximum, 2^x >= max (x = -128:
// no mass), packed into the bits of one stored value: float (E << 8 | x) with |E| < 2^23,
// double (E << 16 | x).
constexpr int kPackEMax = 1 << 23;
__device__ __forceinline__ int mx_exp(float mx) {
    if (!(mx > 0.0f)) return -128;
    const int e = __builtin_amdgcn_frexp_expf(mx);   // mx < 2^e
    return e > 127 ? 127 : (e < -127 ? -127 : e);
}
__device__ __forceinline__ int mx_exp(double mx) {
    if (!(mx > 0.0)) return -32768;
    return __builtin_amdgcn_frexp_exp(mx);
}
__device__ __forceinline__ float pack_em(int E, float mx) { return __int_as_float((int)((unsigned)E << 8) | (mx_exp(mx) & 0xff)); }
__device__ __forceinline__ double pack_em(int E, double mx) {
    return __longlong_as_double((long long)((unsigned long long)(long long)E << 16) | (mx_exp(mx) & 0xffff));
}
__device__ __forceinline__ int em_E(float v) { return __float_as_int(v) >> 8; }
__device__ __forceinline__ int em_E(double v) { return (int)(__double_as_longlong(v) >> 16); }
__device__ __forceinline__ float em_max(float v) {
    const int x = (int)(signed char)(__float_as_int(v) & 0xff);
    return x == -128 ? 0.0f : __builtin_amdgcn_ldexpf(1.0f, x);
}
__device__ __forceinline__ double em_max(double v) {
    const int x = (int)(short)(__double_as_longlong(v) & 0xffff);
    return x == -32768 ? 0.0 : __builtin_amdgcn_ldexp(1.0, x);
}

// The value, hidden from loop-invariant code motion: a per-column comparison against it is
// evaluated where it is used (one compare into VCC) instead of being hoisted out of the row
// loop as S live 64-bit lane masks.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Mass with a separate binary exponent: value = acc * 2^E.
struct ExpAcc {
    double acc;
    int E;
};
__device__ __forceinline__ void expacc_add(ExpAcc& a, double v, int e) {
    if (!(v > 0.0)) return;
    if (!(a.acc > 0.0)) { a.acc = v; a.E = e; }
    else if (e > a.E) { a.acc = __builtin_amdgcn_ldexp(a.acc, a.E - e) + v; a.E = e; }
    else a.acc += __builtin_amdgcn_ldexp(v, e - a.E);
    const int k = __builtin_amdgcn_frexp_exp(a.acc) - 1;
    a.acc = __builtin_amdgcn_ldexp(a.acc, -k);
    a.E += k;
}

// Row segment of one lane: M[S], X[S], Y[S], then {cM, cX, cY, E|max} (the boundary column
// and the packed exponent and maximum bound); NV 16-byte vectors, vector v of lane l at
// (v * 64 + l) * VW values from the row's base.
template <typename SF, int S>
struct Seg {
    static constexpr int VW = 16 / (int)sizeof(SF);
    static constexpr int vals = 3 * S + 4;
    static constexpr int NV = vals / VW;
    static constexpr int row = NV * kWave * VW;   // values per stored row of a wave slot
    static constexpr int iX = S, iY = 2 * S, iC = 3 * S, iE = 3 * S + 3;
    static_assert(vals % VW == 0, "segment is a whole number of vectors");
};

template <typename SF, int S>
__device__ __forceinline__ void store_seg(SF* __restrict__ base, int lane, const SF (&M)[S], const SF (&X)[S],
                                          const SF (&Y)[S], SF cM, SF cX, SF cY, int E, SF mx) {
    using L = Seg<SF, S>;
    SF v[L::vals];
#pragma unroll
    for (int k = 0; k < S; ++k) { v[k] = M[k]; v[L::iX + k] = X[k]; v[L::iY + k] = Y[k]; }
    v[L::iC] = cM; v[L::iC + 1] = cX; v[L::iC + 2] = cY; v[L::iE] = pack_em(E, mx);
    if constexpr (sizeof(SF) == 4) {
#pragma unroll
        for (int q = 0; q < L::NV; ++q)
            reinterpret_cast<float4*>(base)[q * kWave + lane] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    } else {
#pragma unroll
        for (int q = 0; q < L::NV; ++q)
            reinterpret_cast<double2*>(base)[q * kWave + lane] = make_double2(v[2 * q], v[2 * q + 1]);
    }
}

template <typename SF, int S>
__device__ __forceinline__ void load_seg(const SF* __restrict__ base, int lane, SF (&v)[Seg<SF, S>::vals]) {
    using L = Seg<SF, S>;
    if constexpr (sizeof(SF) == 4) {
#pragma unroll
        for (int q = 0; q < L::NV; ++q) {
            const float4 f = reinterpret_cast<const float4*>(base)[q * kWave + lane];
            v[4 * q] = f.x; v[4 * q + 1] = f.y; v[4 * q + 2] = f.z; v[4 * q + 3] = f.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < L::NV; ++q) {
            const double2 f = reinterpret_cast<const double2*>(base)[q * kWave + lane];
            v[2 * q] = f.x; v[2 * q + 1] = f.y;
        }
    }
}

// The boundary record of a row that is not stored in full: {cM, cX, cY, E}, the values the
// full row holds at Seg::iC .. iE (one 16-byte vector for float, two for double).
template <typename SF, int S>
__device__ __forceinline__ void store_rec(SF* __restrict__ base, int lane, SF cM, SF cX, SF cY, int E) {
    using L = Seg<SF, S>;
    if constexpr (sizeof(SF) == 4) {
        reinterpret_cast<float4*>(base)[(L::iC / 4) * kWave + lane] = make_float4(cM, cX, cY, pack_em(E, SF(0)));
    } else {
        reinterpret_cast<double2*>(base)[(L::iC / 2) * kWave + lane] = make_double2(cM, cX);
        reinterpret_cast<double2*>(base)[(L::iC / 2 + 1) * kWave + lane] = make_double2(cY, pack_em(E, SF(0)));
    }
}

template <typename SF, int S>
__device__ __forceinline__ void load_rec(const SF* __restrict__ base, int lane, SF (&r)[4]) {
    using L = Seg<SF, S>;
    if constexpr (sizeof(SF) == 4) {
        const float4 f = reinterpret_cast<const float4*>(base)[(L::iC / 4) * kWave + lane];
        r[0] = f.x; r[1] = f.y; r[2] = f.z; r[3] = f.w;
    } else {
        const double2 a = reinterpret_cast<const double2*>(base)[(L::iC / 2) * kWave + lane];
        const double2 b = reinterpret_cast<const double2*>(base)[(L::iC / 2 + 1) * kWave + lane];
        r[0] = a.x; r[1] = a.y; r[2] = b.x; r[3] = b.y;
    }
}

// One tile in a block slot (idir -1: every panel's forward, Z, every panel's backward, the
// results), or one item of the wide pool, in the tile's own region of it: idir 2 runs the
// whole tile; idir 0 runs panel ipp's forward of a pipelined tile and leaves the pair's Z part
// and range flag; idir 1 runs panel ipp's backward from Z formed out of every panel's part,
// leaves its histogram, sums and flag, and the tile's last backward item to finish adds the
// parts of all panels, in panel order, into the results.  An item waits, bounded in real
// time, only for items queued before it: a tile's first item for the tiles whose pool
// regions it reuses to be done, a forward item for the previous panel's forward rows, a
// backward item for every forward item of its tile and for the next panel's backward rows.
// A wave that holds an item is running, so every wait ends.  A tile in the pool is done when
// its last item has made its writes visible at agent scope: another XCD's L2 then holds no
// dirty line of the region that a later tile reuses.
template <typename SF, int AM, int S, bool TR, bool BAND>
__device__ void soft_tile(const SoftArgs<SF>& A, const SoftView<SF> V, const int32_t* __restrict__ order,
                          const SoftTile* __restrict__ tdp, int64_t slot, const SF* __restrict__ spsub,
                          const SF* __restrict__ sgw,
                          double* __restrict__ hist, double* __restrict__ lacc, int lane, int tix, int ipp,
                          int idir) {
    using R = SoftRange<SF>;
    using LY = Seg<SF, S>;
    constexpr bool L = AM == kLocal, GL = AM == kGlocal;
    // this lane's pair slot
    struct {
        int32_t first, fsteps, npairs, panels;
    } td{tdp->first, tdp->fsteps, tdp->npairs, tdp->panels};
    int q = -1, G = 1, lo = lane;
    for (int k = 0; k < td.npairs; ++k) {
        const int lk = tdp->lo[k], gk = tdp->G[k];
        if (lane >= lk && lane < lk + gk) { q = k; G = gk; lo = lk; }
    }
    const int g = lane - lo;
    const int p = q >= 0 ? order[td.first + q] : -1;
    const int na = A.na, nn = na * na, gs = (TR ? A.K : 1) * (nn + 4);
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
    const int NP = td.panels;
    const int PW = G * S;   // columns of one panel of this pair
    // Banded (aligner.hpp:2644-2690): row r's cells are the columns [lo, hi] around the guide,
    // lo = max(1, g[r] - bw), hi = min(n, max(g[r] + bw, g[r+1] - 1 + bw)) (Global's row m
    // reaches n); column 0 of row r is in the band while g[r] <= bw (border_rows); the other
    // cells hold zero, in both directions.  Local's column-0 starts and Local / Glocal's row 0
    // stay as the full pass has them in the forward, but Local's Z counts a row's cells only
    // inside [jlo0, jhi0] (column 0 while g[r] <= bw; soft_kernel_impl.inl:289, 487), and the
    // backward holds zero outside every row's [jlo0, jhi0] (soft_kernel_impl.inl:826-828).
    const int32_t* bgj = nullptr;
    int bw = 0;
    if constexpr (BAND) {
        if (p >= 0) {
            bgj = A.guide + A.off_a[p] + p;
            bw = A.band < n ? A.band : n;
        }
    }
    auto band_of = [&](int r, int& blo, int& bhi) {
        const int g0 = bgj[r];
        blo = g0 - bw > 1 ? g0 - bw : 1;
        if (!L && !GL && r == m) {
            bhi = n;
            return;
        }
        int h = g0 + bw;
        if (r < m) {
            const int h1 = bgj[r + 1] - 1 + bw;
            h = h1 > h ? h1 : h;
        }
        bhi = h < n ? h : n;
    };
    // the panels whose forward and backward this call runs
    const bool whole = idir == -1 || idir == 2;
    const int pf0 = whole ? 0 : ipp, pf1 = whole ? NP : (idir == 0 ? ipp + 1 : ipp);
    const int pb0 = whole ? 0 : ipp, pb1 = whole ? NP : (idir == 1 ? ipp + 1 : ipp);
    // pipelined: the tile's words -- forward steps done per panel, backward steps done per
    // panel, forward items done, backward items done -- and the launch's error word
    int32_t* const sw = idir < 0 ? nullptr : A.lsync + tdp->sync;
    int32_t* const serr = A.lsync;
    const int64_t poff = (idir == 0 || idir == 1) ? tdp->poff : 0;
    // a tile's first item: the tiles whose pool regions it reuses are done
    if (idir == 2 || (idir == 0 && ipp == 0)) {
        const int w0 = tdp->wait0, nw = tdp->nwait;
        for (int k = 0; k < nw; ++k) bounded_wait(A.lsync + 1 + A.lwait[w0 + k], 1, serr, A.wait_ticks, lane, 1);
    }
    // set 0's gap weights (the untracked kernels read them as before, from the argument buffer)
    const SF oa = TR ? sgw[0] : A.gapw[0], ea = TR ? sgw[1] : A.gapw[1];
    const SF ob = TR ? sgw[2] : A.gapw[2], eb = TR ? sgw[3] : A.gapw[3];
    // Tracks: the pair's row sets.  Row i's profile is set ts's (profile row ts * (na + 1) +
    // its letter), its gaps in B (ob, eb) set ts's, its gaps in A (oa, ea) set gs's.
    const uint16_t* tk = A.trk;
    const uint16_t* gk = A.gtrk;
    if constexpr (TR) {
        if (p >= 0) {
            tk += A.off_a[p];
            gk += A.off_a[p] + p;
        }
    }
    const int KS = TR ? A.K : 1;   // sets
    SF oa0 = oa, ea0 = ea;         // row 0's gaps in A
    if constexpr (TR) {
        if (p >= 0) {
            const int g0 = gk[0];
            oa0 = sgw[4 * g0];
            ea0 = sgw[4 * g0 + 1];
        }
    }
    // Panel pp's row i is row index rbase + pp * rows + i of the pool: a block slot (the
    // pool's row count per panel), or a pipelined tile's own region (its own rows).
    const bool own = idir >= 0;
    const int64_t rows = own ? (int64_t)td.fsteps + 1 : V.rows;
    const int64_t rbase = own ? tdp->rbase : slot * (int64_t)V.panels * V.rows;
    // panel boundary column of row i of panel pp, this pair: 6 values, 2 exponents
    auto bnd = [&](int pp, int i) {
        return V.pbnd + ((rbase + pp * rows + i) * kTilePairs + (q < 0 ? 0 : q)) * 6;
    };
    auto bex = [&](int pp, int i) {
        return V.pexp + ((rbase + pp * rows + i) * kTilePairs + (q < 0 ? 0 : q)) * 2;
    };
    bool bad = false;

    // ── forward ──────────────────────────────────────────────────────────────
    SF M[S], X[S], Y[S], T[S];               // row i-1 of the lane's columns, T = M + X + Y
    SF cM = SF(0), cX = SF(0), cY = SF(0);   // its boundary column j0 (column 0 at j0 = 0)
    int E = 0;                               // the units of M, X, Y, T, c*
    ExpAcc zacc{0.0, 0};
    SF corner = SF(0);                       // Global: the (m, n) cell, held by the lane of column n
    int cornerE = 0;
    for (int pp = pf0; pp < pf1; ++pp) {
        // Pipelined: panel 0 waits until the tile that last used this scratch slot is done,
        // panel pp > 0 for the previous panel's first steps (and, every kPipeChunk steps, for
        // the next chunk's).  The boundary row that lane 0 reads at step t was written by the
        // previous panel's last lane (G <= 64) at its step t + G, so a chunk of steps
        // t .. t+kPipeChunk-1 needs t + kPipeChunk + 64 of its steps done.
        int32_t* fin = nullptr;
        int32_t* fout = nullptr;
        if (idir == 0) {
            if (pp > 0) {
                fin = sw + pp - 1;
                bounded_wait(fin, min(kPipeChunk + kWave, td.fsteps), serr, A.wait_ticks, lane, 1);
            }
            if (pp + 1 < NP) fout = sw + pp;
        }
        const int j0 = pp * PW + g * S;
        const int ncols = n - j0 < 0 ? 0 : (n - j0 > S ? S : n - j0);
        const bool live = p >= 0 && (ncols > 0 || j0 == 0);   // the lane of column 0 always computes
        int coff[S];
#pragma unroll
        for (int k = 0; k < S; ++k) coff[k] = (k < ncols) ? (int)bp[j0 + k] : na;   // na: no column
        SF* __restrict__ seg = V.seg + (rbase + pp * rows) * LY::row;
        E = 0;
        if (L || GL) {
#pragma unroll
            for (int k = 0; k < S; ++k) { M[k] = (k < ncols) ? SF(1) : SF(0); X[k] = SF(0); Y[k] = SF(0); }
            cM = SF(1); cX = SF(0); cY = SF(0);
        } else {
            // Row 0 of Global: M(0,0) = 1, Y(0,j) = oa * ea^(j-1) (soft_kernel_impl.inl:740-744),
            // formed in double with a separate exponent (repeated squaring from the segment's
            // first column) so that far columns keep their own range.
            if (j0 == 0) {
                cM = SF(1); cX = SF(0); cY = SF(0);
            }
            // v = oa * ea^(j0 - 1) as vm * 2^ve (the boundary column; column 1 when j0 == 0)
            int ve = 0, be = 0;
            double vm = (double)oa0, bm = (double)ea0;
            if (j0 > 0) {
                for (int r = j0 - 1; r > 0; r >>= 1) {
                    if (r & 1) {
                        vm *= bm;
                        const int k1 = __builtin_amdgcn_frexp_exp(vm);
                        vm = __builtin_amdgcn_ldexp(vm, -k1);
                        ve += be + k1;
                    }
                    bm *= bm;
                    const int k2 = __builtin_amdgcn_frexp_exp(bm);
                    bm = __builtin_amdgcn_ldexp(bm, -k2);
                    be = 2 * be + k2;
                }
            }
            // Y of the S columns: the boundary's value times ea^(k+1) (column 1 at j0 == 0: oa)
            double ym[S];
            int yex[S];
            {
                double cm = vm;
                int ce = ve;
                const double em = (double)ea0;
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    if (j0 > 0 || k > 0) {
                        cm *= em;
                        const int k1 = __builtin_amdgcn_frexp_exp(cm);
                        cm = __builtin_amdgcn_ldexp(cm, -k1);
                        ce += k1;
                    }
                    ym[k] = cm;
                    yex[k] = ce;
                }
            }
            if (j0 > 0) {
                // the segment's exponent: that of its largest value (the boundary or the last column)
                const int last = ncols > 0 ? ncols - 1 : 0;
                const int kb = __builtin_amdgcn_frexp_exp(vm), kl = __builtin_amdgcn_frexp_exp(ym[last]);
                E = (ve + kb > yex[last] + kl ? ve + kb : yex[last] + kl) - 1;
                cM = SF(0); cX = SF(0);
                cY = (SF)__builtin_amdgcn_ldexp(vm, ve - E);
            }
#pragma unroll
            for (int k = 0; k < S; ++k) {
                M[k] = SF(0); X[k] = SF(0);
                Y[k] = (k < ncols) ? (SF)__builtin_amdgcn_ldexp(ym[k], yex[k] - E) : SF(0);
            }
            if constexpr (BAND) {   // row 0's border runs up to its band's end (border_cols)
                if (p >= 0) {
                    int blo, bhi;
                    band_of(0, blo, bhi);
#pragma unroll
                    for (int k = 0; k < S; ++k)
                        if (j0 + k + 1 > bhi) Y[k] = SF(0);
                    if (j0 > bhi) cY = SF(0);
                }
            }
        }
        SF z = SF(0);
#pragma unroll
        for (int k = 0; k < S; ++k) {
            T[k] = M[k] + X[k] + Y[k];
            if (k < ncols) z = dmax(z, T[k]);
        }
        if (j0 == 0) z = dmax(z, cM + cX + cY);
        if (live) {
            if (L || (GL && m == 0)) {
                SF s = j0 == 0 ? cM + cX + cY : SF(0);
                int zlo = 1, zhi = n;   // banded: row 0's cells in [jlo0(0), jhi0(0)]
                if constexpr (BAND) {
                    band_of(0, zlo, zhi);
                    if (!(bgj[0] <= bw)) s = SF(0);
                }
#pragma unroll
                for (int k = 0; k < S; ++k) s += (k < ncols && j0 + k + 1 >= zlo && j0 + k + 1 <= zhi) ? T[k] : SF(0);
                expacc_add(zacc, (double)s, E);
            }
            store_seg<SF, S>(seg + (int64_t)g * LY::row, lane, M, X, Y, cM, cX, cY, E, z);
        }
        SF oM = M[S - 1], oX = X[S - 1], oY = Y[S - 1];
        int oE = E;
        int kpend = 0;                                                // next row's units: E + kpend
        // the lane's last row held no mass (Local lanes always hold their start term, except
        // outside a band, where they take a neighbour's units as the other modes' do)
        bool fzero = (!L || BAND) && !(z > SF(0)) && !(cM + cX + cY > SF(0));
        const bool pin = g == 0 && j0 > 0;                            // reads the previous panel's column
        const bool pout = g == G - 1 && pp + 1 < NP;                  // writes this panel's last column
        // the codes of the lane's next three rows (loaded three rows ahead: waiting for a load
        // also waits for the stores issued before it, and a row step stores its row)
        int acode = (live && m >= 1) ? (int)ap[0] : 0;
        int acode2 = (live && m >= 2) ? (int)ap[1] : 0;
        int acode3 = (live && m >= 3) ? (int)ap[2] : 0;
        int tnext = 0, gnext = 0;                                     // and (tracks) its sets
        if constexpr (TR) {
            if (live && m >= 1) {
                tnext = tk[0];
                gnext = gk[1];
            }
        }
        SF qM = SF(0), qX = SF(0), qY = SF(0);
        int qE = 0;
        if (pin && live && m >= 1) {
            const SF* bq = bnd(pp - 1, 1);
            qM = bq[0]; qX = bq[1]; qY = bq[2]; qE = bex(pp - 1, 1)[0];
        }
        const int fsteps = td.fsteps;
        // One wavefront step (row i = t - g + 1 on lane g); the loop runs it twice per trip,
        // which lets the compiler keep the step's values in place instead of copying them
        // at the loop's back edge.
        // chk false: a step at which every lane of every pair is on one of its rows (the tile's
        // middle steps), run without the per-lane branch; the lanes beyond a pair's columns and
        // the lanes of no pair then compute too, on their own registers and their own rows of
        // the slot, and only their writes others read (the panel column, Z, the range flag)
        // are held back.
        auto fstep = [&](const int t, auto chk) {
            constexpr bool CHECK = decltype(chk)::value;
            const int i = t - g + 1;
            SF lM = wshift_up(oM), lX = wshift_up(oX), lY = wshift_up(oY);
            int lE = wshift_up(oE);
            if (!CHECK || (live && i >= 1 && i <= m)) {
                int ac = acode;
                acode = acode2;
                acode2 = acode3;
                if (i + 2 < m) acode3 = (int)ap[i + 2];   // three rows ahead
                // the row's weights: its sets' (tracks), else the constants
                SF r_oa = oa, r_ea = ea, r_ob = ob, r_eb = eb;
                if constexpr (TR) {
                    const int ts = tnext, gs = gnext;
                    if (i < m) {
                        tnext = tk[i];
                        gnext = gk[i + 1];
                    }
                    ac += ts * (na + 1);
                    r_oa = sgw[4 * gs]; r_ea = sgw[4 * gs + 1];
                    r_ob = sgw[4 * ts + 2]; r_eb = sgw[4 * ts + 3];
                }
                if (pin) {
                    lM = qM; lX = qX; lY = qY; lE = qE;
                    if (i < m) {
                        const SF* bq = bnd(pp - 1, i + 1);
                        qM = bq[0]; qX = bq[1]; qY = bq[2]; qE = bex(pp - 1, i + 1)[0];
                    }
                }
                int Ec = E + kpend;                 // this row's units
                SF s = sldexp(SF(1), -kpend);       // the previous row's values into them
                SF nM, nX, nY;
                if (j0 != 0) {
                    if (fzero) {
                        // no mass yet: take the neighbour's units, normalised to its boundary values
                        const SF ml = dmax(lM, dmax(lX, lY));
                        Ec = ml > SF(0) ? lE + silogb(ml) : lE;
                        s = SF(1);
                    } else {
                        // Headroom.  Untracked, on the neighbour's units: every row's gap chain
                        // decays at the same rate, so the units a lane carries from its previous
                        // row keep pace with the row.  Tracked, on the boundary's magnitude
                        // (above its units when its values exceed 1): a row whose gap weights
                        // decay slower than the row before's outgrows the carried units a few
                        // lanes along it.
                        int lmag = lE;
                        if constexpr (TR) {
                            const SF ml = dmax(lM, dmax(lX, lY));
                            if (ml > SF(1)) lmag += silogb(ml);
                        }
                        if (lmag - R::headroom > Ec) {
                            Ec = lmag - R::headroom;
                            s = sldexp(SF(1), E - Ec);
                        }
                    }
                    nM = sldexp(lM, lE - Ec); nX = sldexp(lX, lE - Ec); nY = sldexp(lY, lE - Ec);
                }
                const SF ebs = r_eb * s, obs = r_ob * s;
                if (j0 == 0) {   // column 0 of row i (soft_kernel_impl.inl:669-674)
                    if constexpr (L) { nM = sldexp(SF(1), -Ec); nX = SF(0); nY = SF(0); }
                    else {
                        nM = SF(0); nY = SF(0); nX = sfma(cX, ebs, (cM + cY) * obs);
                        if constexpr (BAND) {
                            if (!(bgj[i] <= bw)) nX = SF(0);
                        }
                    }
                }
                int blo = 1, bhi = n;
                if constexpr (BAND) band_of(i, blo, bhi);
                const SF fr = L ? sldexp(SF(1), -Ec) : SF(0);
                const SF* __restrict__ prow = spsub + ac * (na + 1);
                SF pw[S];
#pragma unroll
                for (int k = 0; k < S; ++k) pw[k] = prow[coff[k]];
                SF dT = cM + cX + cY;   // the diagonal of column j0+1: row i-1, column j0
                SF lm = nM, lx = nX, ly = nY;
                SF mx = (j0 == 0) ? nM + nX + nY : SF(0);
                const int nc = opaque(ncols);
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    const SF uM = M[k], uX = X[k], uY = Y[k];
                    SF mv = sfma(dT, pw[k] * s, fr);
                    SF xv = sfma(uX, ebs, (uM + uY) * obs);
                    SF yv = sfma(ly, r_ea, (lm + lx) * r_oa);
                    if constexpr (BAND) {
                        if (j0 + k + 1 < blo || j0 + k + 1 > bhi) { mv = SF(0); xv = SF(0); yv = SF(0); }
                    }
                    dT = T[k];
                    const SF tv = mv + xv + yv;
                    M[k] = mv; X[k] = xv; Y[k] = yv; T[k] = tv;
                    lm = mv; lx = xv; ly = yv;
                    mx = dmax(mx, tv);
                }
                cM = nM; cX = nX; cY = nY;
                if (live && !(mx <= R::fmax)) bad = true;
                if (live && sizeof(SF) == 4 && (Ec >= kPackEMax || Ec <= -kPackEMax)) bad = true;   // no room to pack E
                E = Ec;
                kpend = 0;
                if (mx > R::hi || (mx < R::lo && mx > SF(0))) {
                    const int k2 = silogb(mx);
                    kpend = k2 > R::kmax ? R::kmax : (k2 < -R::kmax ? -R::kmax : k2);
                }
                fzero = (!L || BAND) && !(mx > SF(0)) && !(cM + cX + cY > SF(0));
                if (live && (L || (GL && i == m))) {
                    SF sr = j0 == 0 ? cM + cX + cY : SF(0);
                    if constexpr (BAND) {   // column 0 counts inside the band only (jlo0)
                        if (j0 == 0 && !(bgj[i] <= bw)) sr = SF(0);
                    }
#pragma unroll
                    for (int k = 0; k < S; ++k) sr += (k < nc) ? T[k] : SF(0);
                    expacc_add(zacc, (double)sr, E);
                }
                // Rows the backward reaches at the last step of a band of kCK steps are stored in
                // full, the others as a boundary record from which the backward recomputes them
                // (row 0 always in full).
                if ((m + G - 1 - g - i) % kCK == kCK - 1) store_seg<SF, S>(seg + (int64_t)(i + g) * LY::row, lane, M, X, Y, cM, cX, cY, E, mx);
                else store_rec<SF, S>(seg + (int64_t)(i + g) * LY::row, lane, cM, cX, cY, E);
                oM = M[S - 1]; oX = X[S - 1]; oY = Y[S - 1]; oE = E;
                if (pout && live) {
                    SF* bo = bnd(pp, i);
                    bo[0] = oM; bo[1] = oX; bo[2] = oY;
                    bex(pp, i)[0] = oE;
                }
            }
        };
        // Pipelined: every kPipeChunk steps, publish the steps done and wait for the next
        // chunk's input -- one uniform comparison per trip (sync_at: the next such step).
        int sync_at = (fin != nullptr || fout != nullptr) ? kPipeChunk : 0x7fffffff;
        auto sync = [&](int t) {
            if (fout != nullptr) publish(fout, t, lane);
            if (fin != nullptr) bounded_wait(fin, min(t + kPipeChunk + kWave, fsteps), serr, A.wait_ticks, lane, 1);
            sync_at += kPipeChunk;
        };
        // The middle steps [t0, t1): every pair of the tile on a row of every lane, from the
        // widest pair's last lane's row 1 to the shortest pair's first lane's row m (both ends
        // even, so the sync test, which meets only even steps, sees every chunk boundary).
        // Banded tiles keep the per-lane branch throughout (a lane of no pair has no guide).
        int t0 = 0, t1 = 0;
        if constexpr (!BAND) {
            int gq = p >= 0 ? G : 0, mq = p >= 0 ? m : 0x7fffffff;
            for (int off = 32; off > 0; off >>= 1) {
                gq = max(gq, __shfl_xor(gq, off, kWave));
                mq = min(mq, __shfl_xor(mq, off, kWave));
            }
            t0 = (wave_uniform(gq) - 1 + 1) & ~1;
            t1 = wave_uniform(mq) & ~1;
            if (t1 > fsteps) t1 = fsteps & ~1;
            if (t1 < t0) t1 = t0;
        }
        const std::true_type checked;
        const std::false_type unchecked;
        int t = 0;
        for (; t < t0; t += 2) {
            if (t == sync_at) sync(t);
            fstep(t, checked);
            if (t + 1 < fsteps) fstep(t + 1, checked);
        }
        for (; t < t1; t += 2) {
            if (t == sync_at) sync(t);
            fstep(t, unchecked);
            fstep(t + 1, unchecked);
        }
        for (; t + 1 < fsteps; t += 2) {
            if (t == sync_at) sync(t);
            fstep(t, checked);
            fstep(t + 1, checked);
        }
        if (t < fsteps) {
            if (t == sync_at) sync(t);
            fstep(t, checked);
        }
        if (fout != nullptr) publish(fout, fsteps, lane);
        // Global's end cell: row m of column n, in the panel that holds column n (a narrow
        // pair in a wide tile ends before the last panel)
        if constexpr (!L && !GL) {
            if (p >= 0 && ((n == 0) ? (j0 == 0) : (j0 < n && n <= j0 + S))) {
                if (n == 0) corner = cM + cX + cY;
                else {
#pragma unroll
                    for (int k = 0; k < S; ++k)
                        if (k == n - 1 - j0) corner = T[k];
                }
                cornerE = E;
            }
        }
        if (NP > 1) __threadfence();   // the panel's boundary column is read by the next panel
    }

    // ── Z ───────────────────────────────────────────────────────────────────
    if (idir == 1) {
        // every panel's part, in panel order, once every forward item of the tile is done
        bounded_wait(sw + 2 * NP, NP, serr, A.wait_ticks, lane, 1);
        ExpAcc z{0.0, 0};
        if (q >= 0)
            for (int k = 0; k < NP; ++k) {
                const double* zp = A.lzpart + ((poff + k) * kTilePairs + q) * 4;
                expacc_add(z, zp[0], (int)zp[1]);
            }
        zacc = z;
    } else if constexpr (!L && !GL) {
        const int gn = n == 0 ? 0 : ((n - 1) % PW) / S;   // the lane of column n
        zacc.acc = __shfl((double)corner, lo + gn, kWave);
        zacc.E = __shfl(cornerE, lo + gn, kWave);
    } else {
        for (int off = 1; off < kWave; off <<= 1) {
            const double oacc = __shfl_down(zacc.acc, off, kWave);
            const int oe = __shfl_down(zacc.E, off, kWave);
            if (g + off < G) expacc_add(zacc, oacc, oe);
        }
        zacc.acc = __shfl(zacc.acc, lo, kWave);
        zacc.E = __shfl(zacc.E, lo, kWave);
    }
    if (idir == 0) {
        // a forward item: the pair's Z part (this panel's) and range flag, then the tile's count
        int fb = bad ? 1 : 0;
        for (int off = 1; off < kWave; off <<= 1) {
            const int o = __shfl_down(fb, off, kWave);
            if (g + off < G) fb |= o;
        }
        if (g == 0 && q >= 0) {
            double* zp = A.lzpart + ((poff + ipp) * kTilePairs + q) * 4;
            zp[0] = zacc.acc;
            zp[1] = (double)zacc.E;
            zp[2] = (double)fb;
        }
        __threadfence();
        if (lane == 0) __hip_atomic_fetch_add(sw + 2 * NP, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const bool zbad = !(zacc.acc > 0.0) || !(zacc.acc <= 1.7976931348623157e308);
    // normalise: Z = zr * 2^ze, zr in [1, 2)
    int ze = zacc.E;
    double zr = zacc.acc;
    if (!zbad) {
        const int kz = __builtin_amdgcn_frexp_exp(zr) - 1;
        zr = __builtin_amdgcn_ldexp(zr, -kz);
        ze += kz;
    }
    const double izr = zbad ? 0.0 : 1.0 / zr;

    // ── backward ────────────────────────────────────────────────────────────
    // The lane's running sums: the gap counts oa, ea, ob, eb and the lost-mass bound.  Untracked
    // they stay in registers; tracked, the K x 4 counts live in LDS (per lane, added without
    // return), since the set a row's counts go to varies along the pair.
    double* __restrict__ acc = lacc + lane;
    double racc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if constexpr (TR) {
#pragma unroll
        for (int c = 0; c < 4 * KS + 1; ++c) acc[c * kWave] = 0.0;
    }
    const int kmin = (L || GL) ? 1 : 0;   // first row whose Y opens count
    const bool col0_opens = !L;           // column-0 X opens count (kmin0 == 0)
    const int hrow = na + 1;   // histogram rows carry a spill column for the no-column letter
    double* __restrict__ hq = hist + (q < 0 ? 0 : q) * KS * na * hrow;   // K sets of na rows
    // Rows run m .. 0 at steps u = G-1-g .. G-1-g+m; the step count is rounded up to two
    // bands (steps in which no lane is active) so that the loop body holds both parities.
    const int bsteps = (td.fsteps + 2 * kCK) / (2 * kCK) * (2 * kCK);
    for (int pp = pb1 - 1; pp >= pb0; --pp) {
        // Pipelined: the next panel's backward steps, as the forward waits (mirrored: the
        // row that lane G-1 reads at step u was written by the next panel's lane 0 at its
        // step u + G).
        int32_t* bin = nullptr;
        int32_t* bout = nullptr;
        if (idir == 1) {
            if (pp + 1 < NP) {
                bin = sw + NP + pp + 1;
                bounded_wait(bin, min(kPipeChunk + kWave, bsteps), serr, A.wait_ticks, lane, 1);
            }
            if (pp > 0) bout = sw + NP + pp;
        }
        const int j0 = pp * PW + g * S;
        const int ncols = n - j0 < 0 ? 0 : (n - j0 > S ? S : n - j0);
        const bool live = p >= 0 && (ncols > 0 || j0 == 0);
        int coff[S + 1];   // the lane's column letters and the next lane's first
#pragma unroll
        for (int k = 0; k < S; ++k) coff[k] = (k < ncols) ? (int)bp[j0 + k] : na;   // na: no column
        coff[S] = (j0 + S < n) ? (int)bp[j0 + S] : na;
        const SF* __restrict__ seg = V.seg + (rbase + pp * rows) * LY::row;
        // Row i+1 of the lane's columns.  Row m starts from a virtual row m+1 whose profile
        // is 1 (letter index na) and whose BM holds the end weights one column to the right,
        // so that row m's d = P BM(m+1, j+1) is its end weight: Global 1 at column n only,
        // Glocal 1 at every column 0..n (soft_kernel_impl.inl:815-818); Local's end weight is
        // the init term of every row.
        SF BM[S], BX[S];
        SF rdM = SF(0);                // BM(i+1, j0+S+1)
#pragma unroll
        for (int k = 0; k < S; ++k) {
            BX[k] = SF(0);
            BM[k] = SF(0);
            if constexpr (GL) BM[k] = (j0 + k <= n) ? SF(1) : SF(0);
            else if constexpr (!L) BM[k] = (j0 + k == n) ? SF(1) : SF(0);
        }
        if constexpr (GL) rdM = (j0 + S <= n) ? SF(1) : SF(0);
        else if constexpr (!L) rdM = (j0 + S == n) ? SF(1) : SF(0);
        SF c0M = SF(0), c0X = SF(0);   // column 0 of row i+1 (the lane of column 0)
        int EB = 0, kpend = 0;         // units of BM, BX, c0*, rdM; next row's: EB + kpend
        bool bzero = true;             // the lane's last backward row held no mass
        const bool owns_end = (n == 0) ? (j0 == 0) : (j0 < n && n <= j0 + S);
        const bool has_right = j0 + S < n;
        const bool pin = g == G - 1 && pp + 1 < NP;   // reads the next panel's column
        const bool pout = g == 0 && pp > 0;           // writes this panel's first column
        // Forward rows.  Step u handles row i = m+G-1-g-u; steps form bands of kCK (u / kCK, the
        // same for every lane).  The row met at a band's last step (and row 0) is stored in
        // full; the band's other rows are recomputed from it, forward, through their boundary
        // records.  Two buffers alternate with the band's parity: S holds the band's full row
        // (loaded during the previous band), W receives the recomputed rows.
        SF Fa[LY::vals], Fb[LY::vals];
        SF rec[kCK - 1][4];     // records of the rows above the band's full row, in order
        int rcode[kCK - 1];     // their profile rows
        int rsets[kCK - 1];     // (tracks) their sets: ts | gs << 16
#pragma unroll
        for (int j = 0; j < kCK - 1; ++j) {
            rec[j][0] = rec[j][1] = rec[j][2] = rec[j][3] = SF(0);
            rcode[j] = 0;
            rsets[j] = 0;
        }
        auto load_recs = [&](int src, int last) {   // the records of rows src+1 .. last
#pragma unroll
            for (int j = 0; j < kCK - 1; ++j) {
                const int r = src + 1 + j;
                if (r <= last) {
                    load_rec<SF, S>(seg + (int64_t)(r + g) * LY::row, lane, rec[j]);
                    rcode[j] = (int)ap[r - 1];
                    if constexpr (TR) {
                        const int ts = tk[r - 1];
                        rcode[j] += ts * (na + 1);
                        rsets[j] = ts | (int)gk[r] << 16;
                    }
                }
            }
        };
        if (live) {   // the first band: its full row (row 0 when the band reaches it) and records
            const int u0 = G - 1 - g;
            const int c0 = m + u0 % kCK - (kCK - 1);
            const int src = c0 > 0 ? c0 : 0;
            if (((u0 / kCK) & 1) == 0) load_seg<SF, S>(seg + (int64_t)(src + g) * LY::row, lane, Fa);
            else load_seg<SF, S>(seg + (int64_t)(src + g) * LY::row, lane, Fb);
            load_recs(src, m);
        }
        SF qM = SF(0), qY = SF(0);
        int qE = 0;
        if (pin && live) {
            const SF* bq = bnd(pp + 1, m);
            qM = bq[3]; qY = bq[4]; qE = bex(pp + 1, m)[1];
        }
        int acode = na;                // code of row i+1 (row m: the virtual row)
        int acode2 = (live && m >= 1) ? (int)ap[m - 1] : na;   // and of rows i, i-1 (loaded ahead)
        int acode3 = (live && m >= 2) ? (int)ap[m - 2] : na;
        // (tracks) row i+1's set, and row i's sets, loaded one row ahead
        int btn = 0, bti = 0, bgi = 0;
        if constexpr (TR) {
            if (live) {
                bti = m >= 1 ? (int)tk[m - 1] : 0;
                bgi = gk[m];
            }
        }
        SF pubM = SF(0), pubY = SF(0);
        int pubE = 0;

        // Recompute forward row i (i >= 1) into Fd from row i-1 in Fs and row i's record: the
        // forward's row step (same operations, same order) in the record's units.
        auto recompute = [&](const SF(&Fs)[LY::vals], SF(&Fd)[LY::vals], const SF(&rc)[4], int code, int sets, int r) {
            const int Ec = em_E(rc[3]);
            const SF s = sldexp(SF(1), em_E(Fs[LY::iE]) - Ec);   // row i-1's values into row i's units
            SF r_oa = oa, r_ea = ea, r_ob = ob, r_eb = eb;
            if constexpr (TR) {
                const int ts = sets & 0xffff, gs = sets >> 16;
                r_oa = sgw[4 * gs]; r_ea = sgw[4 * gs + 1];
                r_ob = sgw[4 * ts + 2]; r_eb = sgw[4 * ts + 3];
            }
            const SF ebs = r_eb * s, obs = r_ob * s;
            const SF fr = L ? sldexp(SF(1), -Ec) : SF(0);
            const SF* __restrict__ prow = spsub + code * (na + 1);
            SF pw[S];
#pragma unroll
            for (int k = 0; k < S; ++k) pw[k] = prow[coff[k]];
            SF dT = Fs[LY::iC] + Fs[LY::iC + 1] + Fs[LY::iC + 2];
            SF lm = rc[0], lx = rc[1], ly = rc[2];
            SF mx = (j0 == 0) ? rc[0] + rc[1] + rc[2] : SF(0);
            const int nc = opaque(ncols);
            int blo = 1, bhi = n;
            if constexpr (BAND) band_of(r, blo, bhi);
#pragma unroll
            for (int k = 0; k < S; ++k) {
                const SF uM = Fs[k], uX = Fs[LY::iX + k], uY = Fs[LY::iY + k];
                SF mv = sfma(dT, pw[k] * s, fr);
                SF xv = sfma(uX, ebs, (uM + uY) * obs);
                SF yv = sfma(ly, r_ea, (lm + lx) * r_oa);
                if constexpr (BAND) {
                    if (j0 + k + 1 < blo || j0 + k + 1 > bhi) { mv = SF(0); xv = SF(0); yv = SF(0); }
                }
                dT = uM + uX + uY;
                const SF tv = mv + xv + yv;
                Fd[k] = mv; Fd[LY::iX + k] = xv; Fd[LY::iY + k] = yv;
                lm = mv; lx = xv; ly = yv;
                mx = dmax(mx, tv);
            }
            Fd[LY::iC] = rc[0]; Fd[LY::iC + 1] = rc[1]; Fd[LY::iC + 2] = rc[2];
            Fd[LY::iE] = pack_em(Ec, mx);
        };
        auto copy_row = [&](const SF(&Fs)[LY::vals], SF(&Fd)[LY::vals]) {
#pragma unroll
            for (int k = 0; k < LY::vals; ++k) Fd[k] = Fs[k];
        };
        // One backward step on row i: Fc holds forward row i.
        auto step = [&](int i, SF rM, SF rY, int rE, SF(&Fc)[LY::vals]) {
            const int ac = acode;
            acode = acode2;
            acode2 = acode3;
            if (i >= 3) acode3 = (int)ap[i - 3];   // three rows ahead
            // The step's weights.  Tracks: row i+1's set tn (its profile, its gaps in B, its
            // match counts and X opens), row i's sets ti (its X moves) and gi (its gaps in A).
            int tn = 0, ti = 0, gi = 0;
            SF r_oa = oa, r_ea = ea, r_ob = ob, r_eb = eb;
            if constexpr (TR) {
                tn = btn;
                ti = bti;
                gi = bgi;
                bti = i >= 2 ? (int)tk[i - 2] : 0;
                if (i >= 1) bgi = gk[i - 1];
                btn = ti;
                r_oa = sgw[4 * gi]; r_ea = sgw[4 * gi + 1];
                r_ob = sgw[4 * tn + 2]; r_eb = sgw[4 * tn + 3];
            }
            if (pin) {
                rM = qM; rY = qY; rE = qE;
                if (i >= 1) {
                    const SF* bq = bnd(pp + 1, i - 1);
                    qM = bq[3]; qY = bq[4]; qE = bex(pp + 1, i - 1)[1];
                }
            }
            const SF* __restrict__ pn = spsub + (ac + tn * (na + 1)) * (na + 1);   // row i+1's profile
            SF pw[S + 1];
#pragma unroll
            for (int k = 0; k <= S; ++k) pw[k] = pn[coff[k]];
            const int efc = em_E(Fc[LY::iE]);
            const SF mfc = em_max(Fc[LY::iE]);   // a power of two at least the segment's maximum
            // This row's units.  A segment without mass of its own (Global's row m left of
            // column n, or rows whose mass has not arrived) takes the neighbour's exponent; one
            // whose neighbour's values would not fit moves up to them minus the headroom.
            const bool own_zero = (i == m) ? (!L && !GL && !owns_end) : ((!L || BAND) && bzero);
            int Ec = (i == m) ? 0 : EB + kpend;
            SF s = (i == m) ? SF(1) : sldexp(SF(1), -kpend);   // row i+1's values into them
            bool adopt = false;
            if (has_right && own_zero) {
                // no mass yet: take the neighbour's units, normalised to its boundary values
                // (the lane's own row i+1 values are zero; the one it carries from the right
                // neighbour, BM(i+1, j0+S+1), is converted on its own below)
                const SF mr = dmax(rM, rY);
                Ec = mr > SF(0) ? rE + silogb(mr) : rE;
                s = SF(1);
                adopt = true;
            } else if (has_right) {
                // headroom, as the forward's (tracked: on the boundary's magnitude)
                int rmag = rE;
                if constexpr (TR) {
                    const SF mr = dmax(rM, rY);
                    if (mr > SF(1)) rmag += silogb(mr);
                }
                if (rmag - R::headroom > Ec) {
                    Ec = rmag - R::headroom;
                    s = sldexp(SF(1), EB - Ec);
                }
            }
            const SF rYc = has_right ? sldexp(rY, rE - Ec) : SF(0);   // BY(i, j0+S+1)
            const SF rMc = has_right ? sldexp(rM, rE - Ec) : SF(0);   // BM(i, j0+S+1)
            const SF init = L ? sldexp(SF(1), -Ec) : SF(0);
            const SF obs = r_ob * s, ebs = r_eb * s;
            // posterior factors: of row i's units, and of row i+1's (the late sums); none of
            // the late sums exist at row m
            // Banded: a lane whose forward or backward segment is empty (outside the band) keeps
            // borrowed units in that direction, whose sum with the other's can exceed the double
            // range while every product it scales is an exact zero; the factors are held to
            // 2^1000 there (a product that is not zero has values in range and a factor far
            // below it), so that no 0 x inf appears.
            int xnew = efc + Ec - ze, xold = efc + EB - ze;
            if constexpr (BAND) {
                xnew = xnew < 1000 ? xnew : 1000;
                xold = xold < 1000 ? xold : 1000;
            }
            const double fnew = zbad ? 0.0 : __builtin_amdgcn_ldexp(izr, xnew);
            const double fold = (zbad || i == m) ? 0.0 : __builtin_amdgcn_ldexp(izr, xold);
            const double fmatch = i == m ? 0.0 : fnew;
            const int nc = opaque(ncols);
            int blo = 1, bhi = n;
            if constexpr (BAND) band_of(i, blo, bhi);
            SF byr = rYc;   // BY(i, j+1), right to left
            SF dn = adopt ? sldexp(rdM, EB - Ec) : rdM;   // BM(i+1, j+1), row i+1's units (adopting: row i's)
            SF mx = SF(0);
            SF sx = SF(0), sy = SF(0), so = SF(0), sob = SF(0);
            // match posteriors go into the pair's histogram row a_{i+1} (row m: none, added as 0)
            double* __restrict__ h = hq + (ac < na ? tn * na + ac : 0) * hrow;
#pragma unroll
            for (int kk = 0; kk < S; ++kk) {
                const int k = S - 1 - kk;
                const SF pd = (pw[k + 1] * s) * dn;   // P(i+1, j+1) BM(i+1, j+1), row i's units
                const SF d = L ? pd + init : pd;
                const SF v = BX[k];                    // BX(i+1, j), row i+1's units
                const SF b1 = sfma(v, obs, d);
                const SF b2 = sfma(v, ebs, d);
                // row i+1's sums on forward row i (soft_kernel_impl.inl:883-884): X opens into
                // (i+1, j), and the match posterior of (i+1, j+1)
                sob = sfma(Fc[k] + Fc[LY::iY + k], v, sob);
                atomicAdd(h + coff[k + 1], (double)((Fc[k] + Fc[LY::iX + k] + Fc[LY::iY + k]) * pd) * fmatch);
                const SF byn = sfma(r_ea, byr, b1);
                const SF hh = byr * r_oa;
                // Columns past len(B) hold exact zeros in Global and Glocal: their profile entry is
                // 0 and no end weight lies there.  Local's start term reaches every cell, so its
                // columns past len(B) are cleared.
                bool in = !L || k < nc;
                if constexpr (BAND) in = in && j0 + k + 1 >= blo && j0 + k + 1 <= bhi;
                const SF bmn = in ? b1 + hh : SF(0);
                const SF bxn = in ? b2 + hh : SF(0);
                const SF byv = in ? byn : SF(0);
                dn = BM[k];
                BM[k] = bmn; BX[k] = bxn;
                if (k == 0) pubY = byv;
                byr = byv;
                mx = dmax(mx, dmax(bmn, dmax(bxn, byv)));
                // row i's sums (soft_kernel_impl.inl:879-882)
                sy = sfma(Fc[LY::iY + k], byv, sy);
                const SF fml = (k == 0) ? Fc[LY::iC] + Fc[LY::iC + 1] : Fc[k - 1] + Fc[LY::iX + k - 1];
                so = sfma(fml, byv, so);
                sx = sfma(Fc[LY::iX + k], bxn, sx);
            }
            if (j0 == 0) {   // column 0 (soft_kernel_impl.inl:849-857)
                const SF pd = (pw[0] * s) * dn;   // P(i+1, 1) BM(i+1, 1)
                const SF d = L ? pd + init : pd;
                const SF v = c0X;
                const SF b1 = sfma(v, obs, d);
                const SF b2 = sfma(v, ebs, d);
                if (col0_opens) sob = sfma(Fc[LY::iC] + Fc[LY::iC + 2], v, sob);
                atomicAdd(h + coff[0], (double)((Fc[LY::iC] + Fc[LY::iC + 1] + Fc[LY::iC + 2]) * pd) * fmatch);
                const SF hh = byr * r_oa;
                c0M = b1 + hh;
                c0X = b2 + hh;
                if constexpr (BAND) {   // column 0 outside the row's span (jlo0(i) > 0)
                    if (!(bgj[i] <= bw)) { c0M = SF(0); c0X = SF(0); }
                }
                mx = dmax(mx, dmax(c0M, c0X));
                sx = sfma(Fc[LY::iC + 1], c0X, sx);
            }
            // Lost mass (soft_kernel_impl.inl:141-178): any value flushed in this row's forward
            // or backward segment was below MIN in its own units; its effect on any posterior
            // is at most MIN times the other direction's segment maximum (3 mx bounds a cell's
            // M + X + Y) times the row's posterior factor, charged for every operation of the
            // cell.  A product of two stored values that leaves the normal range loses at most
            // the smallest subnormal (soft_kernel_impl.inl:933-948).
            // Banded: a lane whose cells all lie outside the row's span holds exact zeros backward
            // (no mass can be lost there), and its units are borrowed, so it charges nothing.
            bool span = true;
            if constexpr (BAND) span = (j0 + ncols >= blo && j0 + 1 <= bhi) || (j0 == 0 && bgj[i] <= bw);
            if (!zbad && span && (mx > SF(0) || mfc > SF(0))) {
                const double lr = 32.0 * (double)(ncols + 1) * A.vmin * (3.0 * (double)mx + (double)mfc) * fnew +
                                  8.0 * (double)(ncols + 1) * A.vtiny * (fnew + fold);
                const double lc = (lr <= 1.0e300) ? lr : 1.0e300;
                if constexpr (TR) atomicAdd(acc + 4 * KS * kWave, lc);
                else racc[4] += lc;
            }
            if (!zbad) {
                // every term is a sum of non-negative products: one range check covers them
                const double c1 = i >= 1 ? (double)sx * fnew : 0.0;   // no X move ends in row 0
                const double c2 = (double)sy * fnew;
                const double c3 = i >= kmin ? (double)so * (double)r_oa * fnew : 0.0;
                const double c4 = (double)sob * (double)r_ob * fold;
                if (!(c1 + c2 + c3 + c4 <= 1.0e300)) bad = true;
                else {
                    if constexpr (TR) {
                        atomicAdd(acc + (4 * gi) * kWave, c3);
                        atomicAdd(acc + (4 * gi + 1) * kWave, c2);
                        atomicAdd(acc + (4 * tn + 2) * kWave, c4);
                        atomicAdd(acc + (4 * ti + 3) * kWave, c1);
                    } else {
                        racc[0] += c3;
                        racc[1] += c2;
                        racc[2] += c4;
                        racc[3] += c1;
                    }
                }
            }
            if (!(mx <= R::fmax)) bad = true;
            rdM = rMc;
            EB = Ec;
            kpend = 0;
            if (mx > R::hi || (mx < R::lo && mx > SF(0))) {
                const int k2 = silogb(mx);
                kpend = k2 > R::kmax ? R::kmax : (k2 < -R::kmax ? -R::kmax : k2);
            }
            bzero = !(mx > SF(0));
            pubM = BM[0]; pubE = EB;
            if (pout) {
                SF* bo = bnd(pp, i);
                bo[3] = pubM; bo[4] = pubY;
                bex(pp, i)[1] = pubE;
            }
        };

        // One step at phase p of a band, S / W the band's buffers.  Phases before the last
        // recompute the row from S through the records (row 0 is copied: it is stored in full);
        // the last phase uses S and loads the next band's full row into W.
        auto phase = [&](int u, int p, SF(&Sb)[LY::vals], SF(&Wb)[LY::vals]) {
            const int i = m + G - 1 - g - u;
            const SF rM = wshift_down(pubM), rY = wshift_down(pubY);
            const int rE = wshift_down(pubE);
            if (!(live && i >= 0 && i <= m)) return;
            if (p < kCK - 1) {
                const int c = i - (kCK - 1 - p);
                const int c0 = c > 0 ? c : 0;
                const int nr = i - c0;   // rows to recompute, ending at row i
                if (nr >= 1) recompute(Sb, Wb, rec[0], rcode[0], rsets[0], c0 + 1);
#pragma unroll
                for (int j = 1; j < kCK - 1 - p; ++j)
                    if (nr > j) recompute(Wb, Wb, rec[j], rcode[j], rsets[j], c0 + 1 + j);
                if (nr == 0) copy_row(Sb, Wb);
                step(i, rM, rY, rE, Wb);
            } else {
                if (i >= 1) {
                    const int src = i - kCK > 0 ? i - kCK : 0;
                    load_seg<SF, S>(seg + (int64_t)(src + g) * LY::row, lane, Wb);
                    load_recs(src, i - 1);
                }
                step(i, rM, rY, rE, Sb);
            }
        };
        // pipelined: publish and wait every kPipeChunk steps, as the forward
        int bsync_at = (bin != nullptr || bout != nullptr) ? kPipeChunk : 0x7fffffff;
        for (int u = 0; u < bsteps; u += 2 * kCK) {
            if (u == bsync_at) {
                if (bout != nullptr) publish(bout, u, lane);
                if (bin != nullptr) bounded_wait(bin, min(u + kPipeChunk + kWave, bsteps), serr, A.wait_ticks, lane, 1);
                bsync_at += kPipeChunk;
            }
#pragma unroll
            for (int p = 0; p < kCK; ++p) phase(u + p, p, Fa, Fb);
#pragma unroll
            for (int p = 0; p < kCK; ++p) phase(u + kCK + p, p, Fb, Fa);
        }
        if (bout != nullptr) publish(bout, bsteps, lane);
        if (NP > 1) __threadfence();   // the panel's boundary column is read by the previous panel
    }

    // ── per-pair results ────────────────────────────────────────────────────
    // The pair's sums: untracked, the tree over its lanes; tracks, each of the K x 4 gap sums
    // and the loss bound added lane by lane from LDS (pair_sum, on the pair's first lane).
    wave_sync();
    double g_oa = 0.0, g_ea = 0.0, g_ob = 0.0, g_eb = 0.0, loss = 0.0;
    if constexpr (!TR) {
        g_oa = racc[0]; g_ea = racc[1]; g_ob = racc[2]; g_eb = racc[3]; loss = racc[4];
        for (int off = 1; off < kWave; off <<= 1) {
            const double o1 = __shfl_down(g_oa, off, kWave), o2 = __shfl_down(g_ea, off, kWave);
            const double o3 = __shfl_down(g_ob, off, kWave), o4 = __shfl_down(g_eb, off, kWave);
            const double o5 = __shfl_down(loss, off, kWave);
            const int o6 = __shfl_down((int)bad, off, kWave);
            if (g + off < G) {
                g_oa += o1; g_ea += o2; g_ob += o3; g_eb += o4; loss += o5;
                bad = bad || o6 != 0;
            }
        }
    } else {
        for (int off = 1; off < kWave; off <<= 1) {
            const int o6 = __shfl_down((int)bad, off, kWave);
            if (g + off < G) bad = bad || o6 != 0;
        }
    }
    auto pair_sum = [&](int c) -> double {   // accumulator c: 4 * set + field, or 4K the loss
        if constexpr (TR) {
            double v = 0.0;
            for (int l = lo; l < lo + G; ++l) v += lacc[c * kWave + l];
            return v;
        } else {
            return c == 0 ? g_oa : (c == 1 ? g_ea : (c == 2 ? g_ob : (c == 3 ? g_eb : loss)));
        }
    };
    const int nsum = 4 * KS + 1;   // the gap sums of every set, then the loss bound
    if constexpr (TR) {
        if (g == 0 && q >= 0) loss = pair_sum(4 * KS);
    }
    wave_sync();
    const int hsz = KS * na * hrow;   // one pair slot's histogram
    const int knn = KS * nn;
    const double logz = zbad ? 0.0 : ::log(zr) + (double)ze * 0.69314718055994530942;
    if (idir == 1) {
        // A backward item: this panel's histogram, sums and flag as the tile's part, then the
        // tile's count; the last backward item adds the parts of every panel in panel order.
        const int np = td.npairs, pst = knn + nsum + 1;
        const int64_t pstride = (int64_t)A.hist_slots * pst;
        double* __restrict__ P0 = A.lpart + (poff + ipp) * pstride;
        for (int idx = lane; idx < np * hsz; idx += kWave) {
            const int qq = idx / hsz, r = idx - qq * hsz;
            const int a = r / hrow, c = r - a * hrow;
            if (c < na) P0[qq * pst + a * na + c] = hist[idx];
            hist[idx] = 0.0;
        }
        if (g == 0 && q >= 0) {
            double* pq = P0 + q * pst + knn;
            for (int c = 0; c < nsum; ++c) pq[c] = pair_sum(c);
            pq[nsum] = bad ? 1.0 : 0.0;
        }
        __threadfence();
        const int prev = wave_uniform(
            __hip_atomic_fetch_add(sw + 2 * NP + 1, lane == 0 ? 1 : 0, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT));
        if (prev != NP - 1) {
            wave_sync();
            return;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const double* __restrict__ PB = A.lpart + poff * pstride;
        unsigned nonfinite = 0u;
        for (int idx = lane; idx < np * knn; idx += kWave) {
            const int qq = idx / knn, r = idx - qq * knn;
            double v = 0.0;
            for (int k = 0; k < NP; ++k) v += PB[k * pstride + qq * pst + r];
            A.grads[(int64_t)order[td.first + qq] * gs + r] = v;
            if (!(v <= 1.0e300)) nonfinite |= 1u << qq;
        }
        for (int off = 1; off < kWave; off <<= 1) nonfinite |= (unsigned)__shfl_xor((int)nonfinite, off, kWave);
        if (g == 0 && p >= 0) {
            bool bb = false;
            double lsum = 0.0;
            for (int k = 0; k < NP; ++k) {
                const double* pq = PB + k * pstride + q * pst + knn;
                lsum += pq[nsum - 1];
                const double* zp = A.lzpart + ((poff + k) * kTilePairs + q) * 4;
                bb = bb || pq[nsum] != 0.0 || zp[2] != 0.0;
            }
            const bool okp = !bb && !zbad && lsum <= A.tol && !((nonfinite >> q) & 1u);
            A.ok[p] = okp ? 1 : 0;
            A.scores[p] = logz * A.temperature;
            double* gr = A.grads + (int64_t)p * gs + knn;
            for (int c = 0; c < nsum - 1; ++c) {
                double v = 0.0;
                for (int k = 0; k < NP; ++k) v += PB[k * pstride + q * pst + knn + c];
                gr[c] = (A.linear && (c & 1) == 0) ? 0.0 : -v;
            }
        }
        publish(A.lsync + 1 + tix, 1, lane);
        return;
    }
    // Expected match counts out of the histogram (and the spill column cleared); a count that
    // left the double range marks its pair.
    const int np = td.npairs;
    unsigned nonfinite = 0u;
    for (int idx = lane; idx < np * hsz; idx += kWave) {
        const int qq = idx / hsz, r = idx - qq * hsz;
        const int a = r / hrow, c = r - a * hrow;
        const double v = hist[idx];
        if (c < na) {
            const int pq = order[td.first + qq];
            A.grads[(int64_t)pq * gs + a * na + c] = v;
            if (!(v <= 1.0e300)) nonfinite |= 1u << qq;
        }
        hist[idx] = 0.0;
    }
    for (int off = 1; off < kWave; off <<= 1) nonfinite |= (unsigned)__shfl_xor((int)nonfinite, off, kWave);
    if (g == 0 && p >= 0) {
        const bool okp = !bad && !zbad && loss <= A.tol && !((nonfinite >> q) & 1u);
        A.ok[p] = okp ? 1 : 0;
        A.scores[p] = logz * A.temperature;
        double* gr = A.grads + (int64_t)p * gs + knn;
        for (int c = 0; c < nsum - 1; ++c) gr[c] = (A.linear && (c & 1) == 0) ? 0.0 : -pair_sum(c);
    }
    if (idir == 2) publish(A.lsync + 1 + tix, 1, lane);   // its pool region is free
    wave_sync();
}

// The soft weights from the parameter vector on the device (K sets: K matrices, then K rows
// of go_a, ge_a, go_b, ge_b), as the reference forms them (aligner.hpp:6870-6885; tracked
// rows, aligner.hpp:6845-6866): exp(s / T) per matrix entry and exp(-(go+ge)/T), exp(-ge/T)
// per sequence (linear gaps: open = extend), each in double and cast once.  Also zeroes the
// persistent launch's queue counters and the pipelined wide tiles' words (a kernel, not a
// memset, so that a captured step holds no memset node).
template <typename SF>
__global__ __launch_bounds__(256) void soft_weights_kernel(const double* __restrict__ theta, int nn, int K, double it,
                                                           int linear, SF* __restrict__ psub, SF* __restrict__ gapw,
                                                           int32_t* q0, int32_t* q1, int32_t* sync, int64_t nsync) {
    for (int k = threadIdx.x; k < K * nn; k += blockDim.x) psub[k] = (SF)::exp(theta[k] * it);
    for (int64_t k = threadIdx.x; k < nsync; k += blockDim.x) sync[k] = 0;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const double* gp = theta + (size_t)K * nn + 4 * k;
        const double go_a = gp[0], ge_a = gp[1], go_b = gp[2], ge_b = gp[3];
        const double ea = ::exp(-ge_a * it), eb = ::exp(-ge_b * it);
        gapw[4 * k] = (SF)(linear ? ea : ::exp(-(go_a + ge_a) * it));
        gapw[4 * k + 1] = (SF)ea;
        gapw[4 * k + 2] = (SF)(linear ? eb : ::exp(-(go_b + ge_b) * it));
        gapw[4 * k + 3] = (SF)eb;
    }
    if (threadIdx.x == 0) {
        if (q0) *q0 = 0;
        if (q1) {
            q1[0] = 0;
            q1[1] = 0;
        }
    }
}

// Persistent: blocks below n_pwaves first take items of pipelined wide tiles (each tile in its
// own region of the wide pool), then blocks below n_lslots whole wide tiles, then blocks
// below n_mslots one-panel tiles, each of the last two in the block's own scratch slot.
// Plain (no queue): block b runs tile tile_base + b in slot b.  One call site of the tile
// body, so the kernel holds one copy of it.
// Occupancy: the float32 pass is held to two waves per SIMD (256 VGPRs); the double pass, twice
// the register bytes, to one.
// LDS of one block: the histograms (pair slots x K x n_alpha x (n_alpha + 1) doubles), the
// per-lane sums ((4K + 1) x 64 doubles), the K profiles ((n_alpha + 1)^2 each) and the K x 4
// gap weights.  K = 1 without tracks.
template <typename SF>
__host__ __device__ constexpr size_t soft_lds_bytes(int na, int K, int hist_slots) {
    return (size_t)hist_slots * K * na * (na + 1) * sizeof(double) + (size_t)(4 * K + 1) * kWave * sizeof(double) +
           (size_t)K * (na + 1) * (na + 1) * sizeof(SF) + (size_t)4 * K * sizeof(SF);
}

template <typename SF, int AM, int S, bool TR, bool BAND = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(SF) == 4 ? 2 : 1)))
void soft_kernel(SoftArgs<SF> A) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int na = A.na, nn = na * na, K = TR ? A.K : 1;
    double* hist = reinterpret_cast<double*>(lds);
    const int hsz = K * na * (na + 1);
    double* lacc = hist + (size_t)A.hist_slots * hsz;   // (4K + 1) x 64 per-lane sums
    SF* spsub = reinterpret_cast<SF*>(lacc + (4 * K + 1) * kWave);
    SF* sgw = spsub + K * (na + 1) * (na + 1);
    const int lane = threadIdx.x;
    // Each set's profile with a row and a column added: column na (a column past len(B))
    // weighs 0 and row na (the backward's virtual row m+1) weighs 1 everywhere.
    for (int k = lane; k < K * (na + 1) * (na + 1); k += kWave) {
        const int set = k / ((na + 1) * (na + 1)), r = k - set * (na + 1) * (na + 1);
        const int a = r / (na + 1), c = r - a * (na + 1);
        spsub[k] = a == na ? SF(1) : (c == na ? SF(0) : A.psub[set * nn + a * na + c]);
    }
    for (int k = lane; k < 4 * K; k += kWave) sgw[k] = A.gapw[k];
    for (int k = lane; k < A.hist_slots * hsz; k += kWave) hist[k] = 0.0;
    __syncthreads();
    const int blk = (int)blockIdx.x;
    int phase = A.queue == nullptr ? 3 : 0;
    for (;;) {
        int tile = 0, ipp = 0, idir = -1;
        bool wide = false;
        if (phase == 3) {
            tile = A.tile_base + blk;
            if (tile >= A.n_tiles) break;
            phase = 4;   // one tile only
        } else if (phase == 4) {
            break;
        } else {
            if (phase == 0) {   // items of pipelined wide tiles
                phase = 1;
                if (A.litems != nullptr && blk < A.n_pwaves) {
                    const int k = wave_fetch(A.lqueue);
                    if (k < A.n_litems) {
                        const int2 it = A.litems[k];
                        tile = wave_uniform(it.x);
                        const int code = wave_uniform(it.y);
                        ipp = code < 0 ? 0 : code >> 1;
                        idir = code < 0 ? 2 : code & 1;
                        wide = true;
                        phase = 0;
                    }
                }
            }
            if (phase == 1) {   // whole wide tiles
                phase = 2;
                if (blk < A.n_lslots) {
                    const int k = wave_fetch(A.lqueue + 1);
                    if (A.n_ptiles + k < A.n_ltiles) {
                        tile = A.n_ptiles + k;
                        wide = true;
                        phase = 1;
                    }
                }
            }
            if (phase == 2) {   // one-panel tiles
                if (blk >= A.n_mslots) break;
                tile = wave_fetch(A.queue);
                if (tile >= A.n_tiles) break;
            }
        }
        soft_tile<SF, AM, S, TR, BAND>(A, wide ? A.lview : A.view, wide ? A.lorder : A.order,
                                 (wide ? A.ltiles : A.tiles) + tile, (int64_t)blk, spsub, sgw, hist, lacc, lane, tile,
                                 ipp, idir);
    }
}

}  // namespace nwg
