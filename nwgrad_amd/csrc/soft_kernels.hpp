// Soft (forward-backward) pass on gfx950: expected counts of the alignment posterior.
//
// Reference behaviour restated (relative to the reference's src/nwgrad/cpp/nwgrad/):
//   scaled forward          soft_kernel_impl.inl:643-708 (sp_affine_fwd), 739-760 (row 0)
//   backward + fused sums   soft_kernel_impl.inl:790-977
//   Z and the end rules     soft_kernel_impl.inl:294-321 (sp_log_z)
//   gap-target rows         soft_kernel_impl.inl:736-737 (kmin, kmin0)
//   score = T log Z, gap fields negated  aligner.hpp:6554-6626, soft_kernel_impl.inl:987
// Linear gaps run as the affine model with open = 0: with oa = ea and ob = eb the affine
// total M + X + Y obeys H(i,j) = H(i-1,j-1) P + H(i-1,j) eb + H(i,j-1) ea, which is the linear
// recurrence (aligner.hpp:6231-6260), and the expected X / Y moves are its gap counts.
//
// Mapping: the lane-group layout of the hard fill (hard_kernels.hpp).  The reference scales
// each DP row by a power of two chosen from the row maximum; lanes of a slot run skewed and
// never hold a whole row at once, so here every (lane, row) segment carries its own
// power-of-two exponent E (true value = stored value * 2^E).  A lane computes its row in its
// own exponent, converts the boundary values it receives from its neighbour (an exact power-
// of-two product), and renormalises the segment when its maximum leaves [2^-w, 2^w].  Values
// more than the format's range below a segment's maximum flush, as in the reference's row
// scaling; a segment is never coarser than the reference's row.
//
// The forward stores every row segment of the lane (M, X, Y of its S columns plus the
// boundary column it received, converted to its exponent) and the exponent to HBM scratch;
// the backward runs the mirrored skew (lane G-1 first), carries F(i) and F(i-1) in registers,
// and accumulates the gradient sums per row in the working precision, scaled into double
// by the row's posterior factor 2^(E_F + E_B - ze) / zr, with match posteriors added into a
// per-pair double LDS histogram (ds_add_f64).
#pragma once

#include "common.hpp"

namespace nwg {

template <typename SF>
struct SoftArgs {
    const uint8_t* codes_a;
    const uint8_t* codes_b;
    const int64_t* off_a;
    const int64_t* off_b;
    const int32_t* order;
    const TileDesc* tiles;
    int32_t n_tiles;
    int32_t tile_base;
    int32_t* queue;
    const SF* psub;          // exp(matrix / T), n_alpha * n_alpha
    int32_t na;
    int32_t hist_slots;
    SF oa, ea, ob, eb;       // exp(-(go+ge)/T), exp(-ge/T) per sequence
    int32_t linear;          // 1: report the open fields as 0
    double temperature;
    double* scores;
    double* grads;
    uint8_t* ok;             // per pair: 1 in range, 0 the pass lost range (overflow / non-finite)
    SF* ftab;                // forward row segments, per wave slot
    int32_t* etab;           // their exponents
    SF* mtab;                // their maxima (stored units, after renormalisation)
    SF* ltab;                // 2^max(0, -k) of their renormalisation by 2^-k: a value lost
                             // below MIN before a scale-up is below MIN * ltab stored
    double tol;              // lost-mass tolerance (2^-20 float32, 2^-45 double)
    double vmin;             // the working format's smallest normal value
    SF* pbnd;                // panel boundary columns: per row M, X, Y (forward), BM, BY (backward)
    int32_t* pexp;           // their exponents: forward, backward
    int32_t panels;          // panel regions per slot
    int64_t slot_rows;       // rows of one panel region (max over tiles of max_m + G)
};

template <typename SF>
struct SoftRange;
template <>
struct SoftRange<float> {
    static constexpr int w = 16;         // renormalise outside [2^-16, 2^16]
    static constexpr int headroom = 64;  // boundary values at most 2^64 above a segment's scale
    static constexpr float hi = 65536.0f, lo = 1.0f / 65536.0f;
    static constexpr float fmax = 3.4028234663852886e38f;
};
template <>
struct SoftRange<double> {
    static constexpr int w = 256;
    static constexpr int headroom = 512;
    static constexpr double hi = 1.157920892373162e77, lo = 8.636168555094445e-78;
    static constexpr double fmax = 1.7976931348623157e308;
};

__device__ __forceinline__ float sldexp(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
__device__ __forceinline__ double sldexp(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }
__device__ __forceinline__ int silogb(float x) { return __builtin_amdgcn_frexp_expf(x) - 1; }
__device__ __forceinline__ int silogb(double x) { return __builtin_amdgcn_frexp_exp(x) - 1; }
__device__ __forceinline__ float sfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double sfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

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

// Row-segment storage: 3 states x (S + 1) columns (column S: the boundary), padded to
// a multiple of 4 and laid out so each 16-byte vector of the 64 lanes is contiguous.
template <int S>
struct SegLayout {
    static constexpr int cols = S + 1;
    static constexpr int vals = 3 * cols;
    static constexpr int padded = (vals + 3) & ~3;
};

template <typename SF, int S>
__device__ __forceinline__ void store_seg(SF* __restrict__ base, int lane, const SF (&M)[S], const SF (&X)[S],
                                          const SF (&Y)[S], SF cM, SF cX, SF cY) {
    constexpr int P = SegLayout<S>::padded;
    SF v[P];
#pragma unroll
    for (int k = 0; k < S; ++k) { v[k] = M[k]; v[S + 1 + k] = X[k]; v[2 * (S + 1) + k] = Y[k]; }
    v[S] = cM; v[2 * S + 1] = cX; v[3 * S + 2] = cY;
#pragma unroll
    for (int k = SegLayout<S>::vals; k < P; ++k) v[k] = SF(0);
    if constexpr (sizeof(SF) == 4) {
#pragma unroll
        for (int q = 0; q < P / 4; ++q) {
            float4 f = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
            reinterpret_cast<float4*>(base)[q * kWave + lane] = f;
        }
    } else {
#pragma unroll
        for (int q = 0; q < P / 2; ++q) {
            double2 f = make_double2(v[2 * q], v[2 * q + 1]);
            reinterpret_cast<double2*>(base)[q * kWave + lane] = f;
        }
    }
}

template <typename SF, int S>
__device__ __forceinline__ void load_seg(const SF* __restrict__ base, int lane, SF (&v)[SegLayout<S>::padded]) {
    constexpr int P = SegLayout<S>::padded;
    if constexpr (sizeof(SF) == 4) {
#pragma unroll
        for (int q = 0; q < P / 4; ++q) {
            const float4 f = reinterpret_cast<const float4*>(base)[q * kWave + lane];
            v[4 * q] = f.x; v[4 * q + 1] = f.y; v[4 * q + 2] = f.z; v[4 * q + 3] = f.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < P / 2; ++q) {
            const double2 f = reinterpret_cast<const double2*>(base)[q * kWave + lane];
            v[2 * q] = f.x; v[2 * q + 1] = f.y;
        }
    }
}

template <typename SF, int AM, int S>
__device__ void soft_tile(const SoftArgs<SF>& A, const SF* __restrict__ spsub, double* __restrict__ hist, int tile,
                          int64_t slot, int lane) {
    using R = SoftRange<SF>;
    constexpr int P = SegLayout<S>::padded;
    constexpr bool L = AM == kLocal, GL = AM == kGlocal;
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
    // Columns per panel: a pair wider than one wavefront's lanes runs panel by panel (one
    // pair per tile then), passing each row's boundary column through HBM.
    const int PW = G * S;
    const int NP = (G == kWave && n > PW) ? (n + PW - 1) / PW : 1;
    const SF oa = A.oa, ea = A.ea, ob = A.ob, eb = A.eb;
    const int64_t rows = A.slot_rows;
    const int64_t pbase = slot * (int64_t)A.panels;   // this slot's first panel region
    bool bad = false;

    // ── forward ──────────────────────────────────────────────────────────────
    SF M[S], X[S], Y[S];
    SF cM = SF(0), cX = SF(0), cY = SF(0);   // the segment's boundary column j0 (column 0 at j0 = 0)
    int E = 0;
    ExpAcc zacc{0.0, 0};
    for (int pp = 0; pp < NP; ++pp) {
        const int j0 = pp * PW + g * S;
        const int ncols = n - j0 < 0 ? 0 : (n - j0 > S ? S : n - j0);
        const bool live = p >= 0 && (ncols > 0 || j0 == 0);   // the lane of column 0 always computes
        int coff[S];
#pragma unroll
        for (int k = 0; k < S; ++k) coff[k] = (k < ncols) ? (int)bp[j0 + k] : 0;
        SF* __restrict__ ft = A.ftab + (pbase + pp) * rows * P * kWave;
        int* __restrict__ et = A.etab + (pbase + pp) * rows * kWave;
        SF* __restrict__ mt = A.mtab + (pbase + pp) * rows * kWave;
        SF* __restrict__ lt = A.ltab + (pbase + pp) * rows * kWave;
        SF* __restrict__ bout = A.pbnd + (pbase + pp) * rows * 6;         // this panel's last column
        const SF* __restrict__ bin = A.pbnd + (pbase + pp - 1) * rows * 6;  // the previous panel's
        int* __restrict__ eout = A.pexp + (pbase + pp) * rows * 2;
        const int* __restrict__ ein = A.pexp + (pbase + pp - 1) * rows * 2;
        E = 0;
        if (L || GL) {
#pragma unroll
            for (int k = 0; k < S; ++k) { M[k] = (k < ncols) ? SF(1) : SF(0); X[k] = SF(0); Y[k] = SF(0); }
            cM = SF(1); cX = SF(0); cY = SF(0);
        } else {
            // Row 0 of Global: M(0,0) = 1, Y(0,j) = oa * ea^(j-1) (soft_kernel_impl.inl:740-744),
            // evaluated per segment through log2 so that far columns keep their own exponent.
            const double l2oa = ::log2((double)oa), l2ea = ::log2((double)ea);
            if (j0 == 0) {
                cM = SF(1); cX = SF(0); cY = SF(0);
            } else {
                double top = l2oa + (double)(j0 - 1) * l2ea;   // column j0 (the boundary)
                const double last = l2oa + (double)(j0 + (ncols > 0 ? ncols : 1) - 1) * l2ea;
                if (last > top) top = last;
                E = (int)::floor(top);
                cM = SF(0); cX = SF(0);
                cY = (SF)::exp2(l2oa + (double)(j0 - 1) * l2ea - (double)E);
            }
#pragma unroll
            for (int k = 0; k < S; ++k) {
                M[k] = SF(0); X[k] = SF(0);
                Y[k] = (k < ncols) ? (SF)::exp2(l2oa + (double)(j0 + k) * l2ea - (double)E) : SF(0);
            }
        }
        if (live) {
            if (L || (GL && m == 0)) {
                SF s = j0 == 0 ? cM + cX + cY : SF(0);
#pragma unroll
                for (int k = 0; k < S; ++k) s += (k < ncols) ? M[k] + X[k] + Y[k] : SF(0);
                expacc_add(zacc, (double)s, E);
            }
            store_seg<SF, S>(ft + (int64_t)g * P * kWave, lane, M, X, Y, cM, cX, cY);
            et[(int64_t)g * kWave + lane] = E;
            SF mx0 = j0 == 0 ? cM + cX + cY : SF(0);
#pragma unroll
            for (int k = 0; k < S; ++k) if (k < ncols) mx0 = dmax(mx0, M[k] + X[k] + Y[k]);
            mt[(int64_t)g * kWave + lane] = mx0;
            lt[(int64_t)g * kWave + lane] = SF(1);
        }
        SF oM = M[S - 1], oX = X[S - 1], oY = Y[S - 1];
        int oE = E;
        bool fzero;   // the lane's last row held no mass
        {
            SF z = SF(0);
#pragma unroll
            for (int k = 0; k < S; ++k) z = dmax(z, M[k] + X[k] + Y[k]);
            fzero = !L && !(z > SF(0)) && !(cM + cX + cY > SF(0));
        }
        const int fsteps = td.max_m + G - 1;
        for (int t = 0; t < fsteps; ++t) {
            const int i = t - g + 1;
            SF lM = shift_up(oM), lX = shift_up(oX), lY = shift_up(oY);
            int lE = shift_up(oE);
            if (live && i >= 1 && i <= m) {
                if (g == 0 && j0 > 0) {   // the previous panel's last column of row i
                    lM = bin[(int64_t)i * 6 + 0]; lX = bin[(int64_t)i * 6 + 1]; lY = bin[(int64_t)i * 6 + 2];
                    lE = ein[(int64_t)i * 2];
                }
                int Ec = E;
                SF nM, nX, nY;
                if (j0 == 0) {   // column 0 of row i (soft_kernel_impl.inl:669-674)
                    if constexpr (L) { nM = sldexp(SF(1), -Ec); nX = SF(0); nY = SF(0); }
                    else { nM = SF(0); nY = SF(0); nX = sfma(cX, eb, (cM + cY) * ob); }
                } else {
                    if (fzero) {
                        Ec = lE;   // the segment holds no mass yet: take the neighbour's exponent
                    } else if (lE - R::headroom > Ec) {
                        const int nE = lE - R::headroom;
#pragma unroll
                        for (int k = 0; k < S; ++k) {
                            M[k] = sldexp(M[k], Ec - nE); X[k] = sldexp(X[k], Ec - nE); Y[k] = sldexp(Y[k], Ec - nE);
                        }
                        cM = sldexp(cM, Ec - nE); cX = sldexp(cX, Ec - nE); cY = sldexp(cY, Ec - nE);
                        Ec = nE;
                    }
                    nM = sldexp(lM, lE - Ec); nX = sldexp(lX, lE - Ec); nY = sldexp(lY, lE - Ec);
                }
                const SF fr = L ? sldexp(SF(1), -Ec) : SF(0);
                const SF* __restrict__ prow = spsub + (int)ap[i - 1] * na;
                SF dT = cM + cX + cY;
                SF lm = nM, lx = nX, ly = nY;
                SF mx = (j0 == 0) ? nM + nX + nY : SF(0);
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    const SF pw = prow[coff[k]];
                    const SF uM = M[k], uX = X[k], uY = Y[k];
                    const SF mv = sfma(dT, pw, fr);
                    const SF xv = sfma(uX, eb, (uM + uY) * ob);
                    const SF yv = sfma(ly, ea, (lm + lx) * oa);
                    dT = uM + uX + uY;
                    M[k] = mv; X[k] = xv; Y[k] = yv;
                    lm = mv; lx = xv; ly = yv;
                    if (k < ncols) mx = dmax(mx, mv + xv + yv);
                }
                cM = nM; cX = nX; cY = nY;
                if (!(mx <= R::fmax)) bad = true;
                SF lfac = SF(1);
                if (mx > R::hi || (mx < R::lo && mx > SF(0))) {
                    const int k2 = silogb(mx);
#pragma unroll
                    for (int k = 0; k < S; ++k) {
                        M[k] = sldexp(M[k], -k2); X[k] = sldexp(X[k], -k2); Y[k] = sldexp(Y[k], -k2);
                    }
                    cM = sldexp(cM, -k2); cX = sldexp(cX, -k2); cY = sldexp(cY, -k2);
                    mx = sldexp(mx, -k2);
                    if (k2 < 0) lfac = sldexp(SF(1), -k2);
                    Ec += k2;
                }
                lt[(int64_t)(i + g) * kWave + lane] = lfac;
                E = Ec;
                mt[(int64_t)(i + g) * kWave + lane] = mx;
                fzero = !L && !(mx > SF(0)) && !(cM + cX + cY > SF(0));
                if (L || (GL && i == m)) {
                    SF s = j0 == 0 ? cM + cX + cY : SF(0);
#pragma unroll
                    for (int k = 0; k < S; ++k) s += (k < ncols) ? M[k] + X[k] + Y[k] : SF(0);
                    expacc_add(zacc, (double)s, E);
                }
                store_seg<SF, S>(ft + (int64_t)(i + g) * P * kWave, lane, M, X, Y, cM, cX, cY);
                et[(int64_t)(i + g) * kWave + lane] = E;
                oM = M[S - 1]; oX = X[S - 1]; oY = Y[S - 1]; oE = E;
                if (g == G - 1 && pp + 1 < NP) {
                    bout[(int64_t)i * 6 + 0] = oM; bout[(int64_t)i * 6 + 1] = oX; bout[(int64_t)i * 6 + 2] = oY;
                    eout[(int64_t)i * 2] = oE;
                }
            }
        }
        if (NP > 1) __threadfence();   // the panel's boundary column is read by the next panel
    }

    // ── Z ───────────────────────────────────────────────────────────────────
    if constexpr (!L && !GL) {
        // Global: the (m, n) cell, held by the lane of column n in the last panel (lane 0
        // when n == 0).
        const int jl = (NP - 1) * PW;
        const int gn = n == 0 ? 0 : (n - 1 - jl) / S;
        SF corner = SF(0);
        if (p >= 0 && g == gn) {
            if (n == 0) corner = cM + cX + cY;
            else {
#pragma unroll
                for (int k = 0; k < S; ++k)
                    if (k == n - 1 - jl - g * S) corner = M[k] + X[k] + Y[k];
            }
        }
        zacc.acc = (double)corner;
        zacc.E = E;
        const int src = (lane & ~(G - 1)) + gn;
        zacc.acc = __shfl(zacc.acc, src, kWave);
        zacc.E = __shfl(zacc.E, src, kWave);
    } else {
        for (int off = 1; off < G; off <<= 1) {
            const double oacc = __shfl_xor(zacc.acc, off, G);
            const int oe = __shfl_xor(zacc.E, off, G);
            expacc_add(zacc, oacc, oe);
        }
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
    double g_oa = 0.0, g_ea = 0.0, g_ob = 0.0, g_eb = 0.0, loss = 0.0;
    const int kmin = (L || GL) ? 1 : 0;   // first row whose Y opens count
    const bool col0_opens = !L;           // column-0 X opens count (kmin0 == 0)
    for (int pp = NP - 1; pp >= 0; --pp) {
        const int j0 = pp * PW + g * S;
        const int ncols = n - j0 < 0 ? 0 : (n - j0 > S ? S : n - j0);
        const bool live = p >= 0 && (ncols > 0 || j0 == 0);
        int coff[S];
#pragma unroll
        for (int k = 0; k < S; ++k) coff[k] = (k < ncols) ? (int)bp[j0 + k] : 0;
        const int coff_next = (j0 + S < n) ? (int)bp[j0 + S] : 0;
        const SF* __restrict__ ft = A.ftab + (pbase + pp) * rows * P * kWave;
        const int* __restrict__ et = A.etab + (pbase + pp) * rows * kWave;
        const SF* __restrict__ mt = A.mtab + (pbase + pp) * rows * kWave;
        const SF* __restrict__ lt = A.ltab + (pbase + pp) * rows * kWave;
        SF* __restrict__ bout = A.pbnd + (pbase + pp) * rows * 6;          // this panel's first column
        const SF* __restrict__ bin = A.pbnd + (pbase + pp + 1) * rows * 6;  // the next panel's
        int* __restrict__ eout = A.pexp + (pbase + pp) * rows * 2;
        const int* __restrict__ ein = A.pexp + (pbase + pp + 1) * rows * 2;
        SF BM[S], BX[S], BY[S];
#pragma unroll
        for (int k = 0; k < S; ++k) { BM[k] = SF(0); BX[k] = SF(0); BY[k] = SF(0); }
        SF c0M = SF(0), c0X = SF(0);   // column 0 of the row below (the lane of column 0)
        SF rdM = SF(0);                // BM(i+1, j0+S+1) in the segment's exponent
        int EB = 0;
        bool bzero = true;             // the lane's last backward row held no mass
        const bool owns_end = (n == 0) ? (j0 == 0) : (j0 < n && n <= j0 + S);
        SF fc[P], fp[P];
        int efc = 0, efp = 0;
        SF mfc = SF(0), mfp = SF(0), lfc = SF(1), lfp = SF(1);
        bool loaded = false;
        SF pubM = SF(0), pubY = SF(0);
        int pubE = 0;
        const int bsteps = td.max_m + G;
        for (int u = 0; u < bsteps; ++u) {
            const int i = m + G - 1 - g - u;
            SF rM = shift_down(pubM), rY = shift_down(pubY);
            int rE = shift_down(pubE);
            if (!(live && i >= 0 && i <= m)) continue;
            if (g == G - 1 && pp + 1 < NP) {   // the next panel's first column of row i
                rM = bin[(int64_t)i * 6 + 3]; rY = bin[(int64_t)i * 6 + 4];
                rE = ein[(int64_t)i * 2 + 1];
            }
            if (!loaded) {
                load_seg<SF, S>(ft + (int64_t)(i + g) * P * kWave, lane, fc);
                efc = et[(int64_t)(i + g) * kWave + lane];
                mfc = mt[(int64_t)(i + g) * kWave + lane];
                lfc = lt[(int64_t)(i + g) * kWave + lane];
                loaded = true;
            } else {
#pragma unroll
                for (int k = 0; k < P; ++k) fc[k] = fp[k];
                efc = efp;
                mfc = mfp;
                lfc = lfp;
            }
            if (i >= 1) {
                load_seg<SF, S>(ft + (int64_t)(i - 1 + g) * P * kWave, lane, fp);
                efp = et[(int64_t)(i - 1 + g) * kWave + lane];
                mfp = mt[(int64_t)(i - 1 + g) * kWave + lane];
                lfp = lt[(int64_t)(i - 1 + g) * kWave + lane];
            }
            // The right boundary: BY(i, j0+S+1) and BM(i, j0+S+1).  A segment without mass
            // of its own (Global's row m left of column n, or rows whose mass has not
            // arrived) takes the neighbour's exponent; otherwise the neighbour's values are
            // converted, the segment's own moved down when they would not fit.
            const bool has_right = j0 + S < n;
            const bool own_zero = (i == m) ? (!L && !GL && !owns_end) : (!L && bzero);
            int Ec = (i == m) ? 0 : EB;
            if (has_right && own_zero) {
                Ec = rE;
            } else if (has_right && rE - R::headroom > Ec) {
                const int nE = rE - R::headroom;
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    BM[k] = sldexp(BM[k], Ec - nE); BX[k] = sldexp(BX[k], Ec - nE); BY[k] = sldexp(BY[k], Ec - nE);
                }
                c0M = sldexp(c0M, Ec - nE); c0X = sldexp(c0X, Ec - nE); rdM = sldexp(rdM, Ec - nE);
                Ec = nE;
            }
            const SF rYc = has_right ? sldexp(rY, rE - Ec) : SF(0);
            const SF rMc = has_right ? sldexp(rM, rE - Ec) : SF(0);
            const SF init = L ? sldexp(SF(1), -Ec) : SF(0);
            const SF* __restrict__ pn = (i < m) ? spsub + (int)ap[i] * na : spsub;      // row i+1's profile
            const SF* __restrict__ pr = (i >= 1) ? spsub + (int)ap[i - 1] * na : spsub; // row i's profile
            SF byr = rYc;          // BY(i, j+1), right to left
            SF dn = rdM;           // BM(i+1, j+1)
            SF mx = SF(0);
            double sx = 0, sy = 0, so = 0, sob = 0;
            const double gh0 = zbad ? 0.0 : __builtin_amdgcn_ldexp(izr, efc + Ec - ze);
            const double gv0 = (i >= 1 && !zbad) ? __builtin_amdgcn_ldexp(izr, efp + Ec - ze) : 0.0;
            double* __restrict__ h = hist + q * nn + (i >= 1 ? (int)ap[i - 1] * na : 0);
#pragma unroll
            for (int kk = 0; kk < S; ++kk) {
                const int k = S - 1 - kk;
                const int j = j0 + k + 1;
                SF b1, b2, y;
                if (i < m) {
                    const SF pw = pn[k == S - 1 ? coff_next : coff[k + 1]];
                    const SF d = sfma(pw, dn, init);
                    const SF v = BX[k];
                    b1 = sfma(v, ob, d);
                    b2 = sfma(v, eb, d);
                    y = b1;
                } else {   // row m end weights (soft_kernel_impl.inl:815-818)
                    const SF e = L ? init : (GL ? SF(1) : (j == n ? SF(1) : SF(0)));
                    b1 = e; b2 = e; y = e;
                }
                const SF byn = sfma(ea, byr, y);
                const SF hh = byr * oa;
                const SF bmn = (k < ncols) ? b1 + hh : SF(0);
                const SF bxn = (k < ncols) ? b2 + hh : SF(0);
                const SF byv = (k < ncols) ? byn : SF(0);
                dn = BM[k];
                BM[k] = bmn; BX[k] = bxn; BY[k] = byv;
                byr = byv;
                if (k < ncols) mx = dmax(mx, bmn + bxn + byv);
                // fused gradient sums (soft_kernel_impl.inl:871-903); products in double so
                // that a product of two small stored values never flushes
                const double fX = (double)fc[S + 1 + k], fY = (double)fc[2 * (S + 1) + k];
                const double fMl = (k == 0) ? (double)fc[S] : (double)fc[k - 1];
                const double fXl = (k == 0) ? (double)fc[2 * S + 1] : (double)fc[S + 1 + k - 1];
                sy = __builtin_fma(fY, (double)byv, sy);
                so = __builtin_fma(fMl + fXl, (double)byv, so);
                if (i >= 1) {
                    const double qM = (double)fp[k], qY = (double)fp[2 * (S + 1) + k];
                    const double qMl = (k == 0) ? (double)fp[S] : (double)fp[k - 1];
                    const double qXl = (k == 0) ? (double)fp[2 * S + 1] : (double)fp[S + 1 + k - 1];
                    const double qYl = (k == 0) ? (double)fp[3 * S + 2] : (double)fp[2 * (S + 1) + k - 1];
                    sx = __builtin_fma(fX, (double)bxn, sx);
                    sob = __builtin_fma(qM + qY, (double)bxn, sob);
                    if (k < ncols) {
                        const double tmatch = (qMl + qXl + qYl) * (double)pr[coff[k]] * (double)bmn;
                        if (tmatch > 0.0) {
                            const double c = tmatch * gv0;
                            if (!(c <= 1.0e300)) bad = true;
                            else atomicAdd(h + coff[k], c);
                        }
                    }
                }
            }
            if (j0 == 0) {   // column 0 (soft_kernel_impl.inl:849-857)
                SF b1, b2;
                if (i < m) {
                    const SF pw = pn[coff[0]];
                    const SF d = sfma(pw, dn, init);
                    b1 = sfma(c0X, ob, d);
                    b2 = sfma(c0X, eb, d);
                } else {
                    const SF e = L ? init : (GL ? SF(1) : (n == 0 ? SF(1) : SF(0)));
                    b1 = e; b2 = e;
                }
                const SF hh = byr * oa;
                c0M = b1 + hh;
                c0X = b2 + hh;
                mx = dmax(mx, c0M + c0X);
                if (i >= 1) {
                    sx = __builtin_fma((double)fc[2 * S + 1], (double)c0X, sx);
                    if (col0_opens) sob = __builtin_fma((double)fp[S] + (double)fp[3 * S + 2], (double)c0X, sob);
                }
            }
            // Lost mass (soft_kernel_impl.inl:141-178): any value flushed in this row's
            // forward or backward segment was below the format's MIN in its stored units
            // (times the forward's scale-up factor); its effect on any posterior is at most
            // that times the other direction's segment maximum times the row's posterior
            // factor, charged for every operation of the cell.
            if (!zbad && (mx > SF(0) || mfc > SF(0))) {
                const double lr =
                    32.0 * (double)(ncols + 1) * A.vmin * ((double)lfc * (double)mx + (double)mfc) * gh0;
                loss += (lr <= 1.0e300) ? lr : 1.0e300;
            }
            if (!zbad) {
                auto acc = [&](double& dst, double v, double f) {
                    if (v != 0.0) {
                        const double c = v * f;
                        if (!(c <= 1.0e300 && c >= -1.0e300)) bad = true;
                        else dst += c;
                    }
                };
                acc(g_eb, sx, gh0);
                acc(g_ea, sy, gh0);
                if (i >= kmin) acc(g_oa, so * (double)oa, gh0);
                if (i >= 1) acc(g_ob, sob * (double)ob, gv0);
            }
            if (!(mx <= R::fmax)) bad = true;
            rdM = rMc;
            if (mx > R::hi || (mx < R::lo && mx > SF(0))) {
                const int k2 = silogb(mx);
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    BM[k] = sldexp(BM[k], -k2); BX[k] = sldexp(BX[k], -k2); BY[k] = sldexp(BY[k], -k2);
                }
                c0M = sldexp(c0M, -k2); c0X = sldexp(c0X, -k2); rdM = sldexp(rdM, -k2);
                Ec += k2;
            }
            EB = Ec;
            bzero = !(mx > SF(0));
            pubM = BM[0]; pubY = BY[0]; pubE = EB;
            if (g == 0 && pp > 0) {
                bout[(int64_t)i * 6 + 3] = pubM; bout[(int64_t)i * 6 + 4] = pubY;
                eout[(int64_t)i * 2 + 1] = pubE;
            }
        }
        if (NP > 1) __threadfence();   // the panel's boundary column is read by the previous panel
    }

    // ── per-pair results ────────────────────────────────────────────────────
    for (int off = 1; off < G; off <<= 1) {
        g_oa += __shfl_xor(g_oa, off, G);
        g_ea += __shfl_xor(g_ea, off, G);
        g_ob += __shfl_xor(g_ob, off, G);
        g_eb += __shfl_xor(g_eb, off, G);
        loss += __shfl_xor(loss, off, G);
        bad = bad || __shfl_xor((int)bad, off, G);
    }
    wave_sync();
    const int np = td.npairs;
    for (int idx = lane; idx < np * nn; idx += kWave) {
        const int qq = idx / nn, e = idx - qq * nn;
        const int pq = A.order[td.first + qq];
        A.grads[(int64_t)pq * gs + e] = hist[qq * nn + e];
        hist[qq * nn + e] = 0.0;
    }
    if (g == 0 && p >= 0) {
        const bool okp = !bad && !zbad && loss <= A.tol;
        A.ok[p] = okp ? 1 : 0;
        const double logz = zbad ? 0.0 : ::log(zr) + (double)ze * 0.69314718055994530942;
        A.scores[p] = logz * A.temperature;
        double* gr = A.grads + (int64_t)p * gs + nn;
        gr[0] = A.linear ? 0.0 : -g_oa;
        gr[1] = -g_ea;
        gr[2] = A.linear ? 0.0 : -g_ob;
        gr[3] = -g_eb;
    }
    wave_sync();
}

template <typename SF, int AM, int S>
__global__ __launch_bounds__(64) void soft_kernel(SoftArgs<SF> A) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int na = A.na, nn = na * na;
    double* hist = reinterpret_cast<double*>(lds);
    SF* spsub = reinterpret_cast<SF*>(lds + (size_t)A.hist_slots * nn * sizeof(double));
    const int lane = threadIdx.x;
    for (int k = lane; k < nn; k += kWave) spsub[k] = A.psub[k];
    for (int k = lane; k < A.hist_slots * nn; k += kWave) hist[k] = 0.0;
    __syncthreads();
    if (A.queue == nullptr) {
        const int tile = A.tile_base + (int)blockIdx.x;
        if (tile < A.n_tiles) soft_tile<SF, AM, S>(A, spsub, hist, tile, (int64_t)blockIdx.x, lane);
        return;
    }
    for (;;) {
        int tile = 0;
        if (lane == 0) tile = atomicAdd(A.queue, 1);
        tile = __shfl(tile, 0, kWave);
        if (tile >= A.n_tiles) break;
        soft_tile<SF, AM, S>(A, spsub, hist, tile, (int64_t)blockIdx.x, lane);
    }
}

}  // namespace nwg
