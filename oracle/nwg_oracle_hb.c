/* TEST INFRASTRUCTURE ONLY -- the CPU oracle of this repository (Hirschberg traceback).
 *
 * A plain-C restatement of the reference's linear-memory traceback for Global affine pairs:
 * traceback="hirschberg" (the exact `- ge_a` carry) and "hirschberg_pmax" (the closed-form
 * prefix-max carry).  Used by tests/ as the checker of the GPU recursion; the product never
 * links, loads or calls it.  Citations relative to the reference's src/nwgrad/cpp/nwgrad/:
 *   - which pairs split: more than hb_cutoff rows          aligner.hpp:136-141, 3295-3330
 *   - half sweeps, forward and mirrored (both carries)     aligner.hpp:4150-4255
 *   - base case: the Pointers fill of a block, in_x/out_x  aligner.hpp:4313-4400
 *   - the recursion and its join (span refund)             aligner.hpp:4456-4518
 *   - the score replayed along the path, in T              aligner.hpp:4525-4560
 *   - gradient counts and guide of a held path             aligner.hpp:5191-5207, 5233-5250
 * Parity pins (tests/test_oracle_hb.py): fixtures the reference engine wrote with
 * traceback 3 / 4 at several cutoffs, and what the mathematics fixes -- the path is a legal
 * global alignment, its replayed score equals the Pointers fill's optimum where the sums are
 * exact, the counts reproduce the score, and with a unique optimum the path is Pointers'.
 *
 * Arithmetic as in nwg_oracle.c: every penalty converted once, T(double); the reference's
 * operations in its order; -ffp-contract=off.  max2(a, b) = a < b ? b : a (std::max), max3
 * the first of equal maxima (std::max of an initializer list).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int gap_model; /* 1 affine (the only model this file restates) */
    int mode;      /* 0 global */
    int n_alpha;
    const double *matrix;
    double go_a, ge_a, go_b, ge_b;
} nwo_hb_params;

/* the fill guide's gaps (aligner.hpp:2734-2749): rows without an entry interpolate between
   their neighbours, rows past the last entry repeat it */
static void hb_fill_guide_gaps(int32_t *gj, int n) {
    int lo = 0;
    while (lo < n) {
        if (gj[lo] < 0) { ++lo; continue; }
        int hi = lo + 1;
        while (hi < n && gj[hi] < 0) ++hi;
        if (hi == n) {
            for (int k = lo + 1; k < n; ++k) gj[k] = gj[lo];
            break;
        }
        for (int k = lo + 1; k < hi; ++k) gj[k] = gj[lo] + (gj[hi] - gj[lo]) * (k - lo) / (hi - lo);
        lo = hi;
    }
}

#define NWO_DEFINE_HB(T, SFX, NINF)                                                                 \
    static inline T hmax2_##SFX(T a, T b) { return a < b ? b : a; }                                  \
    static inline T hmax3_##SFX(T a, T b, T c) { return hmax2_##SFX(hmax2_##SFX(a, b), c); }        \
    static inline unsigned char hpick_##SFX(T vm, T vx, T vy) {                                      \
        if (vm >= vx && vm >= vy) return 0;                                                          \
        return vx >= vy ? 1 : 2;                                                                     \
    }                                                                                                \
    typedef struct {                                                                                 \
        const uint8_t *a, *b;                                                                        \
        int na, pmax, cutoff;                                                                        \
        T blk[64 * 64];                                                                              \
        T go_a, ge_a, go_b, ge_b;                                                                    \
        T *row[15]; /* sweep rows p/c (M X Y) x 2 halves' outputs + 3 scratch */                    \
        T *rmp;                                                                                      \
        unsigned char *D;                                                                            \
        uint8_t *hops;                                                                               \
        int nh;                                                                                      \
    } hbctx_##SFX;                                                                                   \
    /* One half sweep (aligner.hpp:4150-4255).  rev = 0: rows (i0, i1] consume a[r-1] in order,  \
       column c pairs b[j0 + c - 1]; rev = 1: the mirror, row r = i1-1 .. i0 consumes a[r] and   \
       column d pairs b[j1 - d].  Row H of the sweep lands in oM / oX / oY (index c, resp. d). */ \
    static void hb_sweep_##SFX(hbctx_##SFX *x, int rev, int i0, int i1, int j0, int j1, int flag, \
                               T *oM, T *oX, T *oY) {                                                \
        const int W = j1 - j0, H = i1 - i0, na = x->na;                                             \
        const T go_a = x->go_a, ge_a = x->ge_a, go_b = x->go_b, ge_b = x->ge_b;                    \
        T *pM = x->row[0], *pX = x->row[1], *pY = x->row[2];                                       \
        T *cM = x->row[3], *cX = x->row[4], *cY = x->row[5];                                       \
        pM[0] = flag ? NINF : (T)0;                                                                  \
        pX[0] = NINF;                                                                                \
        pY[0] = NINF;                                                                                \
        for (int c = 1; c <= W; ++c) {                                                               \
            pM[c] = NINF;                                                                            \
            pX[c] = NINF;                                                                            \
            pY[c] = hmax3_##SFX(pM[c - 1] - go_a - ge_a, pX[c - 1] - go_a - ge_a, pY[c - 1] - ge_a); \
        }                                                                                            \
        for (int k = 1; k <= H; ++k) {                                                               \
            const int ai = rev ? x->a[i1 - k] : x->a[i0 + k - 1];                                   \
            const T *srow = x->blk + (size_t)ai * na;                                                \
            cM[0] = NINF;                                                                            \
            cX[0] = hmax3_##SFX(pM[0] - go_b - ge_b, pX[0] - ge_b, pY[0] - go_b - ge_b);            \
            if (flag && k == 1) cX[0] = (T)0 - ge_b;                                                 \
            cY[0] = NINF;                                                                            \
            T P = cY[0] - ge_a;                                                                      \
            for (int c = 1; c <= W; ++c) {                                                           \
                const int bj = rev ? x->b[j1 - c] : x->b[j0 + c - 1];                               \
                cM[c] = hmax3_##SFX(pM[c - 1], pX[c - 1], pY[c - 1]) + srow[bj];                    \
                cX[c] = hmax3_##SFX(pM[c] - go_b - ge_b, pX[c] - ge_b, pY[c] - go_b - ge_b);        \
                if (x->pmax) {                                                                       \
                    const T g = hmax2_##SFX(cM[c - 1], cX[c - 1]) - go_a - ge_a;                    \
                    P = hmax2_##SFX(g + x->rmp[c - 1], P);                                           \
                    cY[c] = P - x->rmp[c - 1];                                                       \
                } else {                                                                             \
                    cY[c] = hmax3_##SFX(cM[c - 1] - go_a - ge_a, cX[c - 1] - go_a - ge_a, cY[c - 1] - ge_a); \
                }                                                                                    \
            }                                                                                        \
            T *t;                                                                                    \
            t = pM; pM = cM; cM = t;                                                                 \
            t = pX; pX = cX; cX = t;                                                                 \
            t = pY; pY = cY; cY = t;                                                                 \
        }                                                                                            \
        memcpy(oM, pM, sizeof(T) * ((size_t)W + 1));                                                 \
        memcpy(oX, pX, sizeof(T) * ((size_t)W + 1));                                                 \
        memcpy(oY, pY, sizeof(T) * ((size_t)W + 1));                                                 \
    }                                                                                                \
    /* The base case (aligner.hpp:4313-4400): the Pointers fill of rows (i0, i1] x columns      \
       (j0, j1] seeded by in_x / out_x, walked back from (H, W); the block's moves are appended   \
       to hops in forward order (0 diagonal, 1 a gap in B, 2 a gap in A). */                      \
    static void hb_base_##SFX(hbctx_##SFX *x, int i0, int i1, int j0, int j1, int in_x, int out_x) { \
        const int H = i1 - i0, W = j1 - j0, na = x->na;                                             \
        const T go_a = x->go_a, ge_a = x->ge_a, go_b = x->go_b, ge_b = x->ge_b;                    \
        if (H == 0 && W == 0) return;                                                                \
        const size_t st = (size_t)W + 1;                                                             \
        unsigned char *D = x->D;                                                                     \
        T *pM = x->row[0], *pX = x->row[1], *pY = x->row[2];                                       \
        T *cM = x->row[3], *cX = x->row[4], *cY = x->row[5];                                       \
        pM[0] = in_x ? NINF : (T)0;                                                                  \
        pX[0] = NINF;                                                                                \
        pY[0] = NINF;                                                                                \
        D[0] = (unsigned char)(3 | 3 << 2 | 3 << 4);                                                 \
        for (int c = 1; c <= W; ++c) {                                                               \
            pM[c] = NINF;                                                                            \
            pX[c] = NINF;                                                                            \
            const T uy = pM[c - 1] - go_a - ge_a, vy = pX[c - 1] - go_a - ge_a, wy = pY[c - 1] - ge_a; \
            pY[c] = hmax3_##SFX(uy, vy, wy);                                                         \
            D[c] = (unsigned char)(3 | 3 << 2 | hpick_##SFX(uy, vy, wy) << 4);                       \
        }                                                                                            \
        for (int r = 1; r <= H; ++r) {                                                               \
            const size_t rb = (size_t)r * st;                                                        \
            const T *srow = x->blk + (size_t)x->a[i0 + r - 1] * na;                                 \
            const T ux0 = pM[0] - go_b - ge_b, vx0 = pX[0] - ge_b, wx0 = pY[0] - go_b - ge_b;       \
            cM[0] = NINF;                                                                            \
            cX[0] = hmax3_##SFX(ux0, vx0, wx0);                                                      \
            cY[0] = NINF;                                                                            \
            D[rb] = (unsigned char)(3 | hpick_##SFX(ux0, vx0, wx0) << 2 | 3 << 4);                   \
            if (in_x && r == 1) { cX[0] = (T)0 - ge_b; D[rb] = (unsigned char)(3 | 1 << 2 | 3 << 4); } \
            for (int c = 1; c <= W; ++c) {                                                           \
                const T dgM = pM[c - 1], dgX = pX[c - 1], dgY = pY[c - 1];                           \
                cM[c] = hmax3_##SFX(dgM, dgX, dgY) + srow[x->b[j0 + c - 1]];                         \
                const T ux = pM[c] - go_b - ge_b, vx = pX[c] - ge_b, wx = pY[c] - go_b - ge_b;       \
                cX[c] = hmax3_##SFX(ux, vx, wx);                                                     \
                const T uy = cM[c - 1] - go_a - ge_a, vy = cX[c - 1] - go_a - ge_a, wy = cY[c - 1] - ge_a; \
                cY[c] = hmax3_##SFX(uy, vy, wy);                                                     \
                D[rb + c] = (unsigned char)(hpick_##SFX(dgM, dgX, dgY) | hpick_##SFX(ux, vx, wx) << 2 | \
                                            hpick_##SFX(uy, vy, wy) << 4);                           \
            }                                                                                        \
            T *t;                                                                                    \
            t = pM; pM = cM; cM = t;                                                                 \
            t = pX; pX = cX; cX = t;                                                                 \
            t = pY; pY = cY; cY = t;                                                                 \
        }                                                                                            \
        int r = H, c = W, tbl;                                                                       \
        if (out_x) tbl = 1;                                                                          \
        else {                                                                                       \
            const T vm = pM[W], vx = pX[W], vy = pY[W];                                              \
            tbl = (vm >= vx && vm >= vy) ? 0 : ((vx >= vy) ? 1 : 2);                                 \
        }                                                                                            \
        const int first = x->nh;                                                                     \
        while (r > 0 || c > 0) {                                                                     \
            const unsigned char p = D[(size_t)r * st + (size_t)c];                                   \
            int code;                                                                                \
            if (tbl == 0) { code = p & 3; x->hops[x->nh++] = 0; --r; --c; }                         \
            else if (tbl == 1) { code = (p >> 2) & 3; x->hops[x->nh++] = 1; --r; }                   \
            else { code = p >> 4; x->hops[x->nh++] = 2; --c; }                                       \
            tbl = (code == 1) ? 1 : (code == 2 ? 2 : 0);                                             \
        }                                                                                            \
        for (int lo = first, hi = x->nh - 1; lo < hi; ++lo, --hi) {                                  \
            const uint8_t t = x->hops[lo];                                                           \
            x->hops[lo] = x->hops[hi];                                                               \
            x->hops[hi] = t;                                                                         \
        }                                                                                            \
    }                                                                                                \
    /* The recursion and its join (aligner.hpp:4456-4518).  The output rows of a level live in   \
       freshly allocated arrays (the recursion may nest 30 levels deep). */                       \
    static int hb_solve_##SFX(hbctx_##SFX *x, int i0, int i1, int j0, int j1, int in_x, int out_x) { \
        const int H = i1 - i0, W = j1 - j0;                                                          \
        if (H <= x->cutoff) { hb_base_##SFX(x, i0, i1, j0, j1, in_x, out_x); return 0; }            \
        const int p = i0 + H / 2;                                                                    \
        T *buf = (T *)malloc(sizeof(T) * 6 * ((size_t)W + 1));                                       \
        if (!buf) return -1;                                                                         \
        T *FM = buf, *FX = buf + (W + 1), *FY = buf + 2 * (W + 1);                                   \
        T *rM = buf + 3 * (W + 1), *rX = buf + 4 * (W + 1), *rY = buf + 5 * (W + 1);                \
        hb_sweep_##SFX(x, 0, i0, p, j0, j1, in_x, FM, FX, FY);                                      \
        hb_sweep_##SFX(x, 1, p, i1, j0, j1, out_x, rM, rX, rY);                                     \
        T best = NINF;                                                                               \
        int cstar = 0, span = 0;                                                                     \
        for (int c = 0; c <= W; ++c) {                                                               \
            const int d = W - c;                                                                     \
            const T f = hmax3_##SFX(FM[c], FX[c], FY[c]);                                            \
            const T g = hmax3_##SFX(rM[d], rX[d], rY[d]);                                            \
            const T plain = f + g;                                                                   \
            if (plain > best) { best = plain; cstar = c; span = 0; }                                 \
            if (FX[c] > NINF && rX[d] > NINF) {                                                      \
                const T joined = (FX[c] + rX[d]) + x->go_b;                                          \
                if (joined > best) { best = joined; cstar = c; span = 1; }                           \
            }                                                                                        \
        }                                                                                            \
        free(buf);                                                                                   \
        const int jstar = j0 + cstar;                                                                \
        if (hb_solve_##SFX(x, i0, p, j0, jstar, in_x, span)) return -1;                             \
        return hb_solve_##SFX(x, p, i1, jstar, j1, span, out_x);                                    \
    }                                                                                                \
    /* One splitting pair: its moves (m + n at most) into hops, *nh of them; the replayed score  \
       (aligner.hpp:4525-4560) into *score.  Returns 0, or -1 on allocation failure. */            \
    static int hb_pair_##SFX(const nwo_hb_params *pp, int pmax, int cutoff, const uint8_t *a, int m,  \
                             const uint8_t *b, int n, uint8_t *hops, int *nh, double *score) {      \
        hbctx_##SFX *x = (hbctx_##SFX *)calloc(1, sizeof(hbctx_##SFX));                            \
        if (!x) return -1;                                                                           \
        x->a = a; x->b = b; x->na = pp->n_alpha; x->pmax = pmax; x->cutoff = cutoff;                \
        for (int k = 0; k < x->na * x->na; ++k) x->blk[k] = (T)pp->matrix[k];                       \
        x->go_a = (T)pp->go_a; x->ge_a = (T)pp->ge_a; x->go_b = (T)pp->go_b; x->ge_b = (T)pp->ge_b; \
        const size_t st = (size_t)n + 1;                                                             \
        T *rows = (T *)malloc(sizeof(T) * 6 * st);                                                   \
        x->rmp = (T *)malloc(sizeof(T) * st);                                                        \
        x->D = (unsigned char *)malloc(((size_t)cutoff + 1) * st);                                   \
        int rc = -1;                                                                                 \
        if (rows && x->rmp && x->D) {                                                                \
            for (int k = 0; k < 6; ++k) x->row[k] = rows + (size_t)k * st;                          \
            /* the ramp k * ge_a, one multiply of an exact integer (aligner.hpp:4428-4436) */      \
            for (int k = 0; k <= n; ++k) x->rmp[k] = (T)k * x->ge_a;                                 \
            x->hops = hops;                                                                          \
            x->nh = 0;                                                                               \
            rc = hb_solve_##SFX(x, 0, m, 0, n, 0, 0);                                                \
        }                                                                                            \
        if (!rc) {                                                                                   \
            T s = (T)0;                                                                              \
            int i = 0, j = 0, prev = 3;                                                              \
            for (int q = 0; q < x->nh; ++q) {                                                        \
                const int op = hops[q];                                                              \
                if (op == 0) { ++i; ++j; s += x->blk[(size_t)a[i - 1] * x->na + b[j - 1]]; }         \
                else if (op == 1) { ++i; if (prev != 1) s -= x->go_b; s -= x->ge_b; }               \
                else { if (prev != 2) s -= x->go_a; s -= x->ge_a; ++j; }                            \
                prev = op;                                                                           \
            }                                                                                        \
            *score = (double)s;                                                                      \
            *nh = x->nh;                                                                             \
        }                                                                                            \
        free(rows); free(x->rmp); free(x->D); free(x);                                               \
        return rc;                                                                                   \
    }

NWO_DEFINE_HB(float, f32, -INFINITY)
NWO_DEFINE_HB(double, f64, -INFINITY)

/* Batch entry: every pair of more than `cutoff` rows through the recursion (traceback
 * "hirschberg" with pmax = 0, "hirschberg_pmax" with pmax = 1); the others are left alone
 * (split[q] = 0) -- the reference runs them as Pointers.  For a split pair q:
 *   scores[q], grads row q (N * (na*na + 4); aligner.hpp:5191-5207),
 *   guides (pair q's m + 1 entries at off_a[q] + q, may be NULL; aligner.hpp:5233-5250),
 *   moves (pair q's at off_a[q] + off_b[q], m + n of room, may be NULL) and n_moves[q].
 * Returns 0, or -1 on allocation failure. */
int nwo_hirschberg(const nwo_hb_params *p, int precision, int pmax, int cutoff, int64_t N, const uint8_t *ca,
                   const int64_t *oa, const uint8_t *cb, const int64_t *ob, double *scores, double *grads,
                   int32_t *guides, uint8_t *moves, int32_t *n_moves, uint8_t *split) {
    const int na = p->n_alpha, nn = na * na, gs = nn + 4;
    for (int64_t q = 0; q < N; ++q) {
        const uint8_t *a = ca + oa[q], *b = cb + ob[q];
        const int m = (int)(oa[q + 1] - oa[q]), n = (int)(ob[q + 1] - ob[q]);
        split[q] = (uint8_t)(m > cutoff);
        if (!split[q]) continue;
        uint8_t *hops = (uint8_t *)malloc((size_t)m + (size_t)n + 1);
        if (!hops) return -1;
        int nh = 0;
        double sc = 0;
        const int rc = precision ? hb_pair_f64(p, pmax, cutoff, a, m, b, n, hops, &nh, &sc)
                                 : hb_pair_f32(p, pmax, cutoff, a, m, b, n, hops, &nh, &sc);
        if (rc) { free(hops); return -1; }
        scores[q] = sc;
        double *g = grads + (size_t)q * gs;
        memset(g, 0, sizeof(double) * (size_t)gs);
        int32_t *gj = guides ? guides + oa[q] + q : NULL;
        if (gj) {
            for (int k = 0; k <= m; ++k) gj[k] = -1;
            gj[0] = 0;
        }
        int i = 0, j = 0, prev = 3;
        for (int k = 0; k < nh; ++k) {
            const int op = hops[k];
            if (op == 0) { ++i; ++j; g[a[i - 1] * na + b[j - 1]] += 1.0; if (gj) gj[i] = j; }
            else if (op == 1) { ++i; if (prev != 1) g[nn + 2] -= 1.0; g[nn + 3] -= 1.0; if (gj) gj[i] = j; }
            else { if (prev != 2) g[nn + 0] -= 1.0; g[nn + 1] -= 1.0; ++j; }
            prev = op;
        }
        if (gj) {
            gj[m] = n;
            hb_fill_guide_gaps(gj, m + 1);
        }
        if (moves) memcpy(moves + oa[q] + ob[q], hops, (size_t)nh);
        if (n_moves) n_moves[q] = nh;
        free(hops);
    }
    return 0;
}
