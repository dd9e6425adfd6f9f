/* TEST INFRASTRUCTURE ONLY -- the CPU oracle of this repository.
 *
 * A plain-C restatement of the reference nwgrad alignment hot path, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.  The product
 * (nwgrad_amd/) never links, loads or calls it.
 *
 * What it restates (all citations relative to the reference's src/nwgrad/cpp/nwgrad/):
 *   - affine Viterbi fill with 1-byte predecessor codes  aligner.hpp:3955-4071
 *   - linear Viterbi fill with direction codes           aligner.hpp:3756-3865
 *   - local end cell (first strict row maximum)          aligner.hpp:5975-5984
 *   - glocal end cell (row m, first strict maximum)       aligner.hpp:6121-6134
 *   - affine walk with border runs                        aligner.hpp:2986-3056
 *   - linear walk                                         aligner.hpp:2826-2855
 *   - hard gradient counts (sign convention)              aligner.hpp:3177-3197, 3886-3925
 *   - log-space forward-backward + soft gradient (double) aligner.hpp:6231-6524
 *   - blocked weighted sum of per-pair gradients           batch_engine.hpp:697-718, 1788-1827
 * Parity pins (tests/test_oracle.py): the reference's known-answer tests, brute-force path
 * enumeration, closed forms, score == sum(counts * params) along the path, and the
 * reference engine built from its own sources (oracle/_ref, when present).
 *
 * Arithmetic: every penalty is converted once, T(double); the cell operations are the
 * reference's, in its order, compiled with -ffp-contract=off.  MAX2(a,b) = a < b ? b : a is
 * std::max; MAX3 is std::max({a,b,c}) (the first of equal maxima).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int gap_model; /* 0 linear, 1 affine */
    int mode;      /* 0 global, 1 local, 2 glocal */
    int n_alpha;
    const double *matrix; /* n_alpha * n_alpha, row-major: matrix[a * n + b] */
    double go_a, ge_a, go_b, ge_b;
} nwo_params;

enum { ST_M = 0, ST_X = 1, ST_Y = 2 };

/* ── the hard fill, restated once per scalar type ─────────────────────────────── */
#define NWO_DEFINE_HARD(T, SUFFIX, NINF)                                                        \
    static inline T max2_##SUFFIX(T a, T b) { return a < b ? b : a; }                            \
    static inline T max3_##SUFFIX(T a, T b, T c) { return max2_##SUFFIX(max2_##SUFFIX(a, b), c); } \
    static inline unsigned char pick_##SUFFIX(T vm, T vx, T vy) {                                \
        const unsigned char kxy = vx >= vy ? 1 : 2;                                            \
        return vm >= max2_##SUFFIX(vx, vy) ? 0 : kxy;                                          \
    }                                                                                          \
    /* affine: aligner.hpp:3955-4071.  D holds (m+1)(n+1) codes M | X<<2 | Y<<4.  Returns the  \
       score; *bi, *bj, *bst the end cell and state. */                                        \
    static T affine_fill_##SUFFIX(const nwo_params *p, const uint8_t *a, int m, const uint8_t *b, \
                                  int n, unsigned char *D, T *rows, int *bi, int *bj, int *bst) { \
        const T go_a = (T)p->go_a, ge_a = (T)p->ge_a, go_b = (T)p->go_b, ge_b = (T)p->ge_b;     \
        const int AM = p->mode, na = p->n_alpha;                                               \
        const size_t st = (size_t)n + 1;                                                       \
        T *pM = rows, *pX = rows + st, *pY = rows + 2 * st;                                    \
        T *cM = rows + 3 * st, *cX = rows + 4 * st, *cY = rows + 5 * st;                       \
        T blk[64 * 64];                                                                        \
        for (int k = 0; k < na * na; ++k) blk[k] = (T)p->matrix[k];                            \
        const int free_row0 = (AM == 1 || AM == 2);                                            \
        for (int j = 0; j <= n; ++j) {                                                         \
            if (free_row0) { pM[j] = 0; pX[j] = NINF; pY[j] = NINF; }                          \
            else {                                                                             \
                pM[j] = (j == 0) ? (T)0 : NINF;                                                \
                pX[j] = NINF;                                                                  \
                pY[j] = (j == 0) ? NINF : -(go_a + (T)j * ge_a);                               \
            }                                                                                  \
            D[j] = (j <= 1 || free_row0) ? 3 : (3 | 2 << 4);                                   \
        }                                                                                      \
        *bi = 0; *bj = 0; *bst = ST_M;                                                         \
        T best_local = 0;                                                                      \
        for (int i = 1; i <= m; ++i) {                                                         \
            const size_t rb = (size_t)i * st;                                                  \
            const T *srow = blk + (size_t)a[i - 1] * na;                                       \
            if (AM == 1) { cM[0] = 0; cX[0] = NINF; }                                          \
            else { cM[0] = NINF; cX[0] = -(go_b + (T)i * ge_b); }                              \
            cY[0] = NINF;                                                                      \
            D[rb] = (i <= 1) ? 3 : (3 | 1 << 2);                                               \
            T rowmax = 0;                                                                      \
            for (int j = 1; j <= n; ++j) {                                                     \
                const T dgM = pM[j - 1], dgX = pX[j - 1], dgY = pY[j - 1];                     \
                T m_val = max3_##SUFFIX(dgM, dgX, dgY) + srow[b[j - 1]];                       \
                unsigned char km = pick_##SUFFIX(dgM, dgX, dgY);                               \
                const T ux = (pM[j] - go_b) - ge_b, vx = pX[j] - ge_b, wx = (pY[j] - go_b) - ge_b; \
                const T x_val = max3_##SUFFIX(ux, vx, wx);                                     \
                const unsigned char kx = pick_##SUFFIX(ux, vx, wx);                            \
                const T uy = (cM[j - 1] - go_a) - ge_a, vy = (cX[j - 1] - go_a) - ge_a,        \
                        wy = cY[j - 1] - ge_a;                                                 \
                const T y_val = max3_##SUFFIX(uy, vy, wy);                                     \
                const unsigned char ky = pick_##SUFFIX(uy, vy, wy);                            \
                if (AM == 1) {                                                                 \
                    m_val = max2_##SUFFIX(m_val, (T)0);                                        \
                    rowmax = max2_##SUFFIX(rowmax, max3_##SUFFIX(m_val, x_val, y_val));        \
                    if (m_val <= (T)0) km = 3;                                                 \
                }                                                                              \
                cM[j] = m_val; cX[j] = x_val; cY[j] = y_val;                                   \
                D[rb + j] = (unsigned char)(km | kx << 2 | ky << 4);                          \
            }                                                                                  \
            if (AM == 1 && rowmax > best_local) { /* local_row_best_, aligner.hpp:5975-5984 */ \
                int j = 1;                                                                     \
                while (!(max3_##SUFFIX(cM[j], cX[j], cY[j]) == rowmax)) ++j;                   \
                best_local = rowmax; *bi = i; *bj = j;                                         \
                if (cM[j] >= cX[j] && cM[j] >= cY[j]) *bst = ST_M;                             \
                else if (cX[j] >= cY[j]) *bst = ST_X;                                          \
                else *bst = ST_Y;                                                              \
            }                                                                                  \
            T *t;                                                                              \
            t = pM; pM = cM; cM = t; t = pX; pX = cX; cX = t; t = pY; pY = cY; cY = t;         \
        }                                                                                      \
        if (AM == 0) {                                                                         \
            const T vm = pM[n], vx = pX[n], vy = pY[n];                                        \
            *bi = m; *bj = n;                                                                  \
            if (vm >= vx && vm >= vy) *bst = ST_M;                                             \
            else if (vx >= vy) *bst = ST_X;                                                    \
            else *bst = ST_Y;                                                                  \
            return max3_##SUFFIX(vm, vx, vy);                                                  \
        }                                                                                      \
        if (AM == 2) { /* glocal_end_rows_, aligner.hpp:6121-6129 */                            \
            int bj2 = 0;                                                                       \
            T best = max3_##SUFFIX(pM[0], pX[0], pY[0]);                                       \
            for (int j = 1; j <= n; ++j) {                                                     \
                const T v = max3_##SUFFIX(pM[j], pX[j], pY[j]);                                \
                if (v > best) { best = v; bj2 = j; }                                           \
            }                                                                                  \
            *bi = m; *bj = bj2;                                                                \
            const T vm = pM[bj2], vx = pX[bj2], vy = pY[bj2];                                  \
            if (vm >= vx && vm >= vy) *bst = ST_M;                                             \
            else if (vx >= vy) *bst = ST_X;                                                    \
            else *bst = ST_Y;                                                                  \
            return best;                                                                       \
        }                                                                                      \
        return best_local;                                                                     \
    }                                                                                          \
    /* linear: aligner.hpp:3756-3865 (codes 0 diag, 1 up, 2 left, 3 stop) */                   \
    static T linear_fill_##SUFFIX(const nwo_params *p, const uint8_t *a, int m, const uint8_t *b, \
                                  int n, unsigned char *D, T *rows, int *bi, int *bj) {        \
        const T ge_a = (T)p->ge_a, ge_b = (T)p->ge_b;                                          \
        const int AM = p->mode, na = p->n_alpha, local = (AM == 1), free0 = (AM == 1 || AM == 2); \
        const size_t st = (size_t)n + 1;                                                       \
        T *prev = rows, *cur = rows + st;                                                      \
        T blk[64 * 64];                                                                        \
        for (int k = 0; k < na * na; ++k) blk[k] = (T)p->matrix[k];                            \
        prev[0] = 0; D[0] = 3;                                                                 \
        for (int j = 1; j <= n; ++j) {                                                         \
            prev[j] = free0 ? (T)0 : -(T)j * ge_a;                                             \
            D[j] = free0 ? 3 : 2;                                                              \
        }                                                                                      \
        int bbi = 0, bbj = 0;                                                                  \
        T best_local = 0;                                                                      \
        for (int i = 1; i <= m; ++i) {                                                         \
            unsigned char *d = D + (size_t)i * st;                                             \
            const T *prow = blk + (size_t)a[i - 1] * na;                                       \
            cur[0] = local ? (T)0 : -(T)i * ge_b;                                              \
            d[0] = local ? 3 : 1;                                                              \
            T left = cur[0];                                                                   \
            for (int j = 1; j <= n; ++j) {                                                     \
                const T d0 = prev[j - 1] + prow[b[j - 1]];                                     \
                const T u = prev[j] - ge_b;                                                    \
                const T t = max2_##SUFFIX(d0, u);                                              \
                unsigned char c = (unsigned char)(d0 < u);                                     \
                const T l = left - ge_a;                                                       \
                T v = max2_##SUFFIX(t, l);                                                     \
                if (t < l) c = 2;                                                              \
                if (local) {                                                                   \
                    v = max2_##SUFFIX(v, (T)0);                                                \
                    if (v <= (T)0) c = (unsigned char)(c | 3);                                 \
                    if (v > best_local) { best_local = v; bbi = i; bbj = j; }                  \
                }                                                                              \
                cur[j] = v; d[j] = c;                                                          \
                left = v;                                                                      \
            }                                                                                  \
            T *tt = prev; prev = cur; cur = tt;                                                \
        }                                                                                      \
        if (AM == 2) { /* glocal_end_row_, aligner.hpp:6130-6134 */                             \
            int bj2 = 0;                                                                       \
            T best = prev[0];                                                                  \
            for (int j = 1; j <= n; ++j) if (prev[j] > best) { best = prev[j]; bj2 = j; }      \
            *bi = m; *bj = bj2;                                                                \
            return best;                                                                       \
        }                                                                                      \
        if (local) { *bi = bbi; *bj = bbj; return best_local; }                               \
        *bi = m; *bj = n;                                                                      \
        return prev[n];                                                                        \
    }

NWO_DEFINE_HARD(float, f32, -INFINITY)
NWO_DEFINE_HARD(double, f64, -INFINITY)

/* ── walks and counts (types only enter through the codes) ──────────────────────── */

/* Counts go to g[0 .. na*na) (matrix, row a, column b) and g[na*na + 0..3]
   (go_a, ge_a, go_b, ge_b), all negative for the gap fields: aligner.hpp:3177-3197. */
static void emit_affine(double *g, int na, const uint8_t *a, const uint8_t *b, int st, int i, int j,
                        int open) {
    const int nn = na * na;
    if (st == ST_M) g[a[i - 1] * na + b[j - 1]] += 1.0;
    else if (st == ST_X) { g[nn + 3] -= 1.0; if (open) g[nn + 2] -= 1.0; }
    else { g[nn + 1] -= 1.0; if (open) g[nn + 0] -= 1.0; }
}

/* aligner.hpp:2986-3056 (Ptr, untracked).  Border runs along row 0 / column 0 are
   emitted without reading codes for Global and Glocal (glocal stops at row 0). */
static void walk_affine(int mode, const unsigned char *D, size_t st, const uint8_t *a,
                        const uint8_t *b, int na, int i, int j, int tbl, double *g) {
    for (;;) {
        if (i == 0 && j == 0) break;
        unsigned char c = 0;
        if (mode == 1) {
            c = D[(size_t)i * st + j];
            if (tbl == ST_M && (c & 3) == 3) break;
        } else {
            if (i == 0) {
                if (mode == 2) break;
                for (; j > 0; --j) emit_affine(g, na, a, b, ST_Y, 0, j, j == 1);
                break;
            }
            if (j == 0) {
                for (; i > 0; --i) emit_affine(g, na, a, b, ST_X, i, 0, i == 1);
                break;
            }
            c = D[(size_t)i * st + j];
        }
        if (i == 0) tbl = ST_Y;
        else if (j == 0) tbl = ST_X;
        int prev;
        if (tbl == ST_M) {
            const int k = c & 3;
            prev = k == 1 ? ST_X : (k == 2 ? ST_Y : ST_M);
            emit_affine(g, na, a, b, ST_M, i, j, 0);
            --i; --j;
        } else if (tbl == ST_X) {
            const int k = (c >> 2) & 3;
            prev = k == 1 ? ST_X : (k == 2 ? ST_Y : ST_M);
            emit_affine(g, na, a, b, ST_X, i, j, prev != ST_X);
            --i;
        } else {
            const int k = c >> 4;
            prev = k == 1 ? ST_X : (k == 2 ? ST_Y : ST_M);
            emit_affine(g, na, a, b, ST_Y, i, j, prev != ST_Y);
            --j;
        }
        tbl = prev;
    }
}

/* aligner.hpp:2826-2855 (pointers): codes read at every cell, borders included. */
static void walk_linear(int mode, const unsigned char *D, size_t st, const uint8_t *a,
                        const uint8_t *b, int na, int i, int j, double *g) {
    const int nn = na * na;
    for (;;) {
        if (i == 0 && j == 0) break;
        if (mode == 2 && i == 0) break;
        const int k = D[(size_t)i * st + j];
        if (k == 3) break;
        if (k == 0) { g[a[i - 1] * na + b[j - 1]] += 1.0; --i; --j; }
        else if (k == 1) { g[nn + 3] -= 1.0; --i; }
        else { g[nn + 1] -= 1.0; --j; }
    }
}

/* Batch entry: hard (with_grad = 1) or score-only (with_grad = 0) for N pairs.
   precision 0 = float32 (SeqPairBatch), 1 = float64 (SeqPairBatchDouble).
   grads: N * (na*na + 4) doubles, row per pair (may be NULL when with_grad = 0).
   Returns 0, or -1 on allocation failure. */
int nwo_hard(const nwo_params *p, int precision, int64_t N, const uint8_t *ca, const int64_t *oa,
             const uint8_t *cb, const int64_t *ob, int with_grad, double *scores, double *grads) {
    const int na = p->n_alpha, gs = na * na + 4;
    size_t maxcells = 1, maxrow = 1;
    for (int64_t q = 0; q < N; ++q) {
        const size_t m = (size_t)(oa[q + 1] - oa[q]), n = (size_t)(ob[q + 1] - ob[q]);
        if ((m + 1) * (n + 1) > maxcells) maxcells = (m + 1) * (n + 1);
        if (n + 1 > maxrow) maxrow = n + 1;
    }
    unsigned char *D = (unsigned char *)malloc(maxcells);
    double *rows = (double *)malloc(6 * maxrow * sizeof(double));
    if (!D || !rows) { free(D); free(rows); return -1; }
    for (int64_t q = 0; q < N; ++q) {
        const uint8_t *a = ca + oa[q], *b = cb + ob[q];
        const int m = (int)(oa[q + 1] - oa[q]), n = (int)(ob[q + 1] - ob[q]);
        int bi, bj, bst = ST_M;
        double sc;
        if (p->gap_model == 1) {
            sc = precision ? affine_fill_f64(p, a, m, b, n, D, rows, &bi, &bj, &bst)
                           : (double)affine_fill_f32(p, a, m, b, n, D, (float *)rows, &bi, &bj, &bst);
        } else {
            sc = precision ? linear_fill_f64(p, a, m, b, n, D, rows, &bi, &bj)
                           : (double)linear_fill_f32(p, a, m, b, n, D, (float *)rows, &bi, &bj);
        }
        scores[q] = sc;
        if (with_grad) {
            double *g = grads + (size_t)q * gs;
            memset(g, 0, (size_t)gs * sizeof(double));
            if (p->gap_model == 1) walk_affine(p->mode, D, (size_t)n + 1, a, b, na, bi, bj, bst, g);
            else walk_linear(p->mode, D, (size_t)n + 1, a, b, na, bi, bj, g);
        }
    }
    free(D);
    free(rows);
    return 0;
}

/* ── soft: log-space forward-backward in double, aligner.hpp:6231-6524 ─────────────
   With temperature Tm the path weights are params/Tm; the score is Tm*log Z and the
   gradient the expected counts under params/Tm (aligner.hpp:6589-6626). */
static inline double lse2(double a, double b) {
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    return a > b ? a + log1p(exp(b - a)) : b + log1p(exp(a - b));
}
static inline double lse3(double a, double b, double c) { return lse2(lse2(a, b), c); }

static void soft_affine_pair(const nwo_params *p, double Tm, const uint8_t *a, int m, const uint8_t *b,
                             int n, double *buf, double *score, double *g) {
    const int AM = p->mode, na = p->n_alpha, nn = na * na;
    const double it = 1.0 / Tm;
    const double goa = p->go_a * it, gea = p->ge_a * it, gob = p->go_b * it, geb = p->ge_b * it;
    const size_t st = (size_t)n + 1, sz = (size_t)(m + 1) * st;
    double *FM = buf, *FX = buf + sz, *FY = buf + 2 * sz, *BM = buf + 3 * sz, *BX = buf + 4 * sz,
           *BY = buf + 5 * sz;
#define AT(t, i, j) (t)[(size_t)(i) * st + (size_t)(j)]
#define SUB(i, j) (p->matrix[a[(i) - 1] * na + b[(j) - 1]] * it)
    for (size_t k = 0; k < sz; ++k) { FM[k] = FX[k] = FY[k] = -INFINITY; }
    if (AM != 1) {
        AT(FM, 0, 0) = 0.0;
        for (int i = 1; i <= m; ++i) AT(FX, i, 0) = -(gob + (double)i * geb);
        if (AM == 0) for (int j = 1; j <= n; ++j) AT(FY, 0, j) = -(goa + (double)j * gea);
        if (AM == 2) for (int j = 0; j <= n; ++j) AT(FM, 0, j) = 0.0;
    } else {
        for (int i = 0; i <= m; ++i) AT(FM, i, 0) = 0.0;
        for (int j = 0; j <= n; ++j) AT(FM, 0, j) = 0.0;
    }
    for (int i = 1; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const double diag = lse3(AT(FM, i - 1, j - 1), AT(FX, i - 1, j - 1), AT(FY, i - 1, j - 1));
            double mv = diag + SUB(i, j);
            const double xv = lse3(AT(FM, i - 1, j) - gob - geb, AT(FX, i - 1, j) - geb,
                                   AT(FY, i - 1, j) - gob - geb);
            const double yv = lse3(AT(FM, i, j - 1) - goa - gea, AT(FX, i, j - 1) - goa - gea,
                                   AT(FY, i, j - 1) - gea);
            if (AM == 1) mv = lse2(mv, 0.0);
            AT(FM, i, j) = mv; AT(FX, i, j) = xv; AT(FY, i, j) = yv;
        }
    double lz;
    if (AM == 0) lz = lse3(AT(FM, m, n), AT(FX, m, n), AT(FY, m, n));
    else if (AM == 2) {
        lz = -INFINITY;
        for (int j = 0; j <= n; ++j) lz = lse2(lz, lse3(AT(FM, m, j), AT(FX, m, j), AT(FY, m, j)));
    } else {
        lz = -INFINITY;
        for (int i = 0; i <= m; ++i)
            for (int j = 0; j <= n; ++j) lz = lse2(lz, lse3(AT(FM, i, j), AT(FX, i, j), AT(FY, i, j)));
    }
    const double e0 = (AM == 1) ? 0.0 : -INFINITY;
    for (size_t k = 0; k < sz; ++k) { BM[k] = BX[k] = BY[k] = e0; }
    if (AM == 0) { AT(BM, m, n) = AT(BX, m, n) = AT(BY, m, n) = 0.0; }
    if (AM == 2) for (int j = 0; j <= n; ++j) { AT(BM, m, j) = AT(BX, m, j) = AT(BY, m, j) = 0.0; }
    for (int i = m; i >= 0; --i)
        for (int j = n; j >= 0; --j) {
            const double bm = AT(BM, i, j), bx = AT(BX, i, j), by = AT(BY, i, j);
            if (bm != -INFINITY && i > 0 && j > 0) {
                const double c = bm + SUB(i, j);
                AT(BM, i - 1, j - 1) = lse2(AT(BM, i - 1, j - 1), c);
                AT(BX, i - 1, j - 1) = lse2(AT(BX, i - 1, j - 1), c);
                AT(BY, i - 1, j - 1) = lse2(AT(BY, i - 1, j - 1), c);
            }
            if (bx != -INFINITY && i > 0) {
                AT(BM, i - 1, j) = lse2(AT(BM, i - 1, j), bx - gob - geb);
                AT(BX, i - 1, j) = lse2(AT(BX, i - 1, j), bx - geb);
                AT(BY, i - 1, j) = lse2(AT(BY, i - 1, j), bx - gob - geb);
            }
            if (by != -INFINITY && j > 0) {
                AT(BM, i, j - 1) = lse2(AT(BM, i, j - 1), by - goa - gea);
                AT(BX, i, j - 1) = lse2(AT(BX, i, j - 1), by - goa - gea);
                AT(BY, i, j - 1) = lse2(AT(BY, i, j - 1), by - gea);
            }
        }
    /* soft_grad_affine, aligner.hpp:6467-6524; gap targets per aligner.hpp:6357-6375 */
    const int gmin_j = (AM == 1) ? 1 : 0, gmin_i = (AM == 0) ? 0 : 1;
    memset(g, 0, (size_t)(nn + 4) * sizeof(double));
    for (int i = 1; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const double bm = AT(BM, i, j);
            if (bm == -INFINITY) continue;
            const double pred = lse3(AT(FM, i - 1, j - 1), AT(FX, i - 1, j - 1), AT(FY, i - 1, j - 1));
            g[a[i - 1] * na + b[j - 1]] += exp(pred + SUB(i, j) + bm - lz);
        }
    for (int i = 1; i <= m; ++i)
        for (int j = 0; j <= n; ++j) {
            const double fx = AT(FX, i, j), bx = AT(BX, i, j);
            if (fx == -INFINITY || bx == -INFINITY) continue;
            g[nn + 3] -= exp(fx + bx - lz);
        }
    for (int i = 0; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const double fy = AT(FY, i, j), by = AT(BY, i, j);
            if (fy == -INFINITY || by == -INFINITY) continue;
            g[nn + 1] -= exp(fy + by - lz);
        }
    for (int i = 1; i <= m; ++i)
        for (int j = gmin_j; j <= n; ++j) {
            const double bx = AT(BX, i, j);
            if (bx == -INFINITY) continue;
            const double lo = lse2(AT(FM, i - 1, j), AT(FY, i - 1, j)) - gob - geb + bx - lz;
            if (lo > -700) g[nn + 2] -= exp(lo);
        }
    for (int i = gmin_i; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const double by = AT(BY, i, j);
            if (by == -INFINITY) continue;
            const double lo = lse2(AT(FM, i, j - 1), AT(FX, i, j - 1)) - goa - gea + by - lz;
            if (lo > -700) g[nn + 0] -= exp(lo);
        }
    *score = lz * Tm;
#undef AT
#undef SUB
}

static void soft_linear_pair(const nwo_params *p, double Tm, const uint8_t *a, int m, const uint8_t *b,
                             int n, double *buf, double *score, double *g) {
    const int AM = p->mode, na = p->n_alpha, nn = na * na;
    const double it = 1.0 / Tm;
    const double gea = p->ge_a * it, geb = p->ge_b * it;
    const size_t st = (size_t)n + 1, sz = (size_t)(m + 1) * st;
    double *F = buf, *B = buf + sz;
#define AT(t, i, j) (t)[(size_t)(i) * st + (size_t)(j)]
#define SUB(i, j) (p->matrix[a[(i) - 1] * na + b[(j) - 1]] * it)
    for (size_t k = 0; k < sz; ++k) F[k] = -INFINITY;
    if (AM != 1) {
        AT(F, 0, 0) = 0.0;
        for (int i = 1; i <= m; ++i) AT(F, i, 0) = -(double)i * geb;
        if (AM == 0) for (int j = 1; j <= n; ++j) AT(F, 0, j) = -(double)j * gea;
        if (AM == 2) for (int j = 0; j <= n; ++j) AT(F, 0, j) = 0.0;
    } else {
        for (int i = 0; i <= m; ++i) AT(F, i, 0) = 0.0;
        for (int j = 0; j <= n; ++j) AT(F, 0, j) = 0.0;
    }
    for (int i = 1; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            double v = lse3(AT(F, i - 1, j - 1) + SUB(i, j), AT(F, i - 1, j) - geb, AT(F, i, j - 1) - gea);
            if (AM == 1) v = lse2(v, 0.0);
            AT(F, i, j) = v;
        }
    double lz;
    if (AM == 0) lz = AT(F, m, n);
    else if (AM == 2) { lz = -INFINITY; for (int j = 0; j <= n; ++j) lz = lse2(lz, AT(F, m, j)); }
    else {
        lz = -INFINITY;
        for (int i = 0; i <= m; ++i) for (int j = 0; j <= n; ++j) lz = lse2(lz, AT(F, i, j));
    }
    const double e0 = (AM == 1) ? 0.0 : -INFINITY;
    for (size_t k = 0; k < sz; ++k) B[k] = e0;
    if (AM == 0) AT(B, m, n) = 0.0;
    if (AM == 2) for (int j = 0; j <= n; ++j) AT(B, m, j) = 0.0;
    for (int i = m; i >= 0; --i)
        for (int j = n; j >= 0; --j) {
            const double bv = AT(B, i, j);
            if (bv == -INFINITY) continue;
            if (i > 0 && j > 0) AT(B, i - 1, j - 1) = lse2(AT(B, i - 1, j - 1), bv + SUB(i, j));
            if (i > 0) AT(B, i - 1, j) = lse2(AT(B, i - 1, j), bv - geb);
            if (j > 0) AT(B, i, j - 1) = lse2(AT(B, i, j - 1), bv - gea);
        }
    const int gmin_j = (AM == 1) ? 1 : 0, gmin_i = (AM == 0) ? 0 : 1;
    memset(g, 0, (size_t)(nn + 4) * sizeof(double));
    for (int i = 1; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const double bv = AT(B, i, j);
            if (bv == -INFINITY) continue;
            g[a[i - 1] * na + b[j - 1]] += exp(AT(F, i - 1, j - 1) + SUB(i, j) + bv - lz);
        }
    for (int i = 1; i <= m; ++i)
        for (int j = gmin_j; j <= n; ++j) {
            const double bv = AT(B, i, j), fv = AT(F, i - 1, j);
            if (bv == -INFINITY || fv == -INFINITY) continue;
            g[nn + 3] -= exp(fv - geb + bv - lz);
        }
    for (int i = gmin_i; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const double bv = AT(B, i, j), fv = AT(F, i, j - 1);
            if (bv == -INFINITY || fv == -INFINITY) continue;
            g[nn + 1] -= exp(fv - gea + bv - lz);
        }
    *score = lz * Tm;
#undef AT
#undef SUB
}

/* Batch soft entry (log space, double).  Returns 0, or -1 on allocation failure. */
int nwo_soft_log(const nwo_params *p, double temperature, int64_t N, const uint8_t *ca,
                 const int64_t *oa, const uint8_t *cb, const int64_t *ob, double *scores,
                 double *grads) {
    const int na = p->n_alpha, gs = na * na + 4;
    size_t maxcells = 1;
    for (int64_t q = 0; q < N; ++q) {
        const size_t m = (size_t)(oa[q + 1] - oa[q]), n = (size_t)(ob[q + 1] - ob[q]);
        if ((m + 1) * (n + 1) > maxcells) maxcells = (m + 1) * (n + 1);
    }
    double *buf = (double *)malloc(6 * maxcells * sizeof(double));
    if (!buf) return -1;
    for (int64_t q = 0; q < N; ++q) {
        const uint8_t *a = ca + oa[q], *b = cb + ob[q];
        const int m = (int)(oa[q + 1] - oa[q]), n = (int)(ob[q + 1] - ob[q]);
        double *g = grads + (size_t)q * gs;
        if (p->gap_model == 1) soft_affine_pair(p, temperature, a, m, b, n, buf, &scores[q], g);
        else soft_linear_pair(p, temperature, a, m, b, n, buf, &scores[q], g);
    }
    free(buf);
    return 0;
}

/* batch_engine.hpp:1788-1827: sum_i w_i * g_i, in blocks of 4096 pairs summed in pair
   order, then the block sums in block order; the product is rounded before the add
   (the reference forms it out of line, batch_engine.hpp:697-701).  w NULL: plain sum. */
void nwo_weighted_sum(int64_t N, int gs, const double *grads, const double *w, double *out) {
    const int64_t B = 4096;
    for (int k = 0; k < gs; ++k) out[k] = 0.0;
    double *acc = (double *)calloc((size_t)gs, sizeof(double));
    for (int64_t b0 = 0; b0 < N; b0 += B) {
        for (int k = 0; k < gs; ++k) acc[k] = 0.0;
        const int64_t e = b0 + B < N ? b0 + B : N;
        for (int64_t i = b0; i < e; ++i) {
            const double *g = grads + (size_t)i * gs;
            for (int k = 0; k < gs; ++k) {
                volatile double prod = w ? g[k] * w[i] : g[k];
                acc[k] += prod;
            }
        }
        for (int k = 0; k < gs; ++k) out[k] += acc[k];
    }
    free(acc);
}

/* ── tracks: K parameter sets, one per DP row (aligner.hpp:1800-1826, 1903-2008) ──────
   Row i (1..m) scores matches with matrix trk[i-1] and charges a gap in B entering it
   set trk[i-1]'s go_b / ge_b; a gap in A within row i (0..m) charges set gtrk[i]'s go_a /
   ge_a.  The Global column-0 border of row i is -(go_b(1) + C_i), C_i the sum of ge_b over
   rows 1..i accumulated by runs of bitwise-equal extends (setup_tracks_, aligner.hpp:
   1981-2007), so a track whose sets are all equal reproduces the untracked fill bit for bit.
   Each gradient count lands in the set of its row (hard_grad_affine_impl, aligner.hpp:
   3199-3230; hard_grad_linear, aligner.hpp:3896-3911).  grads: K blocks of na*na, then K
   rows of go_a, ge_a, go_b, ge_b. */
typedef struct {
    int gap_model, mode, n_alpha, K;
    const double *matrices; /* K * na * na */
    const double *gaps;     /* K * 4 */
} nwo_tparams;

#define NWO_DEFINE_TRACKED(T, SUFFIX, NINF)                                                    \
    /* per-row penalties of one pair, setup_tracks_ (aligner.hpp:1919-2008) */                 \
    static void track_rows_##SUFFIX(const nwo_tparams *p, const uint16_t *trk, const uint16_t *gtrk, \
                                    int m, T *goa, T *gea, T *gob, T *geb, T *bx0) {           \
        for (int i = 0; i <= m; ++i) {                                                         \
            goa[i] = (T)p->gaps[4 * gtrk[i] + 0];                                              \
            gea[i] = (T)p->gaps[4 * gtrk[i] + 1];                                              \
        }                                                                                      \
        gob[0] = geb[0] = bx0[0] = (T)0;                                                       \
        for (int i = 1; i <= m; ++i) {                                                         \
            gob[i] = (T)p->gaps[4 * trk[i - 1] + 2];                                           \
            geb[i] = (T)p->gaps[4 * trk[i - 1] + 3];                                           \
        }                                                                                      \
        T csum = (T)0;                                                                         \
        int s = 1;                                                                             \
        for (int i = 1; i <= m; ++i) {                                                         \
            if (i > 1 && memcmp(&geb[i], &geb[i - 1], sizeof(T)) != 0) {                      \
                csum = (s == 1) ? (T)(i - 1) * geb[i - 1] : csum + (T)(i - s) * geb[i - 1];  \
                s = i;                                                                         \
            }                                                                                  \
            const T ge = geb[i];                                                               \
            if (p->gap_model == 0) bx0[i] = (s == 1) ? -(T)i * ge : -(csum + (T)(i - s + 1) * ge); \
            else {                                                                             \
                const T go = gob[1];                                                           \
                bx0[i] = (s == 1) ? -(go + (T)i * ge) : -(go + (csum + (T)(i - s + 1) * ge)); \
            }                                                                                  \
        }                                                                                      \
    }                                                                                          \
    static T affine_fill_tr_##SUFFIX(const nwo_tparams *p, const T *blk, const uint8_t *a,    \
                                     const uint16_t *trk, const T *goa, const T *gea, const T *gob, \
                                     const T *geb, const T *bx0, int m, const uint8_t *b, int n, \
                                     unsigned char *D, T *rows, int *bi, int *bj, int *bst) {  \
        const int AM = p->mode, na = p->n_alpha;                                               \
        const size_t st = (size_t)n + 1;                                                       \
        T *pM = rows, *pX = rows + st, *pY = rows + 2 * st;                                    \
        T *cM = rows + 3 * st, *cX = rows + 4 * st, *cY = rows + 5 * st;                       \
        const int free_row0 = (AM == 1 || AM == 2);                                            \
        for (int j = 0; j <= n; ++j) {                                                         \
            if (free_row0) { pM[j] = 0; pX[j] = NINF; pY[j] = NINF; }                          \
            else {                                                                             \
                pM[j] = (j == 0) ? (T)0 : NINF;                                                \
                pX[j] = NINF;                                                                  \
                pY[j] = (j == 0) ? NINF : -(goa[0] + (T)j * gea[0]);                           \
            }                                                                                  \
            D[j] = (j <= 1 || free_row0) ? 3 : (3 | 2 << 4);                                   \
        }                                                                                      \
        *bi = 0; *bj = 0; *bst = ST_M;                                                         \
        T best_local = 0;                                                                      \
        for (int i = 1; i <= m; ++i) {                                                         \
            const size_t rb = (size_t)i * st;                                                  \
            const T go_a = goa[i], ge_a = gea[i], go_b = gob[i], ge_b = geb[i];               \
            const T *srow = blk + ((size_t)trk[i - 1] * na + a[i - 1]) * na;                   \
            if (AM == 1) { cM[0] = 0; cX[0] = NINF; }                                          \
            else { cM[0] = NINF; cX[0] = bx0[i]; }                                             \
            cY[0] = NINF;                                                                      \
            D[rb] = (i <= 1) ? 3 : (3 | 1 << 2);                                               \
            T rowmax = 0;                                                                      \
            for (int j = 1; j <= n; ++j) {                                                     \
                const T dgM = pM[j - 1], dgX = pX[j - 1], dgY = pY[j - 1];                     \
                T m_val = max3_##SUFFIX(dgM, dgX, dgY) + srow[b[j - 1]];                       \
                unsigned char km = pick_##SUFFIX(dgM, dgX, dgY);                               \
                const T ux = (pM[j] - go_b) - ge_b, vx = pX[j] - ge_b, wx = (pY[j] - go_b) - ge_b; \
                const T x_val = max3_##SUFFIX(ux, vx, wx);                                     \
                const unsigned char kx = pick_##SUFFIX(ux, vx, wx);                            \
                const T uy = (cM[j - 1] - go_a) - ge_a, vy = (cX[j - 1] - go_a) - ge_a,        \
                        wy = cY[j - 1] - ge_a;                                                 \
                const T y_val = max3_##SUFFIX(uy, vy, wy);                                     \
                const unsigned char ky = pick_##SUFFIX(uy, vy, wy);                            \
                if (AM == 1) {                                                                 \
                    m_val = max2_##SUFFIX(m_val, (T)0);                                        \
                    rowmax = max2_##SUFFIX(rowmax, max3_##SUFFIX(m_val, x_val, y_val));        \
                    if (m_val <= (T)0) km = 3;                                                 \
                }                                                                              \
                cM[j] = m_val; cX[j] = x_val; cY[j] = y_val;                                   \
                D[rb + j] = (unsigned char)(km | kx << 2 | ky << 4);                          \
            }                                                                                  \
            if (AM == 1 && rowmax > best_local) {                                              \
                int j = 1;                                                                     \
                while (!(max3_##SUFFIX(cM[j], cX[j], cY[j]) == rowmax)) ++j;                   \
                best_local = rowmax; *bi = i; *bj = j;                                         \
                if (cM[j] >= cX[j] && cM[j] >= cY[j]) *bst = ST_M;                             \
                else if (cX[j] >= cY[j]) *bst = ST_X;                                          \
                else *bst = ST_Y;                                                              \
            }                                                                                  \
            T *t;                                                                              \
            t = pM; pM = cM; cM = t; t = pX; pX = cX; cX = t; t = pY; pY = cY; cY = t;         \
        }                                                                                      \
        if (AM == 0) {                                                                         \
            const T vm = pM[n], vx = pX[n], vy = pY[n];                                        \
            *bi = m; *bj = n;                                                                  \
            if (vm >= vx && vm >= vy) *bst = ST_M;                                             \
            else if (vx >= vy) *bst = ST_X;                                                    \
            else *bst = ST_Y;                                                                  \
            return max3_##SUFFIX(vm, vx, vy);                                                  \
        }                                                                                      \
        if (AM == 2) {                                                                         \
            int bj2 = 0;                                                                       \
            T best = max3_##SUFFIX(pM[0], pX[0], pY[0]);                                       \
            for (int j = 1; j <= n; ++j) {                                                     \
                const T v = max3_##SUFFIX(pM[j], pX[j], pY[j]);                                \
                if (v > best) { best = v; bj2 = j; }                                           \
            }                                                                                  \
            *bi = m; *bj = bj2;                                                                \
            const T vm = pM[bj2], vx = pX[bj2], vy = pY[bj2];                                  \
            if (vm >= vx && vm >= vy) *bst = ST_M;                                             \
            else if (vx >= vy) *bst = ST_X;                                                    \
            else *bst = ST_Y;                                                                  \
            return best;                                                                       \
        }                                                                                      \
        return best_local;                                                                     \
    }                                                                                          \
    static T linear_fill_tr_##SUFFIX(const nwo_tparams *p, const T *blk, const uint8_t *a,    \
                                     const uint16_t *trk, const T *gea, const T *geb, const T *bx0, \
                                     int m, const uint8_t *b, int n, unsigned char *D, T *rows, \
                                     int *bi, int *bj) {                                       \
        const int AM = p->mode, na = p->n_alpha, local = (AM == 1), free0 = (AM == 1 || AM == 2); \
        const size_t st = (size_t)n + 1;                                                       \
        T *prev = rows, *cur = rows + st;                                                      \
        prev[0] = 0; D[0] = 3;                                                                 \
        for (int j = 1; j <= n; ++j) {                                                         \
            prev[j] = free0 ? (T)0 : -(T)j * gea[0];                                           \
            D[j] = free0 ? 3 : 2;                                                              \
        }                                                                                      \
        int bbi = 0, bbj = 0;                                                                  \
        T best_local = 0;                                                                      \
        for (int i = 1; i <= m; ++i) {                                                         \
            unsigned char *d = D + (size_t)i * st;                                             \
            const T *prow = blk + ((size_t)trk[i - 1] * na + a[i - 1]) * na;                   \
            const T ge_a = gea[i], ge_b = geb[i];                                              \
            cur[0] = local ? (T)0 : bx0[i];                                                    \
            d[0] = local ? 3 : 1;                                                              \
            T left = cur[0];                                                                   \
            for (int j = 1; j <= n; ++j) {                                                     \
                const T d0 = prev[j - 1] + prow[b[j - 1]];                                     \
                const T u = prev[j] - ge_b;                                                    \
                const T t = max2_##SUFFIX(d0, u);                                              \
                unsigned char c = (unsigned char)(d0 < u);                                     \
                const T l = left - ge_a;                                                       \
                T v = max2_##SUFFIX(t, l);                                                     \
                if (t < l) c = 2;                                                              \
                if (local) {                                                                   \
                    v = max2_##SUFFIX(v, (T)0);                                                \
                    if (v <= (T)0) c = (unsigned char)(c | 3);                                 \
                    if (v > best_local) { best_local = v; bbi = i; bbj = j; }                  \
                }                                                                              \
                cur[j] = v; d[j] = c;                                                          \
                left = v;                                                                      \
            }                                                                                  \
            T *tt = prev; prev = cur; cur = tt;                                                \
        }                                                                                      \
        if (AM == 2) {                                                                         \
            int bj2 = 0;                                                                       \
            T best = prev[0];                                                                  \
            for (int j = 1; j <= n; ++j) if (prev[j] > best) { best = prev[j]; bj2 = j; }      \
            *bi = m; *bj = bj2;                                                                \
            return best;                                                                       \
        }                                                                                      \
        if (local) { *bi = bbi; *bj = bbj; return best_local; }                               \
        *bi = m; *bj = n;                                                                      \
        return prev[n];                                                                        \
    }

NWO_DEFINE_TRACKED(float, f32, -INFINITY)
NWO_DEFINE_TRACKED(double, f64, -INFINITY)

/* Counts of a tracked walk: matrix block trk[i-1] for a match; a gap in B (X, into row i)
   the gap fields of set trk[i-1]; a gap in A (Y, within row i) those of set gtrk[i]. */
static void emit_tracked(double *g, int na, int K, const uint8_t *a, const uint8_t *b, const uint16_t *trk,
                         const uint16_t *gtrk, int st, int i, int j, int open) {
    const int nn = na * na;
    double *gap = g + (size_t)K * nn;
    if (st == ST_M) g[(size_t)trk[i - 1] * nn + a[i - 1] * na + b[j - 1]] += 1.0;
    else if (st == ST_X) { gap[4 * trk[i - 1] + 3] -= 1.0; if (open) gap[4 * trk[i - 1] + 2] -= 1.0; }
    else { gap[4 * gtrk[i] + 1] -= 1.0; if (open) gap[4 * gtrk[i] + 0] -= 1.0; }
}

/* walk_affine_impl<Ptr, Tr=true> (aligner.hpp:2986-3056): the untracked walk's moves,
   each count in its row's set. */
static void walk_affine_tr(int mode, const unsigned char *D, size_t st, const uint8_t *a, const uint8_t *b,
                           int na, int K, const uint16_t *trk, const uint16_t *gtrk, int i, int j, int tbl,
                           double *g) {
    for (;;) {
        if (i == 0 && j == 0) break;
        unsigned char c = 0;
        if (mode == 1) {
            c = D[(size_t)i * st + j];
            if (tbl == ST_M && (c & 3) == 3) break;
        } else {
            if (i == 0) {
                if (mode == 2) break;
                for (; j > 0; --j) emit_tracked(g, na, K, a, b, trk, gtrk, ST_Y, 0, j, j == 1);
                break;
            }
            if (j == 0) {
                for (; i > 0; --i) emit_tracked(g, na, K, a, b, trk, gtrk, ST_X, i, 0, i == 1);
                break;
            }
            c = D[(size_t)i * st + j];
        }
        if (i == 0) tbl = ST_Y;
        else if (j == 0) tbl = ST_X;
        int prev;
        if (tbl == ST_M) {
            const int k = c & 3;
            prev = k == 1 ? ST_X : (k == 2 ? ST_Y : ST_M);
            emit_tracked(g, na, K, a, b, trk, gtrk, ST_M, i, j, 0);
            --i; --j;
        } else if (tbl == ST_X) {
            const int k = (c >> 2) & 3;
            prev = k == 1 ? ST_X : (k == 2 ? ST_Y : ST_M);
            emit_tracked(g, na, K, a, b, trk, gtrk, ST_X, i, j, prev != ST_X);
            --i;
        } else {
            const int k = c >> 4;
            prev = k == 1 ? ST_X : (k == 2 ? ST_Y : ST_M);
            emit_tracked(g, na, K, a, b, trk, gtrk, ST_Y, i, j, prev != ST_Y);
            --j;
        }
        tbl = prev;
    }
}

/* walk_linear with hard_grad_linear's tracked emit (aligner.hpp:2826-2855, 3896-3911). */
static void walk_linear_tr(int mode, const unsigned char *D, size_t st, const uint8_t *a, const uint8_t *b,
                           int na, int K, const uint16_t *trk, const uint16_t *gtrk, int i, int j, double *g) {
    const int nn = na * na;
    double *gap = g + (size_t)K * nn;
    for (;;) {
        if (i == 0 && j == 0) break;
        if (mode == 2 && i == 0) break;
        const int k = D[(size_t)i * st + j];
        if (k == 3) break;
        if (k == 0) { g[(size_t)trk[i - 1] * nn + a[i - 1] * na + b[j - 1]] += 1.0; --i; --j; }
        else if (k == 1) { gap[4 * trk[i - 1] + 3] -= 1.0; --i; }
        else { gap[4 * gtrk[i] + 1] -= 1.0; --j; }
    }
}

/* Tracked batch entry.  Pair q's track is trk[oa[q] .. oa[q+1]) (one set per residue of
   A) and its gap track gtrk[oa[q] + q .. oa[q+1] + q] (one per DP row 0..m).  grads: N
   rows of K*(na*na + 4) doubles.  Returns 0, -1 on allocation failure, -2 on a set index
   outside [0, K). */
int nwo_hard_tracked(const nwo_tparams *p, int precision, int64_t N, const uint8_t *ca, const int64_t *oa,
                     const uint8_t *cb, const int64_t *ob, const uint16_t *trk, const uint16_t *gtrk,
                     int with_grad, double *scores, double *grads) {
    const int na = p->n_alpha, K = p->K, nn = na * na;
    const size_t gs = (size_t)K * (nn + 4);
    size_t maxcells = 1, maxrow = 1, maxm = 1;
    for (int64_t q = 0; q < N; ++q) {
        const size_t m = (size_t)(oa[q + 1] - oa[q]), n = (size_t)(ob[q + 1] - ob[q]);
        if ((m + 1) * (n + 1) > maxcells) maxcells = (m + 1) * (n + 1);
        if (n + 1 > maxrow) maxrow = n + 1;
        if (m + 1 > maxm) maxm = m + 1;
        for (size_t i = 0; i < m; ++i) if (trk[oa[q] + i] >= K) return -2;
        for (size_t i = 0; i <= m; ++i) if (gtrk[oa[q] + q + i] >= K) return -2;
    }
    unsigned char *D = (unsigned char *)malloc(maxcells);
    double *rows = (double *)malloc(6 * maxrow * sizeof(double));
    double *rowp = (double *)malloc(5 * maxm * sizeof(double));
    float *blk32 = (float *)malloc((size_t)K * nn * sizeof(float));
    double *blk64 = (double *)malloc((size_t)K * nn * sizeof(double));
    if (!D || !rows || !rowp || !blk32 || !blk64) {
        free(D); free(rows); free(rowp); free(blk32); free(blk64);
        return -1;
    }
    for (size_t k = 0; k < (size_t)K * nn; ++k) { blk32[k] = (float)p->matrices[k]; blk64[k] = p->matrices[k]; }
    for (int64_t q = 0; q < N; ++q) {
        const uint8_t *a = ca + oa[q], *b = cb + ob[q];
        const uint16_t *tq = trk + oa[q], *gq = gtrk + oa[q] + q;
        const int m = (int)(oa[q + 1] - oa[q]), n = (int)(ob[q + 1] - ob[q]);
        int bi, bj, bst = ST_M;
        double sc;
        if (precision) {
            double *goa = rowp, *gea = rowp + maxm, *gob = rowp + 2 * maxm, *geb = rowp + 3 * maxm, *bx0 = rowp + 4 * maxm;
            track_rows_f64(p, tq, gq, m, goa, gea, gob, geb, bx0);
            sc = p->gap_model == 1
                     ? affine_fill_tr_f64(p, blk64, a, tq, goa, gea, gob, geb, bx0, m, b, n, D, rows, &bi, &bj, &bst)
                     : linear_fill_tr_f64(p, blk64, a, tq, gea, geb, bx0, m, b, n, D, rows, &bi, &bj);
        } else {
            float *r32 = (float *)rowp;
            float *goa = r32, *gea = r32 + maxm, *gob = r32 + 2 * maxm, *geb = r32 + 3 * maxm, *bx0 = r32 + 4 * maxm;
            track_rows_f32(p, tq, gq, m, goa, gea, gob, geb, bx0);
            sc = p->gap_model == 1
                     ? (double)affine_fill_tr_f32(p, blk32, a, tq, goa, gea, gob, geb, bx0, m, b, n, D, (float *)rows,
                                                  &bi, &bj, &bst)
                     : (double)linear_fill_tr_f32(p, blk32, a, tq, gea, geb, bx0, m, b, n, D, (float *)rows, &bi, &bj);
        }
        scores[q] = sc;
        if (with_grad) {
            double *g = grads + (size_t)q * gs;
            memset(g, 0, gs * sizeof(double));
            if (p->gap_model == 1) walk_affine_tr(p->mode, D, (size_t)n + 1, a, b, na, K, tq, gq, bi, bj, bst, g);
            else walk_linear_tr(p->mode, D, (size_t)n + 1, a, b, na, K, tq, gq, bi, bj, g);
        }
    }
    free(D); free(rows); free(rowp); free(blk32); free(blk64);
    return 0;
}

/* ── tracked soft: the log-space pass of aligner.hpp:6231-6524 with each row's penalties
   (drow_go_a / drow_ge_a of gap set gtrk[i], drow_go_b / drow_ge_b and the matrix of set
   trk[i-1], aligner.hpp:1903-1916, as the tracked scaled job takes them, aligner.hpp:
   6845-6881), the record kept per row (soft_record_log_, aligner.hpp:6647-6700) and added
   into the sets of its rows (add_soft_record_, aligner.hpp:6628-6645): matches of row i into
   set trk[i-1]'s block, row i's gap-in-A counts into set gtrk[i], its gap-in-B counts into
   set trk[i-1].  Column 0 of Global / Glocal is the vertical chain of row penalties.
   grads: K blocks of na*na, then K rows of go_a, ge_a, go_b, ge_b. */
static void soft_affine_pair_tr(const nwo_tparams *p, double Tm, const uint8_t *a, const uint16_t *trk,
                                const uint16_t *gtrk, int m, const uint8_t *b, int n, const double *goa,
                                const double *gea, const double *gob, const double *geb, double *buf, double *score,
                                double *g) {
    const int AM = p->mode, na = p->n_alpha, nn = na * na, K = p->K;
    const double it = 1.0 / Tm;
    const size_t st = (size_t)n + 1, sz = (size_t)(m + 1) * st;
    double *FM = buf, *FX = buf + sz, *FY = buf + 2 * sz, *BM = buf + 3 * sz, *BX = buf + 4 * sz,
           *BY = buf + 5 * sz;
#define AT(t, i, j) (t)[(size_t)(i) * st + (size_t)(j)]
#define SUB(i, j) (p->matrices[((size_t)trk[(i) - 1] * na + a[(i) - 1]) * na + b[(j) - 1]] * it)
#define GOA(i) (goa[i] * it)
#define GEA(i) (gea[i] * it)
#define GOB(i) (gob[i] * it)
#define GEB(i) (geb[i] * it)
    for (size_t k = 0; k < sz; ++k) { FM[k] = FX[k] = FY[k] = -INFINITY; }
    if (AM != 1) {
        AT(FM, 0, 0) = 0.0;
        for (int i = 1; i <= m; ++i) AT(FX, i, 0) = (i == 1) ? -(GOB(1) + GEB(1)) : AT(FX, i - 1, 0) - GEB(i);
        if (AM == 0) for (int j = 1; j <= n; ++j) AT(FY, 0, j) = -(GOA(0) + (double)j * GEA(0));
        if (AM == 2) for (int j = 0; j <= n; ++j) AT(FM, 0, j) = 0.0;
    } else {
        for (int i = 0; i <= m; ++i) AT(FM, i, 0) = 0.0;
        for (int j = 0; j <= n; ++j) AT(FM, 0, j) = 0.0;
    }
    for (int i = 1; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const double diag = lse3(AT(FM, i - 1, j - 1), AT(FX, i - 1, j - 1), AT(FY, i - 1, j - 1));
            double mv = diag + SUB(i, j);
            const double xv = lse3(AT(FM, i - 1, j) - GOB(i) - GEB(i), AT(FX, i - 1, j) - GEB(i),
                                   AT(FY, i - 1, j) - GOB(i) - GEB(i));
            const double yv = lse3(AT(FM, i, j - 1) - GOA(i) - GEA(i), AT(FX, i, j - 1) - GOA(i) - GEA(i),
                                   AT(FY, i, j - 1) - GEA(i));
            if (AM == 1) mv = lse2(mv, 0.0);
            AT(FM, i, j) = mv; AT(FX, i, j) = xv; AT(FY, i, j) = yv;
        }
    double lz;
    if (AM == 0) lz = lse3(AT(FM, m, n), AT(FX, m, n), AT(FY, m, n));
    else if (AM == 2) {
        lz = -INFINITY;
        for (int j = 0; j <= n; ++j) lz = lse2(lz, lse3(AT(FM, m, j), AT(FX, m, j), AT(FY, m, j)));
    } else {
        lz = -INFINITY;
        for (int i = 0; i <= m; ++i)
            for (int j = 0; j <= n; ++j) lz = lse2(lz, lse3(AT(FM, i, j), AT(FX, i, j), AT(FY, i, j)));
    }
    const double e0 = (AM == 1) ? 0.0 : -INFINITY;
    for (size_t k = 0; k < sz; ++k) { BM[k] = BX[k] = BY[k] = e0; }
    if (AM == 0) { AT(BM, m, n) = AT(BX, m, n) = AT(BY, m, n) = 0.0; }
    if (AM == 2) for (int j = 0; j <= n; ++j) { AT(BM, m, j) = AT(BX, m, j) = AT(BY, m, j) = 0.0; }
    for (int i = m; i >= 0; --i)
        for (int j = n; j >= 0; --j) {
            const double bm = AT(BM, i, j), bx = AT(BX, i, j), by = AT(BY, i, j);
            if (bm != -INFINITY && i > 0 && j > 0) {
                const double c = bm + SUB(i, j);
                AT(BM, i - 1, j - 1) = lse2(AT(BM, i - 1, j - 1), c);
                AT(BX, i - 1, j - 1) = lse2(AT(BX, i - 1, j - 1), c);
                AT(BY, i - 1, j - 1) = lse2(AT(BY, i - 1, j - 1), c);
            }
            if (bx != -INFINITY && i > 0) {
                AT(BM, i - 1, j) = lse2(AT(BM, i - 1, j), bx - GOB(i) - GEB(i));
                AT(BX, i - 1, j) = lse2(AT(BX, i - 1, j), bx - GEB(i));
                AT(BY, i - 1, j) = lse2(AT(BY, i - 1, j), bx - GOB(i) - GEB(i));
            }
            if (by != -INFINITY && j > 0) {
                AT(BM, i, j - 1) = lse2(AT(BM, i, j - 1), by - GOA(i) - GEA(i));
                AT(BX, i, j - 1) = lse2(AT(BX, i, j - 1), by - GOA(i) - GEA(i));
                AT(BY, i, j - 1) = lse2(AT(BY, i, j - 1), by - GEA(i));
            }
        }
    const int gmin_j = (AM == 1) ? 1 : 0, gmin_i = (AM == 0) ? 0 : 1;
    double *gg = g + (size_t)K * nn;
    memset(g, 0, (size_t)K * (nn + 4) * sizeof(double));
    for (int i = 1; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const double bm = AT(BM, i, j);
            if (bm == -INFINITY) continue;
            const double pred = lse3(AT(FM, i - 1, j - 1), AT(FX, i - 1, j - 1), AT(FY, i - 1, j - 1));
            g[((size_t)trk[i - 1] * na + a[i - 1]) * na + b[j - 1]] += exp(pred + SUB(i, j) + bm - lz);
        }
    for (int i = 1; i <= m; ++i)
        for (int j = 0; j <= n; ++j) {
            const double fx = AT(FX, i, j), bx = AT(BX, i, j);
            if (fx == -INFINITY || bx == -INFINITY) continue;
            gg[4 * trk[i - 1] + 3] -= exp(fx + bx - lz);
        }
    for (int i = 0; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const double fy = AT(FY, i, j), by = AT(BY, i, j);
            if (fy == -INFINITY || by == -INFINITY) continue;
            gg[4 * gtrk[i] + 1] -= exp(fy + by - lz);
        }
    for (int i = 1; i <= m; ++i)
        for (int j = gmin_j; j <= n; ++j) {
            const double bx = AT(BX, i, j);
            if (bx == -INFINITY) continue;
            const double lo = lse2(AT(FM, i - 1, j), AT(FY, i - 1, j)) - GOB(i) - GEB(i) + bx - lz;
            if (lo > -700) gg[4 * trk[i - 1] + 2] -= exp(lo);
        }
    for (int i = gmin_i; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const double by = AT(BY, i, j);
            if (by == -INFINITY) continue;
            const double lo = lse2(AT(FM, i, j - 1), AT(FX, i, j - 1)) - GOA(i) - GEA(i) + by - lz;
            if (lo > -700) gg[4 * gtrk[i] + 0] -= exp(lo);
        }
    *score = lz * Tm;
#undef AT
#undef SUB
}

static void soft_linear_pair_tr(const nwo_tparams *p, double Tm, const uint8_t *a, const uint16_t *trk,
                                const uint16_t *gtrk, int m, const uint8_t *b, int n, const double *gea,
                                const double *geb, double *buf, double *score, double *g) {
    const int AM = p->mode, na = p->n_alpha, nn = na * na, K = p->K;
    const double it = 1.0 / Tm;
    const size_t st = (size_t)n + 1, sz = (size_t)(m + 1) * st;
    double *F = buf, *B = buf + sz;
#define AT(t, i, j) (t)[(size_t)(i) * st + (size_t)(j)]
#define SUB(i, j) (p->matrices[((size_t)trk[(i) - 1] * na + a[(i) - 1]) * na + b[(j) - 1]] * it)
    for (size_t k = 0; k < sz; ++k) F[k] = -INFINITY;
    if (AM != 1) {
        AT(F, 0, 0) = 0.0;
        for (int i = 1; i <= m; ++i) AT(F, i, 0) = AT(F, i - 1, 0) - GEB(i);
        if (AM == 0) for (int j = 1; j <= n; ++j) AT(F, 0, j) = -(double)j * GEA(0);
        if (AM == 2) for (int j = 0; j <= n; ++j) AT(F, 0, j) = 0.0;
    } else {
        for (int i = 0; i <= m; ++i) AT(F, i, 0) = 0.0;
        for (int j = 0; j <= n; ++j) AT(F, 0, j) = 0.0;
    }
    for (int i = 1; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            double v = lse3(AT(F, i - 1, j - 1) + SUB(i, j), AT(F, i - 1, j) - GEB(i), AT(F, i, j - 1) - GEA(i));
            if (AM == 1) v = lse2(v, 0.0);
            AT(F, i, j) = v;
        }
    double lz;
    if (AM == 0) lz = AT(F, m, n);
    else if (AM == 2) { lz = -INFINITY; for (int j = 0; j <= n; ++j) lz = lse2(lz, AT(F, m, j)); }
    else {
        lz = -INFINITY;
        for (int i = 0; i <= m; ++i) for (int j = 0; j <= n; ++j) lz = lse2(lz, AT(F, i, j));
    }
    const double e0 = (AM == 1) ? 0.0 : -INFINITY;
    for (size_t k = 0; k < sz; ++k) B[k] = e0;
    if (AM == 0) AT(B, m, n) = 0.0;
    if (AM == 2) for (int j = 0; j <= n; ++j) AT(B, m, j) = 0.0;
    for (int i = m; i >= 0; --i)
        for (int j = n; j >= 0; --j) {
            const double bv = AT(B, i, j);
            if (bv == -INFINITY) continue;
            if (i > 0 && j > 0) AT(B, i - 1, j - 1) = lse2(AT(B, i - 1, j - 1), bv + SUB(i, j));
            if (i > 0) AT(B, i - 1, j) = lse2(AT(B, i - 1, j), bv - GEB(i));
            if (j > 0) AT(B, i, j - 1) = lse2(AT(B, i, j - 1), bv - GEA(i));
        }
    const int gmin_j = (AM == 1) ? 1 : 0, gmin_i = (AM == 0) ? 0 : 1;
    double *gg = g + (size_t)K * nn;
    memset(g, 0, (size_t)K * (nn + 4) * sizeof(double));
    for (int i = 1; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const double bv = AT(B, i, j);
            if (bv == -INFINITY) continue;
            g[((size_t)trk[i - 1] * na + a[i - 1]) * na + b[j - 1]] += exp(AT(F, i - 1, j - 1) + SUB(i, j) + bv - lz);
        }
    for (int i = 1; i <= m; ++i)
        for (int j = gmin_j; j <= n; ++j) {
            const double bv = AT(B, i, j), fv = AT(F, i - 1, j);
            if (bv == -INFINITY || fv == -INFINITY) continue;
            gg[4 * trk[i - 1] + 3] -= exp(fv - GEB(i) + bv - lz);
        }
    for (int i = gmin_i; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const double bv = AT(B, i, j), fv = AT(F, i, j - 1);
            if (bv == -INFINITY || fv == -INFINITY) continue;
            gg[4 * gtrk[i] + 1] -= exp(fv - GEA(i) + bv - lz);
        }
    *score = lz * Tm;
#undef AT
#undef SUB
#undef GOA
#undef GEA
#undef GOB
#undef GEB
}

/* Tracked batch soft entry (log space, double).  Returns 0, or -1 on allocation failure. */
int nwo_soft_log_tracked(const nwo_tparams *p, double temperature, int64_t N, const uint8_t *ca, const int64_t *oa,
                         const uint8_t *cb, const int64_t *ob, const uint16_t *trk, const uint16_t *gtrk,
                         double *scores, double *grads) {
    const int na = p->n_alpha, K = p->K;
    const size_t gs = (size_t)K * (na * na + 4);
    size_t maxcells = 1, maxm = 1;
    for (int64_t q = 0; q < N; ++q) {
        const size_t m = (size_t)(oa[q + 1] - oa[q]), n = (size_t)(ob[q + 1] - ob[q]);
        if ((m + 1) * (n + 1) > maxcells) maxcells = (m + 1) * (n + 1);
        if (m + 1 > maxm) maxm = m + 1;
    }
    double *buf = (double *)malloc(6 * maxcells * sizeof(double));
    double *rowp = (double *)malloc(5 * maxm * sizeof(double));
    if (!buf || !rowp) { free(buf); free(rowp); return -1; }
    for (int64_t q = 0; q < N; ++q) {
        const uint8_t *a = ca + oa[q], *b = cb + ob[q];
        const int m = (int)(oa[q + 1] - oa[q]), n = (int)(ob[q + 1] - ob[q]);
        const uint16_t *tq = trk + oa[q], *gq = gtrk + oa[q] + q;
        double *goa = rowp, *gea = rowp + maxm, *gob = rowp + 2 * maxm, *geb = rowp + 3 * maxm, *bx0 = rowp + 4 * maxm;
        track_rows_f64(p, tq, gq, m, goa, gea, gob, geb, bx0);
        double *g = grads + (size_t)q * gs;
        if (p->gap_model == 1) soft_affine_pair_tr(p, temperature, a, tq, gq, m, b, n, goa, gea, gob, geb, buf,
                                                   &scores[q], g);
        else soft_linear_pair_tr(p, temperature, a, tq, gq, m, b, n, gea, geb, buf, &scores[q], g);
    }
    free(buf);
    free(rowp);
    return 0;
}
