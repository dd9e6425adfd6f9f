"""Brute-force banded soft alignment (test infrastructure): the forward-backward of one pair over
an explicit graph of (state, i, j) nodes, in log space and double, with the band rules of the
reference's GuideBanded soft pass.  Pure Python loops: small pairs only.

The rules (relative to the reference's src/nwgrad/cpp/nwgrad/):
* the band of row i, aligner.hpp:2644-2690: interior cells j in [jlo(i), jhi(i)], jlo = max(1,
  g[i] - bw), jhi = min(n, max(g[i] + bw, g[i+1] - 1 + bw)) (Global's row m reaches n); bw is
  clamped to n; column 0 carries the X border for rows 1 .. border_rows (g[i] <= bw), row 0 the Y
  border up to border_cols = jhi(0) (Global);
* the forward over those cells with the full pass's borders (aligner.hpp:6365-6400, log space;
  soft_kernel_impl.inl:643-708, scaled): Local's M = lse(diagonal + s, 0), starts at every
  column-0 and row-0 cell; Glocal's row 0 all starts;
* Z: Global the (m, n) cell; Glocal row m over [jlo0(m), jhi0(m)]; Local every row over
  [jlo0(i), jhi0(i)] (jlo0 = max(0, g[i] - bw); soft_kernel_impl.inl:289, 487, 785);
* the backward holds every node outside its row's [jlo0(i), jhi0(i)] at zero
  (soft_kernel_impl.inl:826-828), so a border move outside it (a Global row-0 gap left of
  jlo0(0)) is not counted, as in the reference.
Counts are the gradient of T log Z (T = 1): expected matches per letter pair, and minus the
expected gap opens and extends (go_a, ge_a, go_b, ge_b: gaps in A are moves along B).
"""
from math import exp, inf, log

import numpy as np


def _lse(xs):
    m = max(xs)
    if m == -inf:
        return -inf
    return m + log(sum(exp(x - m) for x in xs))


def band_rules(g, m, n, bw, mode):
    """(jlo, jhi, jlo0) as functions of the row, border_rows, border_cols."""
    bw = min(bw, n)

    def jlo(i):
        return max(1, int(g[i]) - bw)

    def jhi(i):
        if mode == "global" and i == m:
            return n
        h = int(g[i]) + bw
        if i < m:
            h = max(h, int(g[i + 1]) - 1 + bw)
        return min(n, h)

    def jlo0(i):
        return max(0, int(g[i]) - bw)

    bi = 0
    while bi < m and int(g[bi + 1]) <= bw:
        bi += 1
    return jlo, jhi, jlo0, bi, jhi(0)


def banded_soft(a, b, matrix, gaps, mode, guide, bw, gap_model="affine"):
    """log Z and the count row (n_alpha^2 + 4) of pair (a, b) under (matrix, gaps), banded
    around `guide` at half-width bw; bw = None runs the full pass."""
    go_a, ge_a, go_b, ge_b = (float(x) for x in gaps)
    if gap_model == "linear":
        go_a = go_b = 0.0
    a = [int(x) for x in a]
    b = [int(x) for x in b]
    m, n = len(a), len(b)
    if bw is None:
        jlo, jhi, jlo0, bi, bj = (lambda i: 1), (lambda i: n), (lambda i: 0), m, n
    else:
        jlo, jhi, jlo0, bi, bj = band_rules(guide, m, n, bw, mode)
    L, GL = mode == "local", mode == "glocal"
    init = {}     # node -> log weight of a path that starts there
    into = {}     # node -> [(source node, log weight, feature)]

    def edge(v, u, w, f):
        into.setdefault(v, []).append((u, w, f))

    if L:
        for i in range(m + 1):
            init[(0, i, 0)] = 0.0
        for j in range(n + 1):
            init[(0, 0, j)] = 0.0
    else:
        init[(0, 0, 0)] = 0.0
        if GL:
            for j in range(n + 1):
                init[(0, 0, j)] = 0.0
        for i in range(1, bi + 1):   # column 0: moves in B (X) down the border rows
            edge((1, i, 0), (0, i - 1, 0), -(go_b + ge_b), ("X", True))
            edge((1, i, 0), (2, i - 1, 0), -(go_b + ge_b), ("X", True))
            edge((1, i, 0), (1, i - 1, 0), -ge_b, ("X", False))
        if not GL:
            for j in range(1, bj + 1):   # row 0: moves in A (Y) along the border columns
                edge((2, 0, j), (0, 0, j - 1), -(go_a + ge_a), ("Y", True))
                edge((2, 0, j), (1, 0, j - 1), -(go_a + ge_a), ("Y", True))
                edge((2, 0, j), (2, 0, j - 1), -ge_a, ("Y", False))
    for i in range(1, m + 1):
        for j in range(jlo(i), jhi(i) + 1):
            s = float(matrix[a[i - 1], b[j - 1]])
            for st in range(3):
                edge((0, i, j), (st, i - 1, j - 1), s, ("M", a[i - 1], b[j - 1]))
            if L:
                init[(0, i, j)] = 0.0
            edge((1, i, j), (0, i - 1, j), -(go_b + ge_b), ("X", True))
            edge((1, i, j), (2, i - 1, j), -(go_b + ge_b), ("X", True))
            edge((1, i, j), (1, i - 1, j), -ge_b, ("X", False))
            edge((2, i, j), (0, i, j - 1), -(go_a + ge_a), ("Y", True))
            edge((2, i, j), (1, i, j - 1), -(go_a + ge_a), ("Y", True))
            edge((2, i, j), (2, i, j - 1), -ge_a, ("Y", False))
    order = sorted(set(into) | set(init), key=lambda v: (v[1], v[2], v[0] == 2))
    F = {}
    for v in order:
        terms = [init[v]] if v in init else []
        terms += [F[u] + w for (u, w, _) in into.get(v, []) if F.get(u, -inf) > -inf]
        F[v] = _lse(terms) if terms else -inf

    def span_hi(i):
        return bj if i == 0 else jhi(i)

    ends = []
    if L:
        ends = [(st, i, j) for i in range(m + 1) for j in range(jlo0(i), span_hi(i) + 1) for st in range(3)]
    elif GL:
        ends = [(st, m, j) for j in range(jlo0(m), jhi(m) + 1) for st in range(3)]
    else:
        ends = [(st, m, n) for st in range(3)]
    ends = [v for v in ends if v in F]
    logz = _lse([F[v] for v in ends])
    out = {}
    for v, lst in into.items():
        for (u, w, f) in lst:
            out.setdefault(u, []).append((v, w))
    B = {v: -inf for v in F}
    for v in ends:
        B[v] = 0.0
    for u in reversed(order):
        if not (jlo0(u[1]) <= u[2] <= span_hi(u[1])):
            B[u] = -inf   # outside the backward's span
            continue
        terms = ([B[u]] if B[u] > -inf else []) + [w + B[v] for (v, w) in out.get(u, [])]
        B[u] = _lse(terms) if terms else -inf
    na = matrix.shape[0]
    cnt = np.zeros(na * na + 4)
    for v, lst in into.items():
        if B.get(v, -inf) == -inf:
            continue
        for (u, w, f) in lst:
            if F.get(u, -inf) == -inf:
                continue
            p = exp(F[u] + w + B[v] - logz)
            if f[0] == "M":
                cnt[f[1] * na + f[2]] += p
            elif f[0] == "X":   # a gap in B: go_b, ge_b
                cnt[na * na + 3] -= p
                if f[1]:
                    cnt[na * na + 2] -= p
            else:
                cnt[na * na + 1] -= p
                if f[1]:
                    cnt[na * na + 0] -= p
    if gap_model == "linear":
        cnt[na * na + 0] = cnt[na * na + 2] = 0.0
    return logz, cnt
