"""Exhaustive path enumeration for tiny pairs: an oracle that shares no code with any DP.

The model is the one the reference's DP implements (aligner.hpp:21-60 and the fills it
documents; its own enumerator is tests/Python/bruteforce.py):

* moves D (a[i], b[j]) -> +matrix[a_i][b_j]; X consumes A only (gap in B, the b params);
  Y consumes B only (gap in A, the a params);
* affine: every maximal run of k gap moves of one kind costs open + k * extend (an X run
  followed directly by a Y run pays both opens); linear: k * extend;
* global: (0,0) -> (m,n);
* local: start at any cell in the match state (an empty path scores 0), end at any cell;
  no X move along column 0 and no Y move along row 0 (the reference keeps those border
  cells fixed);
* glocal (A within B): start at any (0, j), end at any (m, j); no Y move along row 0.

Soft: Z = sum over every (start, path, end) of exp(score / T), the expected counts the
posterior-weighted counts of D moves per (a, b) letter pair and of gap moves / opens.
"""
from __future__ import annotations

import math

import numpy as np


def _paths(i, j, m, n, mode, first_row_free):
    """Yield move strings from (i, j); in local / glocal every prefix end is also a candidate."""
    yield ""
    if i < m and j < n:
        for rest in _paths(i + 1, j + 1, m, n, mode, first_row_free):
            yield "D" + rest
    if i < m and not (mode == "local" and j == 0):
        for rest in _paths(i + 1, j, m, n, mode, first_row_free):
            yield "X" + rest
    if j < n and not (i == 0 and mode in ("local", "glocal")):
        for rest in _paths(i, j + 1, m, n, mode, first_row_free):
            yield "Y" + rest


def _score_and_counts(path, i, j, a, b, matrix, gaps, affine):
    goa, gea, gob, geb = gaps
    na = matrix.shape[0]
    cnt = np.zeros(na * na + 4)
    s = 0.0
    prev = "M"
    for mv in path:
        if mv == "D":
            s += matrix[a[i], b[j]]
            cnt[a[i] * na + b[j]] += 1
            i += 1
            j += 1
            prev = "M"
        elif mv == "X":
            s -= geb
            cnt[na * na + 3] -= 1
            if affine and prev != "X":
                s -= gob
                cnt[na * na + 2] -= 1
            i += 1
            prev = "X"
        else:
            s -= gea
            cnt[na * na + 1] -= 1
            if affine and prev != "Y":
                s -= goa
                cnt[na * na + 0] -= 1
            j += 1
            prev = "Y"
    return s, cnt, i, j


def enumerate_alignments(a, b, matrix, gaps, *, gap_model="affine", mode="global"):
    """All (score, counts) of the model's alignments of a vs b (lists of codes)."""
    m, n = len(a), len(b)
    affine = gap_model == "affine"
    out = []
    if mode == "global":
        starts = [(0, 0)]
    elif mode == "glocal":
        starts = [(0, j) for j in range(n + 1)]
    else:
        starts = [(i, j) for i in range(m + 1) for j in range(n + 1)]
    for (i0, j0) in starts:
        for p in _paths(i0, j0, m, n, mode, True):
            s, c, i1, j1 = _score_and_counts(p, i0, j0, a, b, matrix, gaps, affine)
            if mode == "global" and (i1, j1) != (m, n):
                continue
            if mode == "glocal" and i1 != m:
                continue
            out.append((s, c))
    return out


def hard(a, b, matrix, gaps, **kw):
    """Best score, and the counts of the best alignment when it is unique (else None)."""
    al = enumerate_alignments(a, b, matrix, gaps, **kw)
    best = max(s for s, _ in al)
    winners = [c for s, c in al if s == best]
    uniq = winners[0] if all(np.array_equal(winners[0], w) for w in winners) else None
    return best, uniq


def soft(a, b, matrix, gaps, temperature=1.0, **kw):
    """T * log Z and the expected counts."""
    al = enumerate_alignments(a, b, matrix, gaps, **kw)
    sc = np.array([s for s, _ in al]) / temperature
    mx = sc.max()
    w = np.exp(sc - mx)
    z = w.sum()
    logz = mx + math.log(z)
    counts = sum(wi * c for wi, (_, c) in zip(w, al)) / z
    return temperature * logz, counts
