"""TEST INFRASTRUCTURE ONLY: the CPU oracle of the alignment hot path.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import this
package, and only as the checker.  The product (``nwgrad_amd``) never imports it.

Two oracles live here:

* ``liboracle.so`` (built from ``nwg_oracle.c`` and ``nwg_oracle_hb.c``): a plain-C restatement
  of the reference's hard fill / walk / counts, score-only fill, log-space soft
  forward-backward, blocked gradient sum and the Hirschberg (linear-memory) traceback of
  Global affine pairs.  Each function cites the reference lines it follows.
* ``_ref/libnwgref.so`` (built by ``Makefile.ref`` from the reference's own sources, when
  they are present): the reference engine itself behind a small C ABI (``ref_harness.cpp``).

Pairs are passed as encoded ``uint8`` codes with ``int64`` offsets (``codes[off[i]:off[i+1]]``
is pair i's sequence), the layout the product uses too.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libnwgref.so")

GAP_MODELS = {"linear": 0, "affine": 1}
MODES = {"global": 0, "local": 1, "glocal": 2}


def build(force: bool = False) -> str:
    """Compile the C restatement (seconds).  -ffp-contract=off keeps every hard-path
    operation a separate rounding, as the reference's build does (CMakeLists.txt:75)."""
    srcs = [os.path.join(HERE, "nwg_oracle.c"), os.path.join(HERE, "nwg_oracle_hb.c")]
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < max(map(os.path.getmtime, srcs)):
        subprocess.check_call(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fPIC", "-shared",
                               "-o", LIB_PATH, *srcs, "-lm"])
    return LIB_PATH


def build_ref() -> str | None:
    """Build oracle/_ref from the reference's sources when they are present (this
    container only; the GPU box receives the built library)."""
    if not os.path.isdir("/root/reference/src/nwgrad/cpp/nwgrad"):
        return REF_PATH if os.path.exists(REF_PATH) else None
    root = os.path.dirname(HERE)
    subprocess.check_call(["make", "-s", "-f", "oracle/Makefile.ref"], cwd=root)
    return REF_PATH


class _Params(ctypes.Structure):
    _fields_ = [("gap_model", ctypes.c_int), ("mode", ctypes.c_int), ("n_alpha", ctypes.c_int),
                ("matrix", ctypes.POINTER(ctypes.c_double)),
                ("go_a", ctypes.c_double), ("ge_a", ctypes.c_double),
                ("go_b", ctypes.c_double), ("ge_b", ctypes.c_double)]


class _TParams(ctypes.Structure):
    _fields_ = [("gap_model", ctypes.c_int), ("mode", ctypes.c_int), ("n_alpha", ctypes.c_int), ("K", ctypes.c_int),
                ("matrices", ctypes.POINTER(ctypes.c_double)), ("gaps", ctypes.POINTER(ctypes.c_double))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.nwo_hard.argtypes = [P(_Params), ctypes.c_int, ctypes.c_int64, P(ctypes.c_uint8), P(ctypes.c_int64),
                               P(ctypes.c_uint8), P(ctypes.c_int64), ctypes.c_int, P(ctypes.c_double),
                               P(ctypes.c_double)]
        L.nwo_hard.restype = ctypes.c_int
        L.nwo_soft_log.argtypes = [P(_Params), ctypes.c_double, ctypes.c_int64, P(ctypes.c_uint8),
                                   P(ctypes.c_int64), P(ctypes.c_uint8), P(ctypes.c_int64),
                                   P(ctypes.c_double), P(ctypes.c_double)]
        L.nwo_soft_log.restype = ctypes.c_int
        L.nwo_weighted_sum.argtypes = [ctypes.c_int64, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double),
                                       P(ctypes.c_double)]
        L.nwo_weighted_sum.restype = None
        L.nwo_hard_tracked.argtypes = [P(_TParams), ctypes.c_int, ctypes.c_int64, P(ctypes.c_uint8),
                                       P(ctypes.c_int64), P(ctypes.c_uint8), P(ctypes.c_int64), P(ctypes.c_uint16),
                                       P(ctypes.c_uint16), ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double)]
        L.nwo_hard_tracked.restype = ctypes.c_int
        L.nwo_soft_log_tracked.argtypes = [P(_TParams), ctypes.c_double, ctypes.c_int64, P(ctypes.c_uint8),
                                           P(ctypes.c_int64), P(ctypes.c_uint8), P(ctypes.c_int64),
                                           P(ctypes.c_uint16), P(ctypes.c_uint16), P(ctypes.c_double),
                                           P(ctypes.c_double)]
        L.nwo_soft_log_tracked.restype = ctypes.c_int
        L.nwo_hirschberg.argtypes = [P(_Params), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                     P(ctypes.c_uint8), P(ctypes.c_int64), P(ctypes.c_uint8), P(ctypes.c_int64),
                                     P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_int32), P(ctypes.c_uint8),
                                     P(ctypes.c_int32), P(ctypes.c_uint8)]
        L.nwo_hirschberg.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _params(matrix, gaps, gap_model, mode):
    mat = np.ascontiguousarray(matrix, dtype=np.float64)
    na = mat.shape[0]
    p = _Params(GAP_MODELS[gap_model], MODES[mode], na, _ptr(mat, ctypes.c_double),
                float(gaps[0]), float(gaps[1]), float(gaps[2]), float(gaps[3]))
    return p, mat


def _pairs(codes_a, off_a, codes_b, off_b):
    ca = np.ascontiguousarray(codes_a, dtype=np.uint8)
    cb = np.ascontiguousarray(codes_b, dtype=np.uint8)
    oa = np.ascontiguousarray(off_a, dtype=np.int64)
    ob = np.ascontiguousarray(off_b, dtype=np.int64)
    if ca.size == 0:
        ca = np.zeros(1, np.uint8)
    if cb.size == 0:
        cb = np.zeros(1, np.uint8)
    return ca, oa, cb, ob


def hard(matrix, gaps, codes_a, off_a, codes_b, off_b, *, gap_model="affine", mode="global",
         precision="float32", with_grad=True):
    """Scores (N,) and per-pair gradients (N, na*na+4) of the hard path (or score-only)."""
    p, mat = _params(matrix, gaps, gap_model, mode)
    ca, oa, cb, ob = _pairs(codes_a, off_a, codes_b, off_b)
    N = len(oa) - 1
    na = mat.shape[0]
    scores = np.zeros(N, np.float64)
    grads = np.zeros((N, na * na + 4), np.float64) if with_grad else np.zeros((1, 1))
    rc = lib().nwo_hard(ctypes.byref(p), 1 if precision in ("float64", "double") else 0, N,
                        _ptr(ca, ctypes.c_uint8), _ptr(oa, ctypes.c_int64), _ptr(cb, ctypes.c_uint8),
                        _ptr(ob, ctypes.c_int64), 1 if with_grad else 0, _ptr(scores, ctypes.c_double),
                        _ptr(grads, ctypes.c_double))
    if rc:
        raise MemoryError("oracle: allocation failed")
    return scores, (grads if with_grad else None)


def soft_log(matrix, gaps, codes_a, off_a, codes_b, off_b, *, gap_model="affine", mode="global",
             temperature=1.0):
    """Log-space double forward-backward: scores T*logZ (N,) and expected counts."""
    p, mat = _params(matrix, gaps, gap_model, mode)
    ca, oa, cb, ob = _pairs(codes_a, off_a, codes_b, off_b)
    N = len(oa) - 1
    na = mat.shape[0]
    scores = np.zeros(N, np.float64)
    grads = np.zeros((N, na * na + 4), np.float64)
    rc = lib().nwo_soft_log(ctypes.byref(p), float(temperature), N, _ptr(ca, ctypes.c_uint8),
                            _ptr(oa, ctypes.c_int64), _ptr(cb, ctypes.c_uint8), _ptr(ob, ctypes.c_int64),
                            _ptr(scores, ctypes.c_double), _ptr(grads, ctypes.c_double))
    if rc:
        raise MemoryError("oracle: allocation failed")
    return scores, grads


def weighted_sum(grads, weights=None):
    """The reference's blocked sum (4096-pair blocks) of per-pair gradient rows."""
    g = np.ascontiguousarray(grads, dtype=np.float64)
    N, gs = g.shape
    out = np.zeros(gs, np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    lib().nwo_weighted_sum(N, gs, _ptr(g, ctypes.c_double),
                           None if w is None else _ptr(w, ctypes.c_double), _ptr(out, ctypes.c_double))
    return out


def have_ref() -> bool:
    return os.path.exists(REF_PATH)


def default_gap_tracks(tracks, off_a):
    """The reference's default gap track of each pair (setup_tracks_, aligner.hpp:1963-1967):
    row 0 takes the first residue's set (0 for an empty A), row i >= 1 the set of residue i."""
    tracks = np.asarray(tracks, np.uint16)
    oa = np.asarray(off_a, np.int64)
    N = len(oa) - 1
    out = np.zeros(int(oa[-1]) + N, np.uint16)
    for q in range(N):
        m = int(oa[q + 1] - oa[q])
        t = tracks[oa[q]:oa[q + 1]]
        g = out[oa[q] + q: oa[q + 1] + q + 1]
        g[0] = t[0] if m else 0
        g[1:] = t
    return out


def hard_tracked(matrices, gaps, codes_a, off_a, codes_b, off_b, tracks, gap_tracks=None, *, gap_model="affine",
                 mode="global", precision="float32", with_grad=True):
    """Tracked hard path (or score-only): K parameter sets, matrices (K, na, na) and gaps (K, 4);
    tracks one set per residue of A, concatenated in pair order (pair q at off_a[q] ..
    off_a[q+1]); gap_tracks one per DP row 0..m (pair q at off_a[q] + q, m + 1 entries), or
    None for the default.  Scores (N,) and gradients (N, K*(na*na + 4)): K matrix blocks,
    then K rows of go_a, ge_a, go_b, ge_b."""
    mats = np.ascontiguousarray(matrices, np.float64)
    K, na = mats.shape[0], mats.shape[1]
    g4 = np.ascontiguousarray(gaps, np.float64).reshape(K, 4)
    p = _TParams(GAP_MODELS[gap_model], MODES[mode], na, K, _ptr(mats, ctypes.c_double), _ptr(g4, ctypes.c_double))
    ca, oa, cb, ob = _pairs(codes_a, off_a, codes_b, off_b)
    N = len(oa) - 1
    trk = np.ascontiguousarray(tracks, np.uint16)
    gtrk = default_gap_tracks(trk, oa) if gap_tracks is None else np.ascontiguousarray(gap_tracks, np.uint16)
    if trk.size == 0:
        trk = np.zeros(1, np.uint16)
    scores = np.zeros(N, np.float64)
    gs = K * (na * na + 4)
    grads = np.zeros((N, gs), np.float64) if with_grad else np.zeros((1, 1))
    rc = lib().nwo_hard_tracked(ctypes.byref(p), 1 if precision in ("float64", "double") else 0, N,
                                _ptr(ca, ctypes.c_uint8), _ptr(oa, ctypes.c_int64), _ptr(cb, ctypes.c_uint8),
                                _ptr(ob, ctypes.c_int64), _ptr(trk, ctypes.c_uint16), _ptr(gtrk, ctypes.c_uint16),
                                1 if with_grad else 0, _ptr(scores, ctypes.c_double), _ptr(grads, ctypes.c_double))
    if rc == -2:
        raise ValueError("oracle: a track names a set outside [0, K)")
    if rc:
        raise MemoryError("oracle: allocation failed")
    return scores, (grads if with_grad else None)


def soft_log_tracked(matrices, gaps, codes_a, off_a, codes_b, off_b, tracks, gap_tracks=None, *,
                     gap_model="affine", mode="global", temperature=1.0):
    """Tracked log-space double forward-backward: K parameter sets (matrices (K, na, na), gaps
    (K, 4)), tracks / gap_tracks as for hard_tracked.  Scores T*logZ (N,) and expected counts
    (N, K*(na*na + 4)): K matrix blocks, then K rows of go_a, ge_a, go_b, ge_b."""
    mats = np.ascontiguousarray(matrices, np.float64)
    K, na = mats.shape[0], mats.shape[1]
    g4 = np.ascontiguousarray(gaps, np.float64).reshape(K, 4)
    p = _TParams(GAP_MODELS[gap_model], MODES[mode], na, K, _ptr(mats, ctypes.c_double), _ptr(g4, ctypes.c_double))
    ca, oa, cb, ob = _pairs(codes_a, off_a, codes_b, off_b)
    N = len(oa) - 1
    trk = np.ascontiguousarray(tracks, np.uint16)
    gtrk = default_gap_tracks(trk, oa) if gap_tracks is None else np.ascontiguousarray(gap_tracks, np.uint16)
    if trk.size == 0:
        trk = np.zeros(1, np.uint16)
    scores = np.zeros(N, np.float64)
    grads = np.zeros((N, K * (na * na + 4)), np.float64)
    rc = lib().nwo_soft_log_tracked(ctypes.byref(p), float(temperature), N, _ptr(ca, ctypes.c_uint8),
                                    _ptr(oa, ctypes.c_int64), _ptr(cb, ctypes.c_uint8), _ptr(ob, ctypes.c_int64),
                                    _ptr(trk, ctypes.c_uint16), _ptr(gtrk, ctypes.c_uint16),
                                    _ptr(scores, ctypes.c_double), _ptr(grads, ctypes.c_double))
    if rc:
        raise MemoryError("oracle: allocation failed")
    return scores, grads


def hirschberg(matrix, gaps, codes_a, off_a, codes_b, off_b, *, precision="float32", pmax=False, cutoff=512):
    """The Hirschberg traceback of Global affine hard pairs (traceback "hirschberg", or
    "hirschberg_pmax" with pmax=True): every pair of more than `cutoff` rows through the
    recursion.  Returns a dict: split (N,) bool -- the pairs that split (the others run as
    Pointers in the reference and are left at zero here) -- scores (N,), grads (N, na*na+4),
    guides (pair i's m + 1 entries at off_a[i] + i), moves (pair i's at off_a[i] + off_b[i],
    0 diagonal / 1 a gap in B / 2 a gap in A) and n_moves (N,)."""
    p, mat = _params(matrix, gaps, "affine", "global")
    ca, oa, cb, ob = _pairs(codes_a, off_a, codes_b, off_b)
    N = len(oa) - 1
    na = mat.shape[0]
    scores = np.zeros(N, np.float64)
    grads = np.zeros((N, na * na + 4), np.float64)
    guides = np.zeros(int(oa[-1]) + N, np.int32)
    moves = np.zeros(max(int(oa[-1] + ob[-1]), 1), np.uint8)
    n_moves = np.zeros(N, np.int32)
    split = np.zeros(max(N, 1), np.uint8)
    rc = lib().nwo_hirschberg(ctypes.byref(p), 1 if precision in ("float64", "double") else 0, 1 if pmax else 0,
                              int(cutoff), N, _ptr(ca, ctypes.c_uint8), _ptr(oa, ctypes.c_int64),
                              _ptr(cb, ctypes.c_uint8), _ptr(ob, ctypes.c_int64), _ptr(scores, ctypes.c_double),
                              _ptr(grads, ctypes.c_double), _ptr(guides, ctypes.c_int32),
                              _ptr(moves, ctypes.c_uint8), _ptr(n_moves, ctypes.c_int32), _ptr(split, ctypes.c_uint8))
    if rc:
        raise MemoryError("oracle: allocation failed")
    return {"split": split[:N].astype(bool), "scores": scores, "grads": grads, "guides": guides, "moves": moves,
            "n_moves": n_moves}
