"""TEST INFRASTRUCTURE ONLY: the reference engine built from its own sources
(``oracle/_ref/libnwgref.so``, see ``Makefile.ref``), driven through ``ref_harness.cpp``.

``RefBatch`` mirrors the reference's ``SeqPairBatch`` / ``SeqPairBatchDouble`` Python class
(py_exports.cpp:787-1263) closely enough for parity tests and the CPU baseline of bench.py.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import REF_PATH

_lib = None
_TB = {"auto": 0, "pointers": 1, "scores": 2, "hirschberg": 3, "hirschberg_pmax": 4}
_GM = {"linear": 0, "affine": 1}
_MODE = {"global": 0, "local": 1, "glocal": 2}
_GRAD = {"none": 0, "hard": 1, "soft": 2}
_FILL = {"interpair": 0, "striped": 1, "rowwise": 2}


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(REF_PATH)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        L.nwgref_last_error.restype = ctypes.c_char_p
        L.nwgref_create.argtypes = [ctypes.c_int, ctypes.c_char_p, P(ctypes.c_double), P(ctypes.c_double),
                                    ctypes.c_long, ctypes.c_char_p, P(ctypes.c_long), ctypes.c_char_p,
                                    P(ctypes.c_long), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, P(vp)]
        L.nwgref_create_tracked.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, P(ctypes.c_double),
                                            P(ctypes.c_double), ctypes.c_long, ctypes.c_char_p, P(ctypes.c_long),
                                            ctypes.c_char_p, P(ctypes.c_long), P(ctypes.c_uint16),
                                            P(ctypes.c_uint16), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, P(vp)]
        L.nwgref_destroy.argtypes = [vp]
        L.nwgref_set_fill.argtypes = [vp, ctypes.c_int]
        L.nwgref_set_hb_cutoff.argtypes = [vp, ctypes.c_int]
        L.nwgref_set_soft.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.nwgref_set_params.argtypes = [vp, ctypes.c_char_p, P(ctypes.c_double), P(ctypes.c_double)]
        L.nwgref_score_and_grad.argtypes = [vp, P(ctypes.c_double)]
        L.nwgref_scores.argtypes = [vp, P(ctypes.c_double)]
        L.nwgref_grads.argtypes = [vp, P(ctypes.c_double), P(ctypes.c_double)]
        L.nwgref_weighted_grad.argtypes = [vp, P(ctypes.c_double), ctypes.c_long, P(ctypes.c_double)]
        L.nwgref_compute_grad.argtypes = [vp, P(ctypes.c_double)]
        L.nwgref_score_and_grad_keep.argtypes = [vp, ctypes.c_int, P(ctypes.c_double)]
        L.nwgref_banded_grad.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_double)]
        L.nwgref_guides.argtypes = [vp, P(ctypes.c_int32)]
        L.nwgref_path.argtypes = [vp, ctypes.c_long, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_long, P(ctypes.c_long),
                                  P(ctypes.c_int64), ctypes.c_long, P(ctypes.c_long), ctypes.c_char_p,
                                  ctypes.c_char_p]
        L.nwgref_align.argtypes = [vp, ctypes.c_char_p, P(ctypes.c_double), P(ctypes.c_double), ctypes.c_long,
                                   ctypes.c_char_p, P(ctypes.c_long), ctypes.c_char_p, P(ctypes.c_long),
                                   ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double)]
        _lib = L
    return _lib


_ERR = {1: ValueError, 2: RuntimeError, 3: ValueError, 4: IndexError, 5: RuntimeError}


def _check(rc):
    if rc:
        raise _ERR.get(rc, RuntimeError)(lib().nwgref_last_error().decode())


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class RefBatch:
    def __init__(self, seqs_a, seqs_b, alphabet, matrix, gaps, *, double=False, gap_model="affine",
                 mode="global", grad_mode="hard", n_threads=0, traceback="auto", tracks=None, gap_tracks=None):
        """matrix (n, n) and gaps (4,), or K sets: matrix (K, n, n) and gaps (K, 4), with
        `tracks` (one set per residue of A, concatenated in pair order) and `gap_tracks` (one
        per DP row 0..m of each pair, concatenated; None: the reference's default)."""
        L = lib()
        self.alphabet = alphabet
        self.n_alpha = len(alphabet)
        self.N = len(seqs_a)
        mat = np.ascontiguousarray(matrix, np.float64)
        self.K = mat.shape[0] if mat.ndim == 3 else 1
        if mat.ndim == 3 or tracks is not None:
            self._create_tracked(seqs_a, seqs_b, mat.reshape(self.K, self.n_alpha, self.n_alpha),
                                 np.ascontiguousarray(gaps, np.float64).reshape(self.K, 4), double, gap_model, mode,
                                 grad_mode, n_threads, tracks, gap_tracks)
            return
        sa = "".join(seqs_a).encode()
        sb = "".join(seqs_b).encode()
        oa = np.zeros(self.N + 1, np.int64)
        ob = np.zeros(self.N + 1, np.int64)
        oa[1:] = np.cumsum([len(s) for s in seqs_a])
        ob[1:] = np.cumsum([len(s) for s in seqs_b])
        self._keep = (sa, sb, oa, ob)
        mat = np.ascontiguousarray(matrix, np.float64)
        g = np.ascontiguousarray(gaps, np.float64)
        h = ctypes.c_void_p()
        _check(L.nwgref_create(1 if double else 0, alphabet.encode(), _d(mat), _d(g), self.N, sa,
                               oa.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), sb,
                               ob.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), _GM[gap_model], _MODE[mode],
                               _GRAD[grad_mode], n_threads, _TB[traceback], ctypes.byref(h)))
        self._h = h

    def _create_tracked(self, seqs_a, seqs_b, mats, gaps, double, gap_model, mode, grad_mode, n_threads, tracks,
                        gap_tracks):
        sa = "".join(seqs_a).encode()
        sb = "".join(seqs_b).encode()
        oa = np.zeros(self.N + 1, np.int64)
        ob = np.zeros(self.N + 1, np.int64)
        oa[1:] = np.cumsum([len(s) for s in seqs_a])
        ob[1:] = np.cumsum([len(s) for s in seqs_b])
        tk = None if tracks is None else np.ascontiguousarray(tracks, np.uint16)
        gk = None if gap_tracks is None else np.ascontiguousarray(gap_tracks, np.uint16)
        if tk is not None and tk.size == 0:
            tk = np.zeros(1, np.uint16)
        self._keep = (sa, sb, oa, ob, mats, gaps, tk, gk)
        P16 = ctypes.POINTER(ctypes.c_uint16)
        h = ctypes.c_void_p()
        _check(lib().nwgref_create_tracked(
            1 if double else 0, self.alphabet.encode(), self.K, _d(mats), _d(gaps), self.N, sa,
            oa.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), sb, ob.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
            None if tk is None else tk.ctypes.data_as(P16), None if gk is None else gk.ctypes.data_as(P16),
            _GM[gap_model], _MODE[mode], _GRAD[grad_mode], n_threads, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().nwgref_destroy(self._h)
            self._h = None

    def set_fill(self, fill):
        _check(lib().nwgref_set_fill(self._h, _FILL[fill]))

    def set_hb_cutoff(self, rows):
        _check(lib().nwgref_set_hb_cutoff(self._h, int(rows)))

    def set_soft(self, temperature=1.0, precision="float32", impl="scaled"):
        _check(lib().nwgref_set_soft(self._h, float(temperature), 0 if precision == "float32" else 1,
                                     {"scaled": 0, "scaled_or_log": 1, "log": 2}[impl]))

    def set_params(self, matrix, gaps):
        mat = np.ascontiguousarray(matrix, np.float64)
        g = np.ascontiguousarray(gaps, np.float64)
        _check(lib().nwgref_set_params(self._h, self.alphabet.encode(), _d(mat), _d(g)))

    def score_and_grad(self):
        t = ctypes.c_double()
        _check(lib().nwgref_score_and_grad(self._h, ctypes.byref(t)))
        return t.value

    def scores(self):
        out = np.zeros(self.N, np.float64)
        _check(lib().nwgref_scores(self._h, _d(out)))
        return out

    def grads(self):
        """(N, n, n) and (N, 4); K sets: (N, K, n, n) and (N, K, 4)."""
        na, K = self.n_alpha, getattr(self, "K", 1)
        mats = np.zeros((self.N, K, na, na) if K > 1 else (self.N, na, na), np.float64)
        gaps = np.zeros((self.N, K, 4) if K > 1 else (self.N, 4), np.float64)
        _check(lib().nwgref_grads(self._h, _d(mats), _d(gaps)))
        return mats, gaps

    def weighted_grad(self, w):
        """K*(n*n + 4) doubles: the K matrices, then K rows of go_a, ge_a, go_b, ge_b."""
        w = np.ascontiguousarray(w, np.float64)
        out = np.zeros(getattr(self, "K", 1) * (self.n_alpha ** 2 + 4), np.float64)
        _check(lib().nwgref_weighted_grad(self._h, _d(w), len(w), _d(out)))
        return out

    def compute_grad(self):
        out = np.zeros(getattr(self, "K", 1) * (self.n_alpha ** 2 + 4), np.float64)
        _check(lib().nwgref_compute_grad(self._h, _d(out)))
        return out

    def score_and_grad_keep(self, keep_paths):
        t = ctypes.c_double()
        _check(lib().nwgref_score_and_grad_keep(self._h, 1 if keep_paths else 0, ctypes.byref(t)))
        return t.value

    def banded_grad(self, bandwidth, keep_paths=False):
        t = ctypes.c_double()
        _check(lib().nwgref_banded_grad(self._h, int(bandwidth), 1 if keep_paths else 0, ctypes.byref(t)))
        return t.value

    def guides(self, off_a):
        """Every pair's guide_j concatenated (pair i's m + 1 entries at off_a[i] + i)."""
        out = np.zeros(int(off_a[-1]) + self.N, np.int32)
        _check(lib().nwgref_guides(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    def path(self, i, cap):
        """Pair i's stored path: (aligned_a, aligned_b, coordinates (2, k), cigar, cigar extended +
        soft-clipped).  cap: room for the aligned strings (len(a) + len(b) suffices)."""
        cap = max(int(cap), 4) + 1
        a = ctypes.create_string_buffer(cap)
        b = ctypes.create_string_buffer(cap)
        c0 = ctypes.create_string_buffer(cap * 8)
        c1 = ctypes.create_string_buffer(cap * 8)
        coords = np.zeros((2, cap), np.int64)
        cols = ctypes.c_long()
        k = ctypes.c_long()
        _check(lib().nwgref_path(self._h, int(i), a, b, cap, ctypes.byref(cols),
                                 coords.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), cap, ctypes.byref(k), c0, c1))
        n = cols.value
        return (a.raw[:n].decode(), b.raw[:n].decode(), coords[:, :k.value].copy(), c0.value.decode(),
                c1.value.decode())

    def align(self, seqs_a, seqs_b, matrix, gaps, band=0):
        """Streaming align() under (matrix, gaps): scores and the summed gradient; band > 0
        bands every pair around the diagonal guide."""
        sa = "".join(seqs_a).encode()
        sb = "".join(seqs_b).encode()
        n = len(seqs_a)
        oa = np.zeros(n + 1, np.int64)
        ob = np.zeros(n + 1, np.int64)
        oa[1:] = np.cumsum([len(x) for x in seqs_a])
        ob[1:] = np.cumsum([len(x) for x in seqs_b])
        mat = np.ascontiguousarray(matrix, np.float64)
        g = np.ascontiguousarray(gaps, np.float64)
        scores = np.zeros(max(n, 1), np.float64)
        grad = np.zeros(self.n_alpha ** 2 + 4, np.float64)
        _check(lib().nwgref_align(self._h, self.alphabet.encode(), _d(mat), _d(g), n, sa,
                                  oa.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), sb,
                                  ob.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), int(band), _d(scores), _d(grad)))
        return scores[:n], grad
