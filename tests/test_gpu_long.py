"""GPU parity of the long-pair hard and score-only paths (len(B) > 1024: column panels
pipelined across waves, then the panel-windowed walk), bit-exact against the oracle."""
import numpy as np
import pytest

import oracle
from nwgrad_amd import workloads as wl
from tests.gpu_util import batch, from_workload, grads_rows, say

pytestmark = pytest.mark.gpu


def _mixed(name, n_pairs, lens_b, max_a, related=0.5):
    """Pairs whose len(B) is drawn from `lens_b` (long and short), len(A) U{0..max_a}; with
    probability `related` B is a mutated copy of A cut or padded to its drawn length."""
    r = wl.rng(name, "seqs")
    A, B = [], []
    for k in range(n_pairs):
        nb = int(lens_b[k % len(lens_b)])
        a = r.integers(0, 20, size=int(r.integers(0, max_a + 1))).astype(np.uint8)
        if r.random() < related and len(a):
            b = wl.mutate(r, a, 20, 0.2, 0.05, 2.0)
            b = np.concatenate([b, r.integers(0, 20, size=max(0, nb - len(b))).astype(np.uint8)])[:nb]
        else:
            b = r.integers(0, 20, size=nb).astype(np.uint8)
        A.append(a)
        B.append(b)
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    return ca, oa, cb, ob


LENS_B = [1025, 1100, 2048, 2049, 3000, 4100, 17, 900, 1024, 1536]


@pytest.mark.parametrize("gm", ["affine", "linear"])
@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
def test_long_pairs_modes(gpu, gm, mode):
    """Long and short pairs in one batch (len(B) from 17 to 4100, one to five panels, len(A)
    from 0 to 700), float32 and float64, hard and score-only, both fills."""
    ca, oa, cb, ob = _mixed(f"long_{gm}_{mode}", 30, LENS_B, 700)
    mat = wl.random_matrix(wl.rng(f"long_{gm}_{mode}", "matrix"))
    gaps = np.array([4.5, 0.5, 3.25, 1.0])
    for precision in ("float32", "float64"):
        say("case long", gm, mode, precision)
        s_ref, g_ref = oracle.hard(mat, gaps, ca, oa, cb, ob, gap_model=gm, mode=mode, precision=precision)
        for fill in ("interpair", "striped"):
            b = batch(wl.AA, mat, gaps, ca, oa, cb, ob, precision=precision, gap_model=gm, mode=mode, fill=fill)
            total = b.score_and_grad()
            np.testing.assert_array_equal(b.scores(), s_ref)
            np.testing.assert_array_equal(grads_rows(b), g_ref)
            assert total == float(np.cumsum(s_ref)[-1])
        b = batch(wl.AA, mat, gaps, ca, oa, cb, ob, precision=precision, gap_model=gm, mode=mode, grad_mode="none")
        b.score_and_grad()
        np.testing.assert_array_equal(b.scores(), s_ref)


def test_long_many_panels(gpu):
    """Forty-panel pairs (len(B) 20,000, 512-column panels) with short and empty len(A), and one len(A) = 6,000:
    the panel pipeline over many waves, rows from 0 to 6,000."""
    r = wl.rng("long_many_panels", "seqs")
    A = [r.integers(0, 20, size=k).astype(np.uint8) for k in (0, 1, 2, 63, 64, 65, 300, 6000)]
    B = [r.integers(0, 20, size=20000).astype(np.uint8) for _ in A]
    B[-1] = np.concatenate([wl.mutate(r, A[-1], 20, 0.1, 0.01, 5.0), B[-1]])[:20000]
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    mat = wl.random_matrix(wl.rng("long_many_panels", "matrix"))
    gaps = np.array([11.0, 1.0, 11.0, 1.0])
    for mode in ("global", "local", "glocal"):
        say("case long_many_panels", mode)
        s_ref, g_ref = oracle.hard(mat, gaps, ca, oa, cb, ob, mode=mode)
        b = batch(wl.AA, mat, gaps, ca, oa, cb, ob, mode=mode)
        b.score_and_grad()
        np.testing.assert_array_equal(b.scores(), s_ref)
        np.testing.assert_array_equal(grads_rows(b), g_ref)


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_long_4096_homologs(gpu, precision):
    """Sixteen 4,096-residue protein pairs, B a mutated copy of A (substitution 0.1, indels
    0.01, mean 5): the 4,096 x 4,096 long-pair workload's shape, every score and count."""
    w = wl.long_pairs(16, 4096)
    say("case long_pairs_4096", precision, w.n_pairs, "pairs", w.cells, "cells")
    s_ref, g_ref = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision=precision)
    b = from_workload(w, precision=precision)
    b.score_and_grad()
    np.testing.assert_array_equal(b.scores(), s_ref)
    np.testing.assert_array_equal(grads_rows(b), g_ref)
    # the score is the path's sum of parameters times counts (exact: integral parameters)
    theta = np.concatenate([w.matrix.ravel(), w.gaps])
    np.testing.assert_array_equal(grads_rows(b) @ theta, s_ref)


def test_long_36000_pair(gpu):
    """One 36,000-residue pair (the longest-pair workload's shape): 36 panels, 1.3e9 cells."""
    w = wl.long_pairs(1, 36000)
    say("case long_pairs_36000", w.cells, "cells")
    s_ref, g_ref = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b)
    b = from_workload(w)
    b.score_and_grad()
    np.testing.assert_array_equal(b.scores(), s_ref)
    np.testing.assert_array_equal(grads_rows(b), g_ref)
