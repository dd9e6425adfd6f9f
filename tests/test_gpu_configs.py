"""The long-pair configuration's arms at their full sizes, against the oracle.

* Soft: 4 pairs of 8,000 x 8,000 (A uniform protein, B = A mutated: substitution 0.1, indels
  0.01 per residue, geometric mean length 5), global affine 11/1.  Pair 0 in float32 and in
  float64 against the oracle's log-space double forward-backward (one 8,000^2 pair takes the
  oracle about half a minute); all four pairs float32 against float64, element by element.
  Float32 is held to 1e-5; float64 against the oracle to 1e-8 (two double computations, each
  rounding the sums of 6.4e7 cells).
* Hard float64: one 36,000 x 36,000 pair of the same law against the oracle's float64 fill,
  bit for bit (float32 at this size is in test_gpu_long.py).
"""
import numpy as np
import pytest

import oracle
from nwgrad_amd import workloads as wl
from tests.gpu_util import assert_soft_close, from_workload, grads_rows, say

pytestmark = pytest.mark.gpu


def test_long_soft_8000(gpu):
    w = wl.long_pairs(4, 8000, grad_mode="soft")
    say("case long_soft_8000", w.n_pairs, "pairs", w.cells, "cells")
    b32 = from_workload(w, grad_mode="soft")
    b32.score_and_grad()
    s32, g32 = b32.scores(), grads_rows(b32)
    say(f"  float32: {b32.soft_reruns} pairs rerun in double, wide tiles {b32.wide_info}")
    b64 = from_workload(w, grad_mode="soft", precision="float64")
    b64.score_and_grad()
    s64, g64 = b64.scores(), grads_rows(b64)
    assert_soft_close(s32, g32, s64, g64, label="float32 vs float64, 4 pairs: ")
    one = w.subset(np.array([0]))
    s_ref, g_ref = oracle.soft_log(one.matrix, one.gaps, one.codes_a, one.off_a, one.codes_b, one.off_b)
    assert_soft_close(s32[:1], g32[:1], s_ref, g_ref, label="float32 vs oracle, pair 0: ")
    # float64 against the double oracle: 1e-8 here, not the 1e-10 of the shorter pairs -- both sides
    # round every one of 6.4e7 cells' sums in double (measured: 9.4e-10 element-wise)
    assert_soft_close(s64[:1], g64[:1], s_ref, g_ref, rtol=1e-8, floor=1e-10, label="float64 vs oracle, pair 0: ")


def test_long_hard_36000_double(gpu):
    w = wl.long_pairs(1, 36000)
    say("case long_pairs_36000 float64", w.cells, "cells")
    s_ref, g_ref = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision="float64")
    b = from_workload(w, precision="float64")
    b.score_and_grad()
    np.testing.assert_array_equal(b.scores(), s_ref)
    np.testing.assert_array_equal(grads_rows(b), g_ref)
    # the score is the path's sum of parameters times counts (exact: integral parameters)
    theta = np.concatenate([w.matrix.ravel(), w.gaps])
    np.testing.assert_array_equal(grads_rows(b) @ theta, s_ref)
