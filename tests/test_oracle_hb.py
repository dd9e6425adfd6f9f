"""CPU pins of the oracle's Hirschberg restatement (oracle/nwg_oracle_hb.c), the checker of the
GPU's linear-memory traceback.

- Fixtures the reference engine wrote (tests/golden/hb_*.npz: traceback "hirschberg" and
  "hirschberg_pmax", base cases of 5 / 24 rows, float32 and float64, integer and fractional
  gaps): every pair that splits matches bit for bit -- score, gradient row, guide, CIGAR.
- What the mathematics fixes, independently of the reference: the path is a global alignment
  of the pair; its counts reproduce its score; its score is the optimum (brute force on tiny
  pairs, the Pointers fill's score where every sum is exact); where the optimum is unique the
  counts are its counts; with integer gaps both carries round nowhere, so they agree.
"""
import glob
import os

import numpy as np
import pytest

import oracle
from nwgrad_amd import workloads as wl
from tests import bruteforce as bf

HERE = os.path.join(os.path.dirname(__file__), "golden")
HB = sorted(glob.glob(os.path.join(HERE, "hb_*.npz")))


def cigar_of(moves):
    """The reference's CIGAR of a move list (0 diagonal 'M', 1 a gap in B 'I', 2 a gap in A 'D')."""
    out, k = [], 0
    while k < len(moves):
        e = k
        while e < len(moves) and moves[e] == moves[k]:
            e += 1
        out.append(f"{e - k}{'MID'[moves[k]]}")
        k = e
    return "".join(out)


def run_fixture(z):
    return oracle.hirschberg(z["matrix"], z["gaps"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"],
                             precision=str(z["precision"]), pmax=str(z["traceback"]) == "hirschberg_pmax",
                             cutoff=int(z["hb_cutoff"]))


@pytest.mark.parametrize("path", HB, ids=lambda p: os.path.basename(p)[:-4])
def test_fixture(path):
    z = np.load(path)
    o = run_fixture(z)
    oa, ob = z["off_a"], z["off_b"]
    split = np.nonzero(o["split"])[0]
    assert len(split) >= 10, "the fixture's pairs should mostly split"
    for q in split:
        assert o["scores"][q] == z["scores"][q], q
        np.testing.assert_array_equal(o["grads"][q], z["grads"][q], err_msg=str(q))
        lo, hi = oa[q] + q, oa[q + 1] + q + 1
        np.testing.assert_array_equal(o["guides"][lo:hi], z["guides"][lo:hi], err_msg=str(q))
        mv = o["moves"][oa[q] + ob[q]: oa[q] + ob[q] + o["n_moves"][q]]
        assert cigar_of(list(mv)) == str(z["cigars"][q]), q
    # the pairs that do not split ran as Pointers in the reference: the Pointers oracle's results
    s, g = oracle.hard(z["matrix"], z["gaps"], z["codes_a"], oa, z["codes_b"], ob, precision=str(z["precision"]))
    for q in np.nonzero(~o["split"])[0]:
        assert s[q] == z["scores"][q]
        np.testing.assert_array_equal(g[q], z["grads"][q])


def check_path(o, q, a, b, matrix, gaps, at=0):
    """The moves of pair q (from index `at`: off_a[q] + off_b[q]) align all of a with all of b,
    and their counts are the gradient row."""
    na = matrix.shape[0]
    mv = o["moves"][at:]
    m, n = len(a), len(b)
    i = j = 0
    prev = 3
    cnt = np.zeros(na * na + 4)
    for op in mv[:o["n_moves"][q]]:
        if op == 0:
            cnt[int(a[i]) * na + int(b[j])] += 1
            i += 1
            j += 1
        elif op == 1:
            cnt[na * na + 3] -= 1
            cnt[na * na + 2] -= prev != 1
            i += 1
        else:
            cnt[na * na + 1] -= 1
            cnt[na * na + 0] -= prev != 2
            j += 1
        prev = op
    assert (i, j) == (m, n)
    np.testing.assert_array_equal(cnt, o["grads"][q])
    return cnt


@pytest.mark.parametrize("pmax", [False, True])
def test_bruteforce_tiny(pmax):
    """Cutoff 1: every pair of two or more rows splits down to single rows.  The path's score is
    the brute-force optimum; counts . params is the score; a unique optimum gives its counts."""
    r = np.random.default_rng(11 + pmax)
    na = 4
    for trial in range(80):
        m, n = int(r.integers(2, 6)), int(r.integers(0, 6))
        a = r.integers(0, na, size=m).astype(np.uint8)
        b = r.integers(0, na, size=n).astype(np.uint8)
        mat = np.round(r.normal(0, 2, size=(na, na)), 3)
        gaps = np.round(r.uniform(0.1, 2.5, size=4), 3)
        o = oracle.hirschberg(mat, gaps, a, [0, m], b, [0, n], precision="float64", pmax=pmax, cutoff=1)
        assert o["split"][0]
        best, uniq = bf.hard(list(a), list(b), mat, gaps)
        assert o["scores"][0] == pytest.approx(best, abs=1e-9), trial
        cnt = check_path(o, 0, a, b, mat, gaps)
        assert float(cnt @ np.concatenate([mat.ravel(), gaps])) == pytest.approx(best, abs=1e-9)
        if uniq is not None:
            np.testing.assert_array_equal(cnt, uniq)


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_exact_sums(precision):
    """Integer params: every sum is exact, so the recursion's path scores the Pointers optimum
    exactly, counts . params equals it, and the two carries (which differ only by rounding)
    return the same paths."""
    w = wl.random_pairs("oracle_hb_exact", 60, wl.AA, (30, 400), (30, 400), gaps=(11.0, 1.0, 11.0, 1.0),
                        related=0.5)
    s0, _ = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision=precision)
    e = oracle.hirschberg(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision=precision, cutoff=7)
    p = oracle.hirschberg(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision=precision, cutoff=7,
                          pmax=True)
    theta = np.concatenate([w.matrix.ravel(), w.gaps])
    for q in range(w.n_pairs):
        assert e["split"][q]
        a = w.codes_a[w.off_a[q]:w.off_a[q + 1]]
        b = w.codes_b[w.off_b[q]:w.off_b[q + 1]]
        cnt = check_path(e, q, a, b, w.matrix, w.gaps, w.off_a[q] + w.off_b[q])
        assert e["scores"][q] == s0[q]
        assert float(cnt @ theta) == e["scores"][q]
        np.testing.assert_array_equal(e["grads"][q], p["grads"][q])
    np.testing.assert_array_equal(e["moves"], p["moves"])


def test_fractional_near_optimal():
    """Fractional gaps in double: the path is a legal alignment whose counts reproduce the
    score to rounding, and the score is the Pointers optimum to rounding (both carries)."""
    w = wl.random_pairs("oracle_hb_frac", 40, wl.AA, (30, 300), (30, 300), gaps=(10.5, 0.7, 9.3, 1.1), related=0.5)
    s0, _ = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision="float64")
    theta = np.concatenate([w.matrix.ravel(), w.gaps])
    for pmax in (False, True):
        o = oracle.hirschberg(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision="float64",
                              cutoff=9, pmax=pmax)
        for q in range(w.n_pairs):
            a = w.codes_a[w.off_a[q]:w.off_a[q + 1]]
            b = w.codes_b[w.off_b[q]:w.off_b[q + 1]]
            cnt = check_path(o, q, a, b, w.matrix, w.gaps, w.off_a[q] + w.off_b[q])
            assert float(cnt @ theta) == pytest.approx(o["scores"][q], abs=1e-9)
            assert o["scores"][q] == pytest.approx(s0[q], abs=1e-9)
