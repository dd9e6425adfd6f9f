"""The linear-memory traceback on the GPU: traceback="hirschberg" / "hirschberg_pmax" of Global
affine hard batches (aligner.hpp:4118-4560).

* Fixtures the reference engine wrote at base cases of 5 and 24 rows (tests/golden/hb_*.npz):
  scores, gradients, guides in the recursion's convention, stored paths and a banded pass that
  follows those guides, bit for bit, split and unsplit pairs alike.
* The oracle's restatement (oracle.hirschberg) on random pairs: several recursion levels,
  column panels, both carries, both precisions, bit for bit.
* The long-pair configuration: a 36,000 x 36,000 mutated pair through the default base case of
  512 rows, bit for bit against the oracle in float32 and float64, in a bounded scratch far
  below the Pointers table's 1.3 GB -- and with integer gaps its score is the Pointers fill's.
* What the path refuses: other problem types, tracks, negative penalties, graph capture.
"""
import glob
import os

import numpy as np
import pytest

import nwgrad_amd as nw
import oracle
from nwgrad_amd import workloads as wl
from tests.gpu_util import batch, grads_rows, params_of, say
from tests.test_oracle_hb import cigar_of

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(__file__), "golden")
HB = sorted(glob.glob(os.path.join(HERE, "hb_*.npz")))


def hb_batch(alphabet, matrix, gaps, ca, oa, cb, ob, precision, traceback, cutoff, **kw):
    cls = nw.SeqPairBatch if precision == "float32" else nw.SeqPairBatchDouble
    b = cls.from_codes(ca, oa, cb, ob, params_of(alphabet, matrix, gaps), traceback=traceback, **kw)
    b.hb_cutoff = int(cutoff)
    return b


def guides_of(b):
    return np.concatenate([np.asarray(b[i].guide_j, np.int32) for i in range(len(b))])


def fixture_batch(z):
    return hb_batch(str(z["alphabet"]), z["matrix"], z["gaps"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"],
                    str(z["precision"]), str(z["traceback"]), int(z["hb_cutoff"]))


@pytest.mark.parametrize("path", HB, ids=lambda p: os.path.basename(p)[:-4])
def test_golden(gpu, path):
    z = np.load(path)
    say("case", os.path.basename(path))
    b = fixture_batch(z)
    total = b.score_and_grad(keep_paths=True)
    split, levels, blocks, scratch, _ = b.hb_info
    say(f"  split {split} pairs, {levels} levels, {blocks} base blocks, {scratch} B scratch")
    assert split >= 10 and levels >= 1
    assert total == float(z["total"])
    np.testing.assert_array_equal(b.scores(), z["scores"])
    np.testing.assert_array_equal(grads_rows(b), z["grads"])
    np.testing.assert_array_equal(guides_of(b), z["guides"])
    assert b.cigars() == list(z["cigars"])
    co = z["coords_off"]
    for i in range(len(b)):
        assert b[i].aligned() == (str(z["aligned_a"][i]), str(z["aligned_b"][i]))
        np.testing.assert_array_equal(b[i].coordinates(), z["coords"][:, co[i]:co[i + 1]])
    # a banded pass after a parameter step follows the recursion's guides
    p1 = params_of(str(z["alphabet"]), z["matrix1"], z["gaps1"])
    for bw in z["bands"]:
        b = fixture_batch(z)
        b.score_and_grad()
        b.set_params(p1)
        t = b.banded_grad(int(bw), keep_paths=True)
        assert t == float(z[f"band{bw}_total"])
        np.testing.assert_array_equal(b.scores(), z[f"band{bw}_scores"])
        np.testing.assert_array_equal(grads_rows(b), z[f"band{bw}_grads"])
        np.testing.assert_array_equal(guides_of(b), z[f"band{bw}_guides"])
        assert b.cigars() == list(z[f"band{bw}_cigars"])


def check_vs_oracle(w, precision, traceback, cutoff, keep=True):
    b = hb_batch(w.alphabet, w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision, traceback, cutoff)
    b.score_and_grad(keep_paths=keep)
    o = oracle.hirschberg(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision=precision,
                          pmax=traceback == "hirschberg_pmax", cutoff=cutoff)
    s0, g0 = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision=precision)
    s, g = b.scores(), grads_rows(b)
    sp = o["split"]
    assert b.hb_info[0] == int(sp.sum())
    np.testing.assert_array_equal(s[sp], o["scores"][sp])
    np.testing.assert_array_equal(g[sp], o["grads"][sp])
    np.testing.assert_array_equal(s[~sp], s0[~sp])
    np.testing.assert_array_equal(g[~sp], g0[~sp])
    gj = guides_of(b)
    cig = b.cigars() if keep else None
    for q in np.nonzero(sp)[0]:
        lo, hi = w.off_a[q] + q, w.off_a[q + 1] + q + 1
        np.testing.assert_array_equal(gj[lo:hi], o["guides"][lo:hi], err_msg=str(q))
        if keep:
            at = w.off_a[q] + w.off_b[q]
            assert cig[q] == cigar_of(list(o["moves"][at:at + o["n_moves"][q]])), q
    return b, o


CASES = [(prec, tb, cutoff) for prec in ("float32", "float64") for tb in ("hirschberg", "hirschberg_pmax")
         for cutoff in (3, 40, 512)]


@pytest.mark.parametrize("precision,traceback,cutoff", CASES)
def test_random_vs_oracle(gpu, precision, traceback, cutoff):
    """Pairs of 1 to 3,000 residues (up to six column panels), related and unrelated, fractional
    gaps (the two carries round differently): many recursion levels at small cutoffs."""
    w = wl.random_pairs(f"hb_rand_{cutoff}", 40, wl.AA, (1, 3000), (1, 3000), gaps=(10.5, 0.7, 9.3, 1.1), related=0.5)
    say("case", precision, traceback, cutoff, w.n_pairs, "pairs", w.cells, "cells")
    b, o = check_vs_oracle(w, precision, traceback, cutoff)
    say("  hb_info", b.hb_info)


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_lopsided_and_ties(gpu, precision):
    """Integer gaps (ties everywhere) and lopsided pairs: a block can be all rows and no columns
    (W = 0), or a single row against thousands of columns (many panels in one base block)."""
    r = np.random.default_rng(5)
    la = [600, 2000, 5, 1500, 700, 3000, 64, 900]
    lb = [3, 1, 2500, 1500, 4000, 20, 3000, 0]
    ca = [r.integers(0, 20, size=n).astype(np.uint8) for n in la]
    cb = [r.integers(0, 20, size=n).astype(np.uint8) for n in lb]
    pa, oa = wl._pack(ca)
    pb, ob = wl._pack(cb)
    w = wl.random_pairs("hb_lop", 1, wl.AA, (1, 2), (1, 2), gaps=(11.0, 1.0, 11.0, 1.0))
    w.codes_a, w.off_a, w.codes_b, w.off_b = pa, oa, pb, ob
    for tb in ("hirschberg", "hirschberg_pmax"):
        say("case lopsided", precision, tb)
        b, o = check_vs_oracle(w, precision, tb, 4)
        # integer gaps: every sum is exact, so the recursion's score is the optimum
        s0, _ = oracle.hard(w.matrix, w.gaps, pa, oa, pb, ob, precision=precision)
        np.testing.assert_array_equal(b.scores(), s0)


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_long_36000_pair(gpu, precision):
    """The long-pair configuration's pair shape: A uniform protein of 36,000 residues, B a mutated
    copy (substitution 0.1, indels 0.01 per residue, mean length 5), global affine 11/1, through
    the default base case of 512 rows (7 levels).  Bit for bit against the oracle; scratch
    linear in the lengths."""
    w = wl.long_pairs(1, 36000)
    say("case 36000", precision, w.n_pairs, "pair", int(w.off_a[1]), "x", int(w.off_b[1]))
    b = hb_batch(w.alphabet, w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision, "hirschberg", 512)
    b.score_and_grad(keep_paths=True)
    split, levels, blocks, scratch, filled = b.hb_info
    say(f"  {filled} cells filled, {levels} levels, {blocks} base blocks, scratch {scratch / 2**20:.1f} MiB")
    o = oracle.hirschberg(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision=precision, cutoff=512)
    assert split == 1 and o["split"][0]
    assert b.scores()[0] == o["scores"][0]
    np.testing.assert_array_equal(grads_rows(b)[0], o["grads"][0])
    np.testing.assert_array_equal(guides_of(b), o["guides"])
    assert b.cigars()[0] == cigar_of(list(o["moves"][:o["n_moves"][0]]))
    cells = int(w.off_a[1]) * int(w.off_b[1])
    assert scratch < cells // 4, "the recursion's scratch should stay far below a byte per cell"
    s0, _ = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision=precision)
    assert b.scores()[0] == s0[0]   # integer params: the recursion's path is optimal


def test_refusals(gpu):
    w = wl.random_pairs("hb_ref", 4, wl.AA, (600, 700), (600, 700), gaps=(11.0, 1.0, 11.0, 1.0))
    p = params_of(w.alphabet, w.matrix, w.gaps)
    for gm, mode in (("linear", "global"), ("affine", "local"), ("affine", "glocal")):
        with pytest.raises(ValueError, match="Global affine"):
            nw.SeqPairBatch.from_codes(w.codes_a, w.off_a, w.codes_b, w.off_b, p, gap_model=gm, mode=mode,
                                       traceback="hirschberg")
    # score-only and soft batches take no traceback: accepted, unaffected
    for gd in ("none", "soft"):
        b = nw.SeqPairBatch.from_codes(w.codes_a, w.off_a, w.codes_b, w.off_b, p, grad_mode=gd, mode="local",
                                       traceback="hirschberg")
        b.score_and_grad()
    neg = params_of(w.alphabet, w.matrix, (11.0, -1.0, 11.0, 1.0))
    b = nw.SeqPairBatch.from_codes(w.codes_a, w.off_a, w.codes_b, w.off_b, neg, traceback="hirschberg")
    with pytest.raises(ValueError, match="non-negative gap"):
        b.score_and_grad()
    b.hb_cutoff = 100000   # no pair splits: runs as Pointers, negative gaps allowed
    b.score_and_grad()
    b = nw.SeqPairBatch.from_codes(w.codes_a, w.off_a, w.codes_b, w.off_b, p, traceback="hirschberg",
                                   tracks=np.zeros(int(w.off_a[-1]), np.uint16))
    with pytest.raises(ValueError, match="tracked"):
        b.score_and_grad()
    with pytest.raises(ValueError, match="unknown traceback"):
        nw.SeqPairBatch.from_codes(w.codes_a, w.off_a, w.codes_b, w.off_b, p, traceback="zigzag")
