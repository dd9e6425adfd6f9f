"""GPU parity of tracked batches (K parameter sets chosen per DP row): hard and score-only
bit-exact against the reference engine's tracked golden vectors and the oracle, short and long
(column-panel) pairs; soft within 1e-5 (float32) / 1e-10 (float64) of the oracle and of the
reference's float32 pass, norm-relative per pair as in test_gpu_soft.py."""
import glob
import os

import numpy as np
import pytest

import oracle
from nwgrad_amd import workloads as wl
from tests.gpu_util import say

pytestmark = pytest.mark.gpu

TRACKED = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                 if "_tracks_" in p and "/soft_" not in p)
SOFT_TRACKED = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                      if "/soft_tracks_" in p)


def _params_k(alphabet, mats, gaps):
    import nwgrad_amd as nw
    return nw.AlignParams.from_arrays(np.asarray(mats, np.float64), np.asarray(gaps, np.float64), str(alphabet))


def _batch(alphabet, mats, gaps, ca, oa, cb, ob, tracks, gap_tracks=None, *, precision="float32", gap_model="affine",
           mode="global", grad_mode="hard", fill="interpair"):
    import nwgrad_amd as nw
    cls = nw.SeqPairBatch if precision == "float32" else nw.SeqPairBatchDouble
    b = cls.from_codes(ca, oa, cb, ob, _params_k(alphabet, mats, gaps), gap_model=gap_model, mode=mode,
                       grad_mode=grad_mode, tracks=tracks, gap_tracks=gap_tracks)
    b.fill = fill
    return b


def _rows(b):
    mats, gaps = b.grads()
    N = len(b)
    return np.concatenate([mats.reshape(N, -1), gaps.reshape(N, -1)], axis=1)


def _flat(p):
    d = p.to_dict()
    if "matrices" in d:
        return np.concatenate([d["matrices"].ravel(), d["gaps"].ravel()])
    return np.concatenate([d["matrix"].ravel(), [d["gap_open_a"], d["gap_extend_a"], d["gap_open_b"],
                                                  d["gap_extend_b"]]])


@pytest.mark.parametrize("fill", ["interpair", "striped"])
@pytest.mark.parametrize("path", TRACKED, ids=lambda p: os.path.basename(p)[:-4])
def test_tracked_golden_bit_exact(gpu, path, fill):
    z = np.load(path)
    gt = z["gap_tracks"] if "gap_tracks" in z else None
    say("case", os.path.basename(path), fill)
    b = _batch(z["alphabet"], z["matrices"], z["gaps_k"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"],
               z["tracks"], gt, precision=str(z["precision"]), gap_model=str(z["gap_model"]), mode=str(z["mode"]),
               grad_mode=str(z["grad_mode"]), fill=fill)
    assert b.tracked and b.num_sets == z["matrices"].shape[0]
    total = b.score_and_grad()
    np.testing.assert_array_equal(b.scores(), z["scores"])
    assert total == float(z["total"])
    if str(z["grad_mode"]) == "hard":
        np.testing.assert_array_equal(_rows(b), z["grads"])
        np.testing.assert_array_equal(_flat(b.compute_grad()), z["compute_grad"])
        np.testing.assert_array_equal(_flat(b.weighted_grad(np.ascontiguousarray(z["weights"]))),
                                      z["weighted_grad"])


@pytest.mark.parametrize("gm", ["affine", "linear"])
@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
def test_tracked_long_pairs(gpu, gm, mode):
    """Tracked pairs wider than 1,024 columns (three to seven 512-column panels) next to short ones,
    segment tracks over K = 4 sets, float32 and float64, against the oracle."""
    r = wl.rng(f"trk_long_{gm}_{mode}", "seqs")
    lens_b = [1100, 2049, 3500, 40, 700, 1025]
    A = [r.integers(0, 20, size=int(r.integers(0, 500))).astype(np.uint8) for _ in lens_b]
    B = [r.integers(0, 20, size=n).astype(np.uint8) for n in lens_b]
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    K = 4
    mats = np.stack([wl.random_matrix(r) for _ in range(K)])
    gaps = np.array([[10.5, 0.75, 9.25, 1.5], [8.0, 1.25, 11.0, 0.5], [12.0, 2.0, 6.5, 1.5], [9.0, 1.0, 9.0, 1.0]])
    if gm == "linear":
        gaps = gaps * np.array([0.0, 1.0, 0.0, 1.0])
    tracks = np.repeat(r.integers(0, K, size=int(oa[-1]) // 7 + 1), 7)[:int(oa[-1])].astype(np.uint16)
    for precision in ("float32", "float64"):
        say("case tracked long", gm, mode, precision)
        s_ref, g_ref = oracle.hard_tracked(mats, gaps, ca, oa, cb, ob, tracks, gap_model=gm, mode=mode,
                                           precision=precision)
        b = _batch(wl.AA, mats, gaps, ca, oa, cb, ob, tracks, precision=precision, gap_model=gm, mode=mode)
        b.score_and_grad()
        np.testing.assert_array_equal(b.scores(), s_ref)
        np.testing.assert_array_equal(_rows(b), g_ref)
        b = _batch(wl.AA, mats, gaps, ca, oa, cb, ob, tracks, precision=precision, gap_model=gm, mode=mode,
                   grad_mode="none")
        b.score_and_grad()
        np.testing.assert_array_equal(b.scores(), s_ref)


def test_equal_sets_match_untracked(gpu):
    """K copies of one set under random tracks: the untracked scores, and per-set counts that
    sum to the untracked counts."""
    from tests.gpu_util import batch, grads_rows
    w = wl.random_pairs("trk_equal_gpu", 200, wl.AA, (0, 300), (0, 1300), related=0.5)
    K = 3
    tracks = wl.rng(w.name, "t").integers(0, K, size=int(w.off_a[-1])).astype(np.uint16)
    b0 = batch(w.alphabet, w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b)
    b0.score_and_grad()
    b1 = _batch(w.alphabet, np.stack([w.matrix] * K), np.stack([w.gaps] * K), w.codes_a, w.off_a, w.codes_b,
                w.off_b, tracks)
    b1.score_and_grad()
    np.testing.assert_array_equal(b1.scores(), b0.scores())
    mats, gaps = b1.grads()
    summed = np.concatenate([mats.sum(axis=1).reshape(len(b1), -1), gaps.sum(axis=1)], axis=1)
    np.testing.assert_array_equal(summed, grads_rows(b0))


def test_track_errors(gpu):
    w = wl.random_pairs("trk_err", 5, wl.AA, (5, 30), (5, 30))
    mats = np.stack([w.matrix] * 2)
    gaps = np.stack([w.gaps] * 2)
    bad = np.full(int(w.off_a[-1]), 2, np.uint16)       # set 2 of 2
    with pytest.raises(ValueError):
        _batch(w.alphabet, mats, gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, bad)
    with pytest.raises(ValueError):                     # wrong length
        _batch(w.alphabet, mats, gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, np.zeros(3, np.uint16))
    ok = np.zeros(int(w.off_a[-1]), np.uint16)
    b = _batch(w.alphabet, mats, gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, ok)
    import nwgrad_amd as nw
    with pytest.raises(ValueError):   # params with fewer sets than a track names
        b2 = _batch(w.alphabet, mats, gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, ok + 1)
        b2.set_params(nw.AlignParams(w.matrix, alphabet=w.alphabet))
    b.score_and_grad()
    assert b.grads()[0].shape == (5, 2, 20, 20)


def _rel(s, g, s_ref, g_ref):
    es = np.max(np.abs(s - s_ref) / np.maximum(1.0, np.abs(s_ref)))
    scale = np.maximum(np.max(np.abs(g_ref), axis=1), 1e-300)
    return float(es), float(np.max(np.max(np.abs(g - g_ref), axis=1) / scale))


@pytest.mark.parametrize("fill", ["interpair", "striped"])
@pytest.mark.parametrize("path", SOFT_TRACKED, ids=lambda p: os.path.basename(p)[:-4])
def test_soft_tracked_golden(gpu, path, fill):
    """Tracked soft (K = 3 sets): against the reference's own result (its float32 scaled pass
    or its float64 pass) and against the oracle."""
    z = np.load(path)
    gt = z["gap_tracks"] if "gap_tracks" in z else None
    prec = str(z["precision"])
    say("case", os.path.basename(path), fill)
    s_log, g_log = oracle.soft_log_tracked(z["matrices"], z["gaps_k"], z["codes_a"], z["off_a"], z["codes_b"],
                                           z["off_b"], z["tracks"], gt, gap_model=str(z["gap_model"]),
                                           mode=str(z["mode"]))
    b = _batch(z["alphabet"], z["matrices"], z["gaps_k"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"],
               z["tracks"], gt, precision=prec, gap_model=str(z["gap_model"]), mode=str(z["mode"]),
               grad_mode="soft", fill=fill)
    assert b.tracked and b.num_sets == 3
    b.score_and_grad()
    s, g = b.scores(), _rows(b)
    tol = 1e-5 if prec == "float32" else 1e-10
    es, eg = _rel(s, g, s_log, g_log)
    say(f"  vs oracle: score {es:.2e}, grad {eg:.2e}")
    assert es <= tol and eg <= tol
    es, eg = _rel(s, g, z["scores"], z["grads"])
    say(f"  vs reference: score {es:.2e}, grad {eg:.2e}")
    assert es <= max(tol, 1e-9) and eg <= max(tol, 1e-9)


@pytest.mark.parametrize("gm", ["affine", "linear"])
@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
def test_soft_tracked_wide_pairs(gpu, gm, mode):
    """Tracked soft pairs from one to five column panels (the persistent launch pipelines the
    longest across waves) beside short ones, K = 4 sets over segment tracks and random gap
    tracks, float32 and float64, both fills, against the oracle."""
    r = wl.rng(f"soft_trk_wide_{gm}_{mode}", "seqs")
    lens_b = [40, 300, 700, 1100, 1600, 2049, 90, 513]
    B = [r.integers(0, 20, size=n).astype(np.uint8) for n in lens_b]
    # related pairs of about equal lengths: a gap run across most of a 2,000-column pair does
    # not fit the scaled pass's range (in the reference either)
    A = [wl.mutate(r, b_, 20, 0.2, 0.02, 2.0) for b_ in B]
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    K = 4
    mats = np.stack([wl.random_matrix(r) for _ in range(K)])
    gaps = np.array([[10.5, 0.75, 9.25, 1.5], [8.0, 1.25, 11.0, 0.5], [12.0, 2.0, 6.5, 1.5], [9.0, 1.0, 9.0, 1.0]])
    if gm == "linear":
        gaps = gaps * np.array([0.0, 1.0, 0.0, 1.0])
    tracks = np.repeat(r.integers(0, K, size=int(oa[-1]) // 7 + 1), 7)[:int(oa[-1])].astype(np.uint16)
    gtr = r.integers(0, K, size=int(oa[-1]) + len(A)).astype(np.uint16)
    s_ref, g_ref = oracle.soft_log_tracked(mats, gaps, ca, oa, cb, ob, tracks, gtr, gap_model=gm, mode=mode)
    for precision, tol in (("float32", 1e-5), ("float64", 1e-10)):
        for fill in ("interpair", "striped"):
            say("case soft tracked wide", gm, mode, precision, fill)
            b = _batch(wl.AA, mats, gaps, ca, oa, cb, ob, tracks, gtr, precision=precision, gap_model=gm, mode=mode,
                       grad_mode="soft", fill=fill)
            b.score_and_grad()
            es, eg = _rel(b.scores(), _rows(b), s_ref, g_ref)
            say(f"  score {es:.2e}, grad {eg:.2e}, wide {b.wide_info if fill == 'interpair' else None}")
            assert es <= tol and eg <= tol


def test_soft_equal_sets_match_untracked(gpu):
    """K copies of one set under random tracks: the untracked soft scores, and per-set counts
    that sum to the untracked counts (within the soft tolerance)."""
    from tests.gpu_util import batch, grads_rows
    w = wl.random_pairs("soft_trk_equal_gpu", 120, wl.AA, (0, 300), (0, 900), related=1.0)
    K = 3
    r = wl.rng(w.name, "t")
    tracks = r.integers(0, K, size=int(w.off_a[-1])).astype(np.uint16)
    gtr = r.integers(0, K, size=int(w.off_a[-1]) + w.n_pairs).astype(np.uint16)
    b0 = batch(w.alphabet, w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, grad_mode="soft")
    b0.score_and_grad()
    b1 = _batch(w.alphabet, np.stack([w.matrix] * K), np.stack([w.gaps] * K), w.codes_a, w.off_a, w.codes_b,
                w.off_b, tracks, gtr, grad_mode="soft")
    b1.score_and_grad()
    mats, gaps = b1.grads()
    summed = np.concatenate([mats.sum(axis=1).reshape(len(b1), -1), gaps.sum(axis=1)], axis=1)
    es, eg = _rel(b1.scores(), summed, b0.scores(), grads_rows(b0))
    say(f"  score {es:.2e}, grad {eg:.2e}")
    assert es <= 1e-5 and eg <= 1e-5


@pytest.mark.parametrize("gm", ["affine", "linear"])
def test_tracked_table_budget(gpu, gm):
    """Tracked short pairs under a pointer-code budget that holds only a few dozen slots: the
    default fill runs persistent, each wave filling many tiles in turn, so every tile's flush
    must leave its per-set histograms and gap counters at zero for the next.  Matrix tracks
    and gap tracks over K = 3 sets, against the oracle, and unchanged from the plain launch."""
    r = wl.rng(f"trk_budget_{gm}", "seqs")
    A = [r.integers(0, 20, size=int(r.integers(1, 120))).astype(np.uint8) for _ in range(3000)]
    B = [r.integers(0, 20, size=int(r.integers(1, 120))).astype(np.uint8) for _ in range(3000)]
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    K = 3
    mats = np.stack([wl.random_matrix(r) for _ in range(K)])
    gaps = np.array([[10.5, 0.75, 9.25, 1.5], [8.0, 1.25, 11.0, 0.5], [12.0, 2.0, 6.5, 1.5]])
    if gm == "linear":
        gaps = gaps * np.array([0.0, 1.0, 0.0, 1.0])
    tracks = r.integers(0, K, size=int(oa[-1])).astype(np.uint16)
    gtracks = r.integers(0, K, size=int(oa[-1]) + len(A)).astype(np.uint16)
    s_ref, g_ref = oracle.hard_tracked(mats, gaps, ca, oa, cb, ob, tracks, gtracks, gap_model=gm)
    b = _batch(wl.AA, mats, gaps, ca, oa, cb, ob, tracks, gtracks, gap_model=gm)
    b.score_and_grad()
    assert b.hard_launches == (1, 0, 0)
    np.testing.assert_array_equal(b.scores(), s_ref)
    np.testing.assert_array_equal(_rows(b), g_ref)
    b.set_table_budget(1 << 20)
    b.score_and_grad()
    launches, persistent, grid = b.hard_launches
    say("case tracked budget", gm, b.tiles, "tiles, persistent grid", grid)
    assert (launches, persistent) == (1, 1) and 1 <= grid < b.tiles // 4
    np.testing.assert_array_equal(b.scores(), s_ref)
    np.testing.assert_array_equal(_rows(b), g_ref)
