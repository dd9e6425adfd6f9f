"""GPU parity of the soft path: expected counts and T*log Z against the oracle's log-space
double forward-backward and the reference's golden vectors.

Tolerance (the target is 1e-5 relative), element by element: scores
|s - s_ref| <= 1e-5 * max(1, |s_ref|); every gradient entry |g - g_ref| <= 1e-5 * |g_ref| +
1e-7 * max_k |g_ref[pair, k]| (a floor of 1e-7 of the pair's largest expected count, for counts
that are rounding noise).  The float64 pass is held to 1e-10 with the same floor relative to
1e-10 (1e-12).
"""
import glob
import os

import numpy as np
import pytest

import oracle
from nwgrad_amd import workloads as wl
from tests.gpu_util import SOFT_FLOOR, SOFT_RTOL, batch, from_workload, grads_rows, say, soft_errors

pytestmark = pytest.mark.gpu

GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")) if "/soft_" in p and "_tracks_" not in p)


def check(b, s_ref, g_ref, tol, lens=None):
    b.score_and_grad()
    s, g = b.scores(), grads_rows(b)
    floor = SOFT_FLOOR * (tol / SOFT_RTOL)
    es, ee, en = soft_errors(s, g, s_ref, g_ref, floor)
    say(f"  max errors: score {es:.3e}, gradient element-wise {ee:.3e}, norm-relative {en:.3e} (tol {tol:g}, "
        f"floor {floor:g})")
    if not (es <= tol and ee <= tol) and lens is not None:
        ps = np.abs(s - s_ref) / np.maximum(1.0, np.abs(s_ref))
        scale = np.max(np.abs(g_ref), axis=1, keepdims=True)
        pe = np.max(np.abs(g - g_ref) / np.maximum(np.abs(g_ref) + floor * scale, 1e-300), axis=1)
        bad = np.nonzero(~((ps <= tol) & (pe <= tol)))[0]
        for k in bad[:12]:
            worst = int(np.argmax(np.abs(g[k] - g_ref[k]) / np.maximum(np.abs(g_ref[k]) + floor * scale[k], 1e-300)))
            say(f"    pair {k}: m={lens[0][k]} n={lens[1][k]} score {s[k]!r} want {s_ref[k]!r} "
                f"score err {ps[k]:.2e} grad err {pe[k]:.2e} worst field {worst} ({g[k][worst]!r} vs "
                f"{g_ref[k][worst]!r})")
    assert es <= tol and ee <= tol, (es, ee, en)
    return es, ee, en


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: os.path.basename(p)[:-4])
def test_golden(gpu, path):
    z = np.load(path)
    T = float(z["temperature"]) if "temperature" in z else 1.0
    say("case", os.path.basename(path))
    for prec, tol in (("float32", 1e-5), ("float64", 1e-10)):
        if prec == "float64" and str(z["precision"]) == "float32":
            continue
        b = batch(str(z["alphabet"]), z["matrix"], z["gaps"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"],
                  precision=prec, gap_model=str(z["gap_model"]), mode=str(z["mode"]), grad_mode="soft")
        b.soft_temperature = T
        # the golden file holds the reference's own float32 or double result; hold float32 to it at
        # 1e-5, and both to the log-space oracle
        s_log, g_log = oracle.soft_log(z["matrix"], z["gaps"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"],
                                       gap_model=str(z["gap_model"]), mode=str(z["mode"]), temperature=T)
        check(b, z["scores"], z["grads"], 1e-5 if prec == "float32" else max(tol, 1e-9))
        check(b, s_log, g_log, tol)


@pytest.mark.parametrize("gm", ["affine", "linear"])
@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
def test_random_vs_oracle(gpu, gm, mode):
    """len(B) 0..1024 (one to three column panels of 512), ragged len(A), related and unrelated
    pairs, several pairs of different lane-group widths in one wavefront."""
    r = wl.rng(f"soft_{gm}_{mode}", "lens")
    nb = np.concatenate([np.arange(0, 24), r.integers(0, 1025, size=24), [1024, 16, 17, 256, 257]])
    na_ = np.concatenate([r.integers(0, 200, size=len(nb) - 4), [0, 1, 2, 700]])
    A = [r.integers(0, 20, size=int(x)).astype(np.uint8) for x in na_]
    B = [r.integers(0, 20, size=int(x)).astype(np.uint8) for x in nb]
    for k in range(0, len(A), 3):   # a third related
        if len(A[k]):
            B[k] = wl.mutate(r, A[k], 20, 0.2, 0.03, 3.0)
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    mat = wl.random_matrix(r)
    gaps = np.array([11.0, 1.0, 10.0, 1.5])
    say("case soft random", gm, mode, len(nb), "pairs")
    s_ref, g_ref = oracle.soft_log(mat, gaps, ca, oa, cb, ob, gap_model=gm, mode=mode)
    for prec, tol in (("float32", 1e-5), ("float64", 1e-10)):
        for fill in ("interpair", "striped"):
            b = batch(wl.AA, mat, gaps, ca, oa, cb, ob, precision=prec, gap_model=gm, mode=mode, grad_mode="soft",
                      fill=fill)
            check(b, s_ref, g_ref, tol, (np.diff(oa), np.diff(ob)))


def test_protein_soft_homologs_sample(gpu):
    """A seeded sample of the soft configuration (homologs and decoys, lengths 50-1500)."""
    w = wl.protein_soft_homologs(256, 256)
    idx = np.concatenate([np.arange(0, 256, 4), np.arange(256, 512, 4)])
    w = w.subset(idx)
    say("case protein_soft_homologs sample", w.n_pairs, "pairs", w.cells, "cells")
    s_ref, g_ref = oracle.soft_log(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b)
    lens = (np.diff(w.off_a), np.diff(w.off_b))
    b = from_workload(w, grad_mode="soft")
    check(b, s_ref, g_ref, 1e-5, lens)
    say(f"  float32 pairs rerun in double: {b.soft_reruns} of {w.n_pairs}")
    b = from_workload(w, grad_mode="soft", precision="float64")
    check(b, s_ref, g_ref, 1e-10, lens)


def test_weighted_grad_soft(gpu):
    w = wl.random_pairs("soft_weighted", 300, wl.AA, (30, 200), (30, 200), related=0.5)
    s_ref, g_ref = oracle.soft_log(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b)
    b = from_workload(w, grad_mode="soft")
    check(b, s_ref, g_ref, 1e-5, (np.diff(w.off_a), np.diff(w.off_b)))
    wts = wl.rng("soft_weighted", "w").normal(size=w.n_pairs)
    got = b.weighted_grad(wts).to_dict()
    flat = np.concatenate([got["matrix"].ravel(), [got["gap_open_a"], got["gap_extend_a"], got["gap_open_b"],
                                                   got["gap_extend_b"]]])
    want = oracle.weighted_sum(g_ref, wts)
    assert np.max(np.abs(flat - want)) <= 1e-5 * np.max(np.abs(want))
    # the device reduction equals the reference's blocked sum of the device rows, bit for bit
    np.testing.assert_array_equal(flat, oracle.weighted_sum(grads_rows(b), wts))


@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
def test_column_panels(gpu, mode):
    """Pairs wider than one wavefront's 512 columns run panel by panel, up to seven panels,
    sharing tiles with narrower pairs."""
    r = wl.rng(f"panels_{mode}", "lens")
    nb = np.array([1025, 1100, 2048, 2049, 2500, 3100, 40, 900, 1024, 1500])
    na_ = np.array([3, 150, 80, 1, 60, 40, 200, 120, 70, 0])
    A = [r.integers(0, 20, size=int(x)).astype(np.uint8) for x in na_]
    B = [r.integers(0, 20, size=int(x)).astype(np.uint8) for x in nb]
    B[1][:len(A[1])] = wl.mutate(r, A[1], 20, 0.1, 0.0, 1.0)[:len(A[1])]
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    mat = wl.random_matrix(r)
    gaps = np.array([11.0, 1.0, 11.0, 1.0])
    say("case soft panels", mode, len(nb), "pairs")
    s_ref, g_ref = oracle.soft_log(mat, gaps, ca, oa, cb, ob, mode=mode)
    for prec, tol in (("float32", 1e-5), ("float64", 1e-10)):
        for fill in ("interpair", "striped"):
            b = batch(wl.AA, mat, gaps, ca, oa, cb, ob, precision=prec, mode=mode, grad_mode="soft", fill=fill)
            check(b, s_ref, g_ref, tol, (np.diff(oa), np.diff(ob)))


@pytest.mark.parametrize("alphabet", ["protein", "dna"])
def test_wide_pipeline_pool(gpu, alphabet):
    """Pairs wider than 512 columns run (persistent fill) as panel items pipelined across waves:
    a panel's forward and its backward are separate items, waiting only for items queued before
    them.  With the soft scratch held small the wide tiles go round a small circular pool -- a
    tile's first item waits for the tiles whose rows it reuses -- down to a pool of the largest
    tile alone.  The results do not depend on the pool: bit for bit the same as with the full
    scratch, and within tolerance of the oracle.  DNA packs several wide pairs into one tile."""
    r = wl.rng(f"soft_ring_{alphabet}", "lens")
    na_alpha = 20 if alphabet == "protein" else 4
    n = 36
    nb = np.concatenate([r.integers(513, 2600, size=n), r.integers(1, 400, size=12)])
    la = np.concatenate([r.integers(0, 260, size=n), r.integers(0, 300, size=12)])
    A = [r.integers(0, na_alpha, size=int(x)).astype(np.uint8) for x in la]
    B = [r.integers(0, na_alpha, size=int(x)).astype(np.uint8) for x in nb]
    for k in range(0, len(A), 2):   # half related: A is a mutated prefix of B
        if len(A[k]):
            A[k] = wl.mutate(r, B[k][:len(A[k])], na_alpha, 0.2, 0.02, 3.0)
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    mat = wl.random_matrix(r, na_alpha)
    gaps = np.array([10.0, 1.0, 9.0, 1.5])
    alpha = wl.AA if alphabet == "protein" else wl.DNA
    say("case soft wide pool", alphabet, len(nb), "pairs")
    s_ref, g_ref = oracle.soft_log(mat, gaps, ca, oa, cb, ob)
    lens = (np.diff(oa), np.diff(ob))
    for prec, tol in (("float32", 1e-5), ("float64", 1e-10)):
        b = batch(alpha, mat, gaps, ca, oa, cb, ob, precision=prec, grad_mode="soft")
        check(b, s_ref, g_ref, tol, lens)
        tiles, pooled, pool, lag = b.wide_info
        say(f"  {prec}: {tiles} pipelined tiles, pool {pool} rows, lag {lag}, launch {b.launch_info}")
        assert pooled == tiles
        assert tiles >= 1 and pool >= 1   # a small batch: every wide tile is pipelined
        s0, g0 = b.scores(), grads_rows(b)
        for scratch in (40 << 20, 1):   # a pool of a few tiles, then of the largest tile alone
            b.set_soft_scratch(scratch)
            b.score_and_grad()
            tiles2, _, pool2, lag2 = b.wide_info
            say(f"  scratch {scratch}: pool {pool2} rows, lag {lag2}")
            assert tiles2 == tiles and 1 <= pool2 < pool
            np.testing.assert_array_equal(b.scores(), s0)
            np.testing.assert_array_equal(grads_rows(b), g0)
        b.set_soft_scratch(0)


def test_wide_pipeline_mixed(gpu):
    """Pipelined and whole wide tiles in one launch: a wide tile whose panels in sequence would
    outlast the launch's work per resident wave (here the two 1,500 x 1,800 pairs) runs as
    panel items across waves, the short wide pairs (20 x 600) run whole, next to 2,500
    one-panel pairs.  The wide pairs against the oracle, the one-panel ones against a batch of
    their own; unchanged bit for bit with the smallest pool."""
    r = wl.rng("soft_mixed_wide", "lens")
    la = np.concatenate([r.integers(150, 250, size=2500), np.full(300, 20), [1500, 1500]])
    lb = np.concatenate([r.integers(250, 350, size=2500), np.full(300, 600), [1800, 1800]])
    A = [r.integers(0, 20, size=int(x)).astype(np.uint8) for x in la]
    B = [wl.mutate(r, a, 20, 0.25, 0.02, 3.0) for a in A]
    B = [np.concatenate([b_, r.integers(0, 20, size=max(0, int(n) - len(b_))).astype(np.uint8)])[:int(n)]
         for b_, n in zip(B, lb)]
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    mat = wl.random_matrix(r)
    gaps = np.array([11.0, 1.0, 11.0, 1.0])
    say("case soft mixed wide", len(A), "pairs")
    # the oracle on the wide pairs (the one-panel ones against a batch of their own)
    wide = np.arange(2500, len(A))
    cwa, owa = wl._pack([A[k] for k in wide])
    cwb, owb = wl._pack([B[k] for k in wide])
    s_w, g_w = oracle.soft_log(mat, gaps, cwa, owa, cwb, owb, mode="glocal")
    na_ = wl._pack(A[:2500])
    nb_ = wl._pack(B[:2500])
    bn = batch(wl.AA, mat, gaps, na_[0], na_[1], nb_[0], nb_[1], mode="glocal", grad_mode="soft")
    bn.score_and_grad()
    s_ref = np.concatenate([bn.scores(), s_w])
    g_ref = np.concatenate([grads_rows(bn), g_w])
    lens = (np.diff(oa), np.diff(ob))
    b = batch(wl.AA, mat, gaps, ca, oa, cb, ob, mode="glocal", grad_mode="soft")
    check(b, s_ref, g_ref, 1e-5, lens)
    tiles, pooled, pool, lag = b.wide_info
    say(f"  {tiles} pipelined tiles, {pooled} in the pool, pool {pool} rows, lag {lag}, launch {b.launch_info}")
    assert tiles == 2 and pooled == 2   # the short wide pairs run whole in block slots
    s0, g0 = b.scores(), grads_rows(b)
    b.set_soft_scratch(1)
    b.score_and_grad()
    assert b.wide_info[2] < pool
    np.testing.assert_array_equal(b.scores(), s0)
    np.testing.assert_array_equal(grads_rows(b), g0)
