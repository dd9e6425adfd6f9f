"""The oracle pinned: the reference's golden vectors, its known answers, brute-force
enumeration and closed forms.  CPU only."""
import glob
import math
import os

import numpy as np
import pytest

import oracle
from nwgrad_amd import workloads as wl
from tests import bruteforce as bf

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def _args(z):
    return (z["matrix"], z["gaps"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"])


@pytest.mark.parametrize("path", [p for p in GOLDEN if ("/hard_" in p or "/score_" in p) and "_tracks_" not in p],
                         ids=lambda p: os.path.basename(p))
def test_hard_matches_reference_golden(path):
    """Bit-exact: scores, per-pair counts, compute_grad and weighted_grad of the reference."""
    z = np.load(path)
    gd = str(z["grad_mode"])
    s, g = oracle.hard(*_args(z), gap_model=str(z["gap_model"]), mode=str(z["mode"]),
                       precision=str(z["precision"]), with_grad=gd == "hard")
    np.testing.assert_array_equal(s, z["scores"])
    assert float(np.sum(s)) == float(z["total"]) or np.isclose(np.sum(s), z["total"])
    if gd == "hard":
        np.testing.assert_array_equal(g, z["grads"])
        np.testing.assert_array_equal(oracle.weighted_sum(g), z["compute_grad"])
        np.testing.assert_array_equal(oracle.weighted_sum(g, z["weights"]), z["weighted_grad"])


@pytest.mark.parametrize("path", [p for p in GOLDEN if "/soft_" in p and "_tracks_" not in p],
                         ids=lambda p: os.path.basename(p))
def test_soft_matches_reference_golden(path):
    """Log-space double vs the reference's scaled pass: 1e-12 against its double pass, and
    against its float32 pass within the reference's own float32 tolerance (2e-4 rel / 1e-5 abs,
    tests/Python/test_soft_float32.py:27-31)."""
    z = np.load(path)
    T = float(z["temperature"]) if "temperature" in z else 1.0
    s, g = oracle.soft_log(*_args(z), gap_model=str(z["gap_model"]), mode=str(z["mode"]), temperature=T)
    if str(z["precision"]) == "float64":
        np.testing.assert_allclose(s, z["scores"], rtol=1e-12, atol=1e-10)
        np.testing.assert_allclose(g, z["grads"], rtol=1e-10, atol=1e-11)
    else:
        np.testing.assert_allclose(s, z["scores"], rtol=2e-4, atol=1e-5)
        np.testing.assert_allclose(g, z["grads"], rtol=2e-4, atol=1e-5)


def _enc(s, alphabet=wl.AA):
    return np.array([alphabet.index(c) for c in s], np.uint8)


def _one(a, b, matrix, gaps, **kw):
    ca, cb = _enc(a), _enc(b)
    return oracle.hard(matrix, gaps, ca, np.array([0, len(ca)]), cb, np.array([0, len(cb)]), **kw)


UNIT = np.where(np.eye(20) > 0, 1.0, -1.0)


def test_known_answers_reference_cpp():
    """Literal counts from the reference's tests/cpp/test_gradient.cpp:35-154."""
    na = 20
    s, g = _one("ACDE", "ACDE", UNIT, [0, 1, 0, 1], gap_model="linear", mode="global")
    assert s[0] == 4.0
    for c in "ACDE":
        assert g[0, wl.AA.index(c) * na + wl.AA.index(c)] == 1.0
    assert g[0, :400].sum() == 4.0
    s, g = _one("A", "", UNIT, [0, 1, 0, 1], gap_model="linear", mode="global")
    assert g[0, :400].sum() == 0.0
    s, g = _one("ADE", "ACDE", UNIT, [0, 0.5, 0, 0.5], gap_model="linear", mode="global")
    assert g[0, :400].sum() == 3.0
    s, g = _one("ADE", "MMMADEM", UNIT, [0, 1, 0, 1], gap_model="linear", mode="local")
    assert g[0, :400].sum() == 3.0 and g[0, wl.AA.index("M") * na + wl.AA.index("M")] == 0.0
    s, g = _one("AAAA", "CCCC", UNIT, [0, 1, 0, 1], gap_model="linear", mode="local")
    assert g[0].sum() == 0.0 and s[0] == 0.0
    s, g = _one("ACDE", "ACDE", UNIT, [10, 1, 10, 1], gap_model="affine", mode="global")
    assert s[0] == 4.0 and g[0, :400].sum() == 4.0
    s, g = _one("A", "", UNIT, [10, 1, 10, 1], gap_model="affine", mode="global")
    assert g[0, :400].sum() == 0.0 and g[0, 402] == -1.0 and g[0, 403] == -1.0
    s, g = _one("ADE", "MMMADEM", UNIT, [10, 1, 10, 1], gap_model="affine", mode="local")
    assert g[0, :400].sum() == 3.0


def test_soft_closed_forms():
    """tests/Python/test_soft_gradient_known.py: single character global linear, Z = e^s + 2e^{-2g};
    "AA" vs "AA" zero matrix: E[diagonal steps] = (2 + 6e^{-2g}) / (1 + 6e^{-2g} + 6e^{-4g})."""
    r = np.random.default_rng(1)
    mat = r.integers(-4, 5, size=(20, 20)).astype(float)
    for a, b in [("A", "A"), ("A", "C"), ("W", "Y")]:
        ge = 1.0
        s = mat[wl.AA.index(a), wl.AA.index(b)]
        denom = math.exp(s) + 2 * math.exp(-2 * ge)
        ca, cb = _enc(a), _enc(b)
        lz, g = oracle.soft_log(mat, [0, ge, 0, ge], ca, [0, 1], cb, [0, 1], gap_model="linear", mode="global")
        assert lz[0] == pytest.approx(math.log(denom), rel=1e-12)
        assert g[0, wl.AA.index(a) * 20 + wl.AA.index(b)] == pytest.approx(math.exp(s) / denom, rel=1e-12)
    g_ = 0.7
    ca = _enc("AA")
    lz, g = oracle.soft_log(np.zeros((20, 20)), [0, g_, 0, g_], ca, [0, 2], ca, [0, 2], gap_model="linear")
    e2, e4 = math.exp(-2 * g_), math.exp(-4 * g_)
    assert g[0, 0] == pytest.approx((2 + 6 * e2) / (1 + 6 * e2 + 6 * e4), rel=1e-12)
    assert lz[0] == pytest.approx(math.log(1 + 6 * e2 + 6 * e4), rel=1e-12)


CASES = [(gm, mode) for gm in ("affine", "linear") for mode in ("global", "local", "glocal")]


@pytest.mark.parametrize("gm,mode", CASES)
def test_hard_vs_bruteforce(gm, mode):
    """Best score equals exhaustive enumeration; where the best alignment is unique, so are its
    counts; and always score == counts . params (the path the walk returns scores the score)."""
    r = np.random.default_rng(hash((gm, mode)) % 2**32)
    na = 4
    for trial in range(60):
        m, n = int(r.integers(0, 5)), int(r.integers(0, 5))
        a = r.integers(0, na, size=m).astype(np.uint8)
        b = r.integers(0, na, size=n).astype(np.uint8)
        mat = np.round(r.normal(0, 2, size=(na, na)), 3)
        gaps = np.round(r.uniform(0.1, 2.5, size=4), 3)
        s, g = oracle.hard(mat, gaps, a, [0, m], b, [0, n], gap_model=gm, mode=mode, precision="float64")
        best, uniq = bf.hard(list(a), list(b), mat, gaps, gap_model=gm, mode=mode)
        assert s[0] == pytest.approx(best, abs=1e-9), (trial, m, n)
        theta = np.concatenate([mat.ravel(), gaps])
        assert float(g[0] @ theta) == pytest.approx(s[0], abs=1e-9)
        if uniq is not None:
            np.testing.assert_array_equal(g[0], uniq)


@pytest.mark.parametrize("gm,mode", CASES)
def test_soft_vs_bruteforce(gm, mode):
    r = np.random.default_rng(hash(("soft", gm, mode)) % 2**32)
    na = 3
    for trial in range(25):
        m, n = int(r.integers(0, 4)), int(r.integers(0, 4))
        a = r.integers(0, na, size=m).astype(np.uint8)
        b = r.integers(0, na, size=n).astype(np.uint8)
        mat = r.normal(0, 1.5, size=(na, na))
        gaps = r.uniform(0.2, 2.0, size=4)
        T = float(r.choice([0.5, 1.0, 2.0]))
        s, g = oracle.soft_log(mat, gaps, a, [0, m], b, [0, n], gap_model=gm, mode=mode, temperature=T)
        es, eg = bf.soft(list(a), list(b), mat, gaps, T, gap_model=gm, mode=mode)
        assert s[0] == pytest.approx(es, rel=1e-11, abs=1e-11)
        np.testing.assert_allclose(g[0], eg, rtol=1e-9, atol=1e-11)


def test_planted_mistake_is_caught():
    """A plausible slip -- the open charged on extension moves too -- fails the identity check:
    the checks above are not restatements of the oracle's own formula."""
    mat = np.where(np.eye(4) > 0, 2.0, -1.0)
    a = np.array([0, 1, 2, 3, 0, 1], np.uint8)
    b = np.array([0, 1, 3, 0, 1], np.uint8)
    s, g = oracle.hard(mat, [2, 1, 2, 1], a, [0, 6], b, [0, 5], precision="float64")
    best, uniq = bf.hard(list(a), list(b), mat, [2, 1, 2, 1])
    assert s[0] == best
    wrong = g[0].copy()
    wrong[16] = wrong[17]  # opens := extends: the slip
    theta = np.concatenate([mat.ravel(), [2, 1, 2, 1]])
    assert float(g[0] @ theta) == s[0]
    assert float(wrong @ theta) != s[0] or np.array_equal(wrong, g[0])


@pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built (needs the reference sources)")
def test_config_sample_matches_reference_engine():
    """A seeded sample of the protein global affine configuration against the reference engine."""
    from oracle import ref
    w = wl.protein_global_hard(64)
    sa, sb = w.strings()
    for prec in ("float32", "float64"):
        rb = ref.RefBatch(sa, sb, w.alphabet, w.matrix, w.gaps, double=prec == "float64", n_threads=2)
        rb.score_and_grad()
        s, g = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision=prec)
        np.testing.assert_array_equal(s, rb.scores())
        mats, gaps = rb.grads()
        np.testing.assert_array_equal(g, np.concatenate([mats.reshape(64, -1), gaps], 1))


@pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built (needs the reference sources)")
def test_long_pairs_match_reference_engine():
    """Two 4,096-residue homolog pairs (the long-pair workload's shape) against the reference
    engine at its default traceback: the oracle's full pointer table gives the same path."""
    from oracle import ref
    w = wl.long_pairs(2, 4096)
    sa, sb = w.strings()
    for prec in ("float32", "float64"):
        rb = ref.RefBatch(sa, sb, w.alphabet, w.matrix, w.gaps, double=prec == "float64", n_threads=2)
        rb.score_and_grad()
        s, g = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision=prec)
        np.testing.assert_array_equal(s, rb.scores())
        mats, gaps = rb.grads()
        np.testing.assert_array_equal(g, np.concatenate([mats.reshape(2, -1), gaps], 1))


TRACKED = [p for p in GOLDEN if "_tracks_" in p and "/soft_" not in p]
SOFT_TRACKED = [p for p in GOLDEN if "/soft_tracks_" in p]


@pytest.mark.parametrize("path", TRACKED, ids=lambda p: os.path.basename(p))
def test_tracked_matches_reference_golden(path):
    """Tracked batches (K = 3 sets, segment tracks, explicit or default gap tracks): the
    oracle's restatement against the reference engine, bit-exact, scores and every count."""
    z = np.load(path)
    gd = str(z["grad_mode"])
    gt = z["gap_tracks"] if "gap_tracks" in z else None
    s, g = oracle.hard_tracked(z["matrices"], z["gaps_k"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"],
                               z["tracks"], gt, gap_model=str(z["gap_model"]), mode=str(z["mode"]),
                               precision=str(z["precision"]), with_grad=gd == "hard")
    np.testing.assert_array_equal(s, z["scores"])
    if gd == "hard":
        np.testing.assert_array_equal(g, z["grads"])
        np.testing.assert_array_equal(oracle.weighted_sum(g), z["compute_grad"])
        np.testing.assert_array_equal(oracle.weighted_sum(g, z["weights"]), z["weighted_grad"])


@pytest.mark.parametrize("gm", ["affine", "linear"])
@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
def test_tracks_with_equal_sets_reduce_to_untracked(gm, mode):
    """K copies of one parameter set under any tracks: the same scores as the untracked
    problem, and per-set counts that sum to the untracked counts (the reference's invariance,
    aligner.hpp:1820-1826: the run-wise column-0 border equals the closed form)."""
    w = wl.random_pairs(f"trk_equal_{gm}_{mode}", 40, wl.AA, (0, 70), (0, 90), gap_model=gm, mode=mode,
                        related=0.5, gaps=(10.5, 0.75, 9.25, 1.5) if gm == "affine" else (0, 0.75, 0, 1.5))
    K, nn = 3, 400
    r = wl.rng(w.name, "tracks")
    tracks = r.integers(0, K, size=int(w.off_a[-1])).astype(np.uint16)
    gtr = r.integers(0, K, size=int(w.off_a[-1]) + w.n_pairs).astype(np.uint16)
    for prec in ("float32", "float64"):
        s0, g0 = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, gap_model=gm, mode=mode,
                             precision=prec)
        s1, g1 = oracle.hard_tracked(np.stack([w.matrix] * K), np.stack([w.gaps] * K), w.codes_a, w.off_a,
                                     w.codes_b, w.off_b, tracks, gtr, gap_model=gm, mode=mode, precision=prec)
        np.testing.assert_array_equal(s1, s0)
        mats = g1[:, :K * nn].reshape(-1, K, nn).sum(axis=1)
        gaps = g1[:, K * nn:].reshape(-1, K, 4).sum(axis=1)
        np.testing.assert_array_equal(np.concatenate([mats, gaps], 1), g0)
        # one set, default tracks: the untracked problem exactly
        s2, g2 = oracle.hard_tracked(w.matrix[None], w.gaps[None], w.codes_a, w.off_a, w.codes_b, w.off_b,
                                     np.zeros(int(w.off_a[-1]), np.uint16), gap_model=gm, mode=mode, precision=prec)
        np.testing.assert_array_equal(s2, s0)
        np.testing.assert_array_equal(g2, g0)


def test_tracked_score_is_counts_times_params():
    """Integral parameters: each pair's score equals its K-set counts times the K sets, so a
    count attributed to the wrong set (or the wrong gap field) changes the sum."""
    w = wl.random_pairs("trk_identity", 40, wl.AA, (1, 70), (1, 90), related=0.5)
    K = 3
    r = wl.rng(w.name, "params")
    mats = np.stack([wl.random_matrix(r) for _ in range(K)])
    gaps = np.array([[10.0, 1.0, 9.0, 2.0], [8.0, 2.0, 11.0, 1.0], [12.0, 3.0, 6.0, 2.0]])
    tracks = r.integers(0, K, size=int(w.off_a[-1])).astype(np.uint16)
    s, g = oracle.hard_tracked(mats, gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, tracks)
    theta = np.concatenate([mats.ravel(), gaps.ravel()])
    np.testing.assert_array_equal(g @ theta, s)


@pytest.mark.parametrize("path", SOFT_TRACKED, ids=lambda p: os.path.basename(p))
def test_soft_tracked_matches_reference_golden(path):
    """Tracked soft batches (K = 3 sets, segment tracks, explicit or default gap tracks): the
    oracle's log-space restatement against the reference engine's float64 pass within 1e-9,
    and its float32 scaled pass within 1e-5 (norm-relative per pair)."""
    z = np.load(path)
    gt = z["gap_tracks"] if "gap_tracks" in z else None
    s, g = oracle.soft_log_tracked(z["matrices"], z["gaps_k"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"],
                                   z["tracks"], gt, gap_model=str(z["gap_model"]), mode=str(z["mode"]))
    tol = 1e-9 if str(z["precision"]) == "float64" else 1e-5
    assert np.max(np.abs(s - z["scores"]) / np.maximum(1.0, np.abs(z["scores"]))) <= tol
    scale = np.maximum(np.max(np.abs(z["grads"]), axis=1), 1e-300)
    assert np.max(np.max(np.abs(g - z["grads"]), axis=1) / scale) <= tol


@pytest.mark.parametrize("gm", ["affine", "linear"])
@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
def test_soft_tracks_with_equal_sets_reduce_to_untracked(gm, mode):
    """K copies of one set under any tracks: the untracked soft scores, and per-set counts that
    sum to the untracked counts (to double rounding)."""
    w = wl.random_pairs(f"soft_trk_equal_{gm}_{mode}", 30, wl.AA, (0, 60), (0, 80), gap_model=gm, mode=mode,
                        related=0.5, gaps=(10.5, 0.75, 9.25, 1.5) if gm == "affine" else (0, 0.75, 0, 1.5))
    K, nn = 3, 400
    r = wl.rng(w.name, "tracks")
    tracks = r.integers(0, K, size=int(w.off_a[-1])).astype(np.uint16)
    gtr = r.integers(0, K, size=int(w.off_a[-1]) + w.n_pairs).astype(np.uint16)
    s0, g0 = oracle.soft_log(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, gap_model=gm, mode=mode)
    s1, g1 = oracle.soft_log_tracked(np.stack([w.matrix] * K), np.stack([w.gaps] * K), w.codes_a, w.off_a,
                                     w.codes_b, w.off_b, tracks, gtr, gap_model=gm, mode=mode)
    np.testing.assert_allclose(s1, s0, rtol=1e-12, atol=1e-12)
    mats = g1[:, :K * nn].reshape(-1, K, nn).sum(axis=1)
    gaps = g1[:, K * nn:].reshape(-1, K, 4).sum(axis=1)
    np.testing.assert_allclose(np.concatenate([mats, gaps], 1), g0, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("gm", ["affine", "linear"])
@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
def test_soft_tracked_counts_are_score_derivatives(gm, mode):
    """Each tracked soft count is the derivative of T log Z with respect to its parameter of its
    set (matches +, gap penalties - the expected gaps): central differences on every gap field
    of every set and on matrix entries of every set, so a count routed to the wrong set, row or
    field fails."""
    r = wl.rng(f"soft_trk_fd_{gm}_{mode}", "all")
    K, na = 3, 4
    mats = np.stack([np.where(np.eye(na) > 0, 2.0, -1.0) + r.normal(0, 0.3, (na, na)) for _ in range(K)])
    gaps = np.array([[2.0, 1.0, 3.0, 0.5], [1.0, 0.75, 2.5, 1.0], [3.0, 0.5, 1.5, 1.25]])
    if gm == "linear":
        gaps = gaps * np.array([0.0, 1.0, 0.0, 1.0])
    A = [r.integers(0, na, size=int(r.integers(1, 9))).astype(np.uint8) for _ in range(3)]
    B = [r.integers(0, na, size=int(r.integers(1, 9))).astype(np.uint8) for _ in range(3)]
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    tracks = r.integers(0, K, size=int(oa[-1])).astype(np.uint16)
    gtr = r.integers(0, K, size=int(oa[-1]) + 3).astype(np.uint16)
    T = 1.7
    s, g = oracle.soft_log_tracked(mats, gaps, ca, oa, cb, ob, tracks, gtr, gap_model=gm, mode=mode, temperature=T)
    theta = np.concatenate([mats.ravel(), gaps.ravel()])
    fields = list(range(K * na * na, len(theta))) + list(r.choice(K * na * na, size=12, replace=False))
    h = 1e-5
    for f in fields:
        if gm == "linear" and f >= K * na * na and (f - K * na * na) % 2 == 0:
            continue   # linear: no open fields
        tp, tm = theta.copy(), theta.copy()
        tp[f] += h
        tm[f] -= h
        sp, _ = oracle.soft_log_tracked(tp[:K * na * na].reshape(K, na, na), tp[K * na * na:].reshape(K, 4), ca, oa,
                                        cb, ob, tracks, gtr, gap_model=gm, mode=mode, temperature=T)
        sm, _ = oracle.soft_log_tracked(tm[:K * na * na].reshape(K, na, na), tm[K * na * na:].reshape(K, 4), ca, oa,
                                        cb, ob, tracks, gtr, gap_model=gm, mode=mode, temperature=T)
        # score = T log Z(theta / T): d score / d theta_f = the expected count (gaps: minus it)
        fd = (sp - sm) / (2 * h)
        np.testing.assert_allclose(g[:, f], fd, rtol=1e-6, atol=1e-7)
