"""Guides, stored paths, banded re-alignment and streaming align() on the GPU, against the
fixtures the reference engine wrote (tests/golden/make_golden.py paths) and against properties
that hold at any size.

* score_and_grad(keep_paths=True): scores and gradients as before, every pair's guide_j
  (aligner.hpp:3155-3160, 2734-2749) and its stored path rendered as aligned strings,
  coordinates and CIGARs (batch_engine.hpp:513-590), all bit for bit.
* banded_grad(bw) after set_params (batch_engine.hpp:633-653): scores, gradients, the new guides
  and paths, bit for bit, for narrow and wide bands, short and long (column-panel) pairs.
* Soft batches' guides: the eager guide, the hard fill under the params scored.
* Streaming align() (batch_engine.hpp:748-1273): scores and the summed gradient.
"""
import glob
import os

import numpy as np
import pytest

import nwgrad_amd as nw
from nwgrad_amd import workloads as wl
from tests.gpu_util import assert_soft_close, batch, grads_rows, params_of, say

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(__file__), "golden")
PATHS = sorted(glob.glob(os.path.join(HERE, "paths_*.npz")))
SOFTGUIDE = sorted(glob.glob(os.path.join(HERE, "softguide_*.npz")))
ALIGN = sorted(glob.glob(os.path.join(HERE, "align_*.npz")))
BANDSCORE = sorted(glob.glob(os.path.join(HERE, "bandscore_*.npz")))
BANDSOFT = sorted(glob.glob(os.path.join(HERE, "bandsoft_*.npz")))


def _guides(b):
    return np.concatenate([np.asarray(b[i].guide_j, np.int32) for i in range(len(b))])


def _strings(z):
    al = np.frombuffer(str(z["alphabet"]).encode(), np.uint8)
    sa = [al[z["codes_a"][z["off_a"][i]:z["off_a"][i + 1]]].tobytes().decode() for i in range(len(z["off_a"]) - 1)]
    sb = [al[z["codes_b"][z["off_b"][i]:z["off_b"][i + 1]]].tobytes().decode() for i in range(len(z["off_b"]) - 1)]
    return sa, sb


def _batch(z):
    return batch(str(z["alphabet"]), z["matrix"], z["gaps"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"],
                 precision=str(z["precision"]), gap_model=str(z["gap_model"]), mode=str(z["mode"]),
                 grad_mode=str(z["grad_mode"]))


@pytest.mark.parametrize("path", PATHS, ids=lambda p: os.path.basename(p)[:-4])
def test_paths_golden(gpu, path):
    z = np.load(path)
    say("case", os.path.basename(path))
    b = _batch(z)
    total = b.score_and_grad(keep_paths=True)
    assert total == float(z["total"])
    np.testing.assert_array_equal(b.scores(), z["scores"])
    np.testing.assert_array_equal(grads_rows(b), z["grads"])
    np.testing.assert_array_equal(_guides(b), z["guides"])
    assert b.cigars() == list(z["cigars"])
    assert b.cigars(extended=True, soft_clip=True) == list(z["cigars_ext"])
    co = z["coords_off"]
    for i in range(len(b)):
        v = b[i]
        assert v.stored_path_valid and v.path_valid
        assert v.aligned() == (str(z["aligned_a"][i]), str(z["aligned_b"][i]))
        np.testing.assert_array_equal(v.coordinates(), z["coords"][:, co[i]:co[i + 1]])
    # banded re-alignment after a parameter step, from the guides of the full pass
    p1 = params_of(str(z["alphabet"]), z["matrix1"], z["gaps1"])
    for bw in z["bands"]:
        b = _batch(z)
        b.score_and_grad()
        b.set_params(p1)
        assert not b[0].stored_path_valid and b[0].guide_j is not None   # paths cleared, guides kept
        t = b.banded_grad(int(bw), keep_paths=True)
        say(f"  band {bw}: total {t}")
        assert t == float(z[f"band{bw}_total"])
        np.testing.assert_array_equal(b.scores(), z[f"band{bw}_scores"])
        np.testing.assert_array_equal(grads_rows(b), z[f"band{bw}_grads"])
        np.testing.assert_array_equal(_guides(b), z[f"band{bw}_guides"])
        assert b.cigars() == list(z[f"band{bw}_cigars"])


@pytest.mark.parametrize("path", BANDSCORE, ids=lambda p: os.path.basename(p)[:-4])
def test_banded_score_only_golden(gpu, path):
    """banded_grad of a score-only batch: the banded Viterbi's scores, new guides and paths."""
    z = np.load(path)
    say("case", os.path.basename(path))
    p1 = params_of(str(z["alphabet"]), z["matrix1"], z["gaps1"])
    for bw in z["bands"]:
        b = _batch(z)   # grad_mode "none"
        b.score_and_grad()
        b.set_params(p1)
        with pytest.raises(RuntimeError):
            b.scores()   # set_params cleared them
        t = b.banded_grad(int(bw), keep_paths=True)
        assert t == float(z[f"band{bw}_total"])
        np.testing.assert_array_equal(b.scores(), z[f"band{bw}_scores"])
        assert [b[i].score for i in range(len(b))] == list(z[f"band{bw}_scores"])
        np.testing.assert_array_equal(_guides(b), z[f"band{bw}_guides"])
        assert b.cigars() == list(z[f"band{bw}_cigars"])
        with pytest.raises(RuntimeError, match="grad_mode"):
            b.grads()
        # a full pass afterwards serves its own scores again
        b.score_and_grad()
        assert b.scores()[0] == b[0].score


@pytest.mark.parametrize("path", BANDSOFT, ids=lambda p: os.path.basename(p)[:-4])
def test_banded_soft_golden(gpu, path):
    """banded_grad of a soft batch: the banded forward-backward's scores and expected counts
    around the eager guides (within the soft tolerance of the reference's float32), the banded
    Viterbi's new guides and paths (bit for bit); streaming align() with a band."""
    z = np.load(path)
    say("case", os.path.basename(path))
    p1 = params_of(str(z["alphabet"]), z["matrix1"], z["gaps1"])
    for bw in z["bands"]:
        b = _batch(z)   # grad_mode "soft"
        b.soft_impl = "scaled_or_log"
        b.score_and_grad()
        b.set_params(p1)
        t = b.banded_grad(int(bw), keep_paths=True)
        assert_soft_close(b.scores(), grads_rows(b), z[f"band{bw}_scores"], z[f"band{bw}_grads"],
                          label=f"band {bw}: ")
        assert abs(t - float(z[f"band{bw}_total"])) <= 1e-5 * max(1.0, abs(float(z[f"band{bw}_total"])))
        np.testing.assert_array_equal(_guides(b), z[f"band{bw}_guides"])
        assert b.cigars() == list(z[f"band{bw}_cigars"])
    sa, sb = _strings(z)
    host = nw.SeqPairBatch(sa[:2], sb[:2], params_of(str(z["alphabet"]), z["matrix"], z["gaps"]),
                           gap_model=str(z["gap_model"]), mode=str(z["mode"]), grad_mode="soft")
    r = host.align(sa, sb, p1, band=8)
    d = r.grad.to_dict()
    g = np.concatenate([d["matrix"].ravel(), [d["gap_open_a"], d["gap_extend_a"], d["gap_open_b"], d["gap_extend_b"]]])
    assert_soft_close(r.scores, g[None, :], z["align8_scores"], z["align8_grad"][None, :], label="align band 8: ")


@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
def test_banded_soft_properties(gpu, mode):
    """At any size: a band as wide as B is the full pass; a narrower band sums over fewer paths,
    so its log Z never exceeds the full one's and grows with the band."""
    # homolog pairs (B a mutated copy of A): unrelated pairs of very different lengths do not fit
    # the scaled pass at narrow bands, in the reference either (its scaled pass refuses 13 of 800
    # such pairs at band 2, in double)
    w = wl.random_pairs(f"banded_soft_props_{mode}", 800, wl.AA, (1, 600), (1, 600), mode=mode, related=1.0,
                        gaps=(11.0, 1.0, 11.0, 1.0))
    b = batch(w.alphabet, w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, mode=mode, grad_mode="soft",
              precision="float64")
    b.score_and_grad()
    s0, g0 = b.scores(), grads_rows(b)
    wide = int(np.max(np.diff(w.off_b))) + 1
    b.banded_grad(wide)
    assert_soft_close(b.scores(), grads_rows(b), s0, g0, rtol=1e-12, floor=1e-12, label="wide band vs full: ")
    prev = None
    for bw in (2, 8, 32):
        b.set_params(params_of(w.alphabet, w.matrix, w.gaps))
        b.score_and_grad()   # the guides of the full pass, each time
        b.banded_grad(bw)
        s = b.scores()
        assert np.all(s <= s0 + 1e-9 * np.maximum(1.0, np.abs(s0)))
        if prev is not None:
            assert np.all(s >= prev - 1e-9 * np.maximum(1.0, np.abs(prev)))
        prev = s


@pytest.mark.parametrize("path", SOFTGUIDE, ids=lambda p: os.path.basename(p)[:-4])
def test_soft_guides_golden(gpu, path):
    z = np.load(path)
    b = batch(str(z["alphabet"]), z["matrix"], z["gaps"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"],
              mode=str(z["mode"]), grad_mode="soft")
    assert b[0].guide_j is None   # no pass yet
    b.score_and_grad()
    np.testing.assert_array_equal(_guides(b), z["guides"])
    # a score-only batch takes its guides from the same hard fill
    s = batch(str(z["alphabet"]), z["matrix"], z["gaps"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"],
              mode=str(z["mode"]), grad_mode="none")
    s.score_and_grad(keep_paths=True)
    np.testing.assert_array_equal(_guides(s), z["guides"])
    assert s[3].aligned()[0].replace("-", "") == s[3].seq_a


@pytest.mark.parametrize("path", ALIGN, ids=lambda p: os.path.basename(p)[:-4])
def test_align_golden(gpu, path):
    z = np.load(path)
    sa, sb = _strings(z)
    p1 = params_of(str(z["alphabet"]), z["matrix1"], z["gaps1"])
    for gd in ("hard", "none", "soft"):
        if f"{gd}_scores" not in z.files:
            continue
        host = nw.SeqPairBatch(sa[:2], sb[:2], params_of(str(z["alphabet"]), z["matrix"], z["gaps"]),
                               gap_model=str(z["gap_model"]), mode=str(z["mode"]), grad_mode=gd)
        r = host.align(sa, sb, p1)
        d = r.grad.to_dict()
        g = np.concatenate([d["matrix"].ravel(), [d["gap_open_a"], d["gap_extend_a"], d["gap_open_b"],
                                                  d["gap_extend_b"]]])
        if gd == "soft":
            assert_soft_close(r.scores, g[None, :], z["soft_scores"], z["soft_grad"][None, :], label="align soft: ")
        else:
            np.testing.assert_array_equal(r.scores, z[f"{gd}_scores"])
            np.testing.assert_array_equal(g, z[f"{gd}_grad"])
        assert len(host) == 2   # nothing was added to the batch


@pytest.mark.parametrize("gm", ["affine", "linear"])
@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
def test_banded_properties(gpu, gm, mode):
    """At any size: a band as wide as B reproduces the full pass (scores, gradients, guides); a
    narrow band never scores above the full optimum; the stored path's own counts equal the
    gradient."""
    w = wl.random_pairs(f"banded_props_{gm}_{mode}", 3000, wl.AA, (0, 400), (0, 1500), gap_model=gm, mode=mode,
                        gaps=(10.5, 0.75, 9.25, 1.5) if gm == "affine" else (0.0, 1.25, 0.0, 1.75), related=0.6)
    b = batch(w.alphabet, w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, gap_model=gm, mode=mode)
    b.score_and_grad(keep_paths=True)
    s0, g0, gj0 = b.scores(), grads_rows(b), _guides(b)
    # the path's counts: matches per letter pair and gap runs, against the gradient rows
    na = 20
    for i in range(0, len(b), 97):
        x, y = b[i].coordinates()[:, 0]
        a, bb = b[i].aligned()
        cnt = np.zeros(na * na + 4)
        prev = "M"
        ca, cb = w.codes_a[w.off_a[i]:w.off_a[i + 1]], w.codes_b[w.off_b[i]:w.off_b[i + 1]]
        for ch_a, ch_b in zip(a, bb):
            if ch_a != "-" and ch_b != "-":
                cnt[int(ca[x]) * na + int(cb[y])] += 1
                x += 1
                y += 1
                prev = "M"
            elif ch_b == "-":   # a gap in B: extend b, open b unless the previous column was one
                cnt[na * na + 3] -= 1
                if prev != "X" and gm == "affine":
                    cnt[na * na + 2] -= 1
                x += 1
                prev = "X"
            else:
                cnt[na * na + 1] -= 1
                if prev != "Y" and gm == "affine":
                    cnt[na * na + 0] -= 1
                y += 1
                prev = "Y"
        np.testing.assert_array_equal(cnt, g0[i])
    wide = int(np.max(np.diff(w.off_b))) + 1
    b.banded_grad(wide)
    np.testing.assert_array_equal(b.scores(), s0)
    np.testing.assert_array_equal(grads_rows(b), g0)
    np.testing.assert_array_equal(_guides(b), gj0)
    b.banded_grad(1)
    assert np.all(b.scores() <= s0)   # a band restricts the paths: never above the full optimum


def test_paths_errors(gpu):
    w = wl.random_pairs("paths_errors", 8, wl.AA, (5, 30), (5, 30))
    b = batch(w.alphabet, w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b)
    with pytest.raises(RuntimeError, match="guide path"):
        b.banded_grad(4)
    b.score_and_grad()
    with pytest.raises(RuntimeError, match="stored path unavailable"):
        b.cigars()
    with pytest.raises(ValueError, match="bandwidth > 0"):
        b.banded_grad(0)
    b.score_and_grad(keep_paths=True)
    assert len(b.cigars()) == 8
    b.drop_paths()
    with pytest.raises(RuntimeError, match="stored path unavailable"):
        b[0].aligned()
    assert b[0].guide_j is not None and b[0].score is not None
    s = batch(w.alphabet, w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, grad_mode="soft")
    with pytest.raises(RuntimeError, match="guide path"):
        s.banded_grad(4)
    s.score_and_grad()
    s.banded_grad(4)   # soft batches band too
    K = 2
    pk = nw.AlignParams.stack([params_of(w.alphabet, w.matrix, w.gaps)] * K)
    t = nw.SeqPairBatch.from_codes(w.codes_a, w.off_a, w.codes_b, w.off_b, pk, grad_mode="soft",
                                   tracks=np.zeros(len(w.codes_a), np.uint16))
    t.score_and_grad()
    with pytest.raises(ValueError, match="tracked"):
        t.banded_grad(4)
