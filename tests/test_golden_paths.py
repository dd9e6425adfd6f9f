"""CPU checks of the path fixtures the reference engine wrote (tests/golden/paths_*.npz): each
stored path is consistent with its pair, its gradient row and its guide, so the fixtures pin the
GPU's guides and paths to the reference's own definitions, and these definitions are the ones the
GPU tests decode (type 0 both advance, 1 a gap in B, 2 a gap in A; coordinates at every change)."""
import glob
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(__file__), "golden")
PATHS = sorted(glob.glob(os.path.join(HERE, "paths_*.npz")))


def _seqs(z):
    al = str(z["alphabet"])
    out = []
    for codes, off in ((z["codes_a"], z["off_a"]), (z["codes_b"], z["off_b"])):
        out.append(["".join(al[c] for c in codes[off[i]:off[i + 1]]) for i in range(len(off) - 1)])
    return out


@pytest.mark.parametrize("path", PATHS, ids=lambda p: os.path.basename(p)[:-4])
def test_path_fixture_consistency(path):
    z = np.load(path)
    sa, sb = _seqs(z)
    al = str(z["alphabet"])
    na = len(al)
    affine = str(z["gap_model"]) == "affine"
    mode = str(z["mode"])
    co = z["coords_off"]
    goff = 0
    for i in range(len(sa)):
        a, b = str(z["aligned_a"][i]), str(z["aligned_b"][i])
        coords = z["coords"][:, co[i]:co[i + 1]]
        x0, y0 = int(coords[0, 0]), int(coords[1, 0])
        x1, y1 = int(coords[0, -1]), int(coords[1, -1])
        # the aligned strings spell the aligned parts of the sequences
        assert a.replace("-", "") == sa[i][x0:x1] and b.replace("-", "") == sb[i][y0:y1]
        if mode == "global":
            assert (x0, y0, x1, y1) == (0, 0, len(sa[i]), len(sb[i]))
        if mode == "glocal":
            assert x0 == 0 and x1 == len(sa[i])
        # counts along the path equal the gradient row
        cnt = np.zeros(na * na + 4)
        prev = "M"
        cig = []
        for ca, cb in zip(a, b):
            op = "M" if ca != "-" and cb != "-" else ("I" if cb == "-" else "D")
            cig.append(op)
            if op == "M":
                cnt[al.index(ca) * na + al.index(cb)] += 1
            elif op == "I":   # a gap in B (A advances): extend b, open b
                cnt[na * na + 3] -= 1
                if affine and prev != "I":
                    cnt[na * na + 2] -= 1
            else:
                cnt[na * na + 1] -= 1
                if affine and prev != "D":
                    cnt[na * na + 0] -= 1
            prev = op
        np.testing.assert_array_equal(cnt, z["grads"][i])
        # the CIGAR is the run-length form of the columns
        runs, k = [], 0
        while k < len(cig):
            e = k
            while e < len(cig) and cig[e] == cig[k]:
                e += 1
            runs.append(f"{e - k}{cig[k]}")
            k = e
        assert "".join(runs) == str(z["cigars"][i])
        # the guide: m + 1 non-decreasing columns in [0, n], ending at the path's end row
        m, n = len(sa[i]), len(sb[i])
        g = z["guides"][goff:goff + m + 1]
        goff += m + 1
        assert np.all(np.diff(g) >= 0) and g.min() >= 0 and g.max() <= n
        assert g[x1] == y1 and (mode != "global" or g[m] == n)
    assert goff == len(z["guides"])


BANDSCORE = sorted(glob.glob(os.path.join(HERE, "bandscore_*.npz")))


@pytest.mark.skipif(not BANDSCORE, reason="no score-only banded fixtures")
@pytest.mark.parametrize("path", BANDSCORE, ids=lambda p: os.path.basename(p)[:-4])
def test_bandscore_fixture_is_the_hard_banded_pass(path):
    """A score-only batch's banded pass is the hard banded pass's scores, guides and paths around
    the same guides: the reference engine itself agrees (the GPU path runs it that way)."""
    import oracle
    if not oracle.have_ref():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    from oracle import ref
    z = np.load(path)
    sa, sb = _seqs(z)
    for bw in z["bands"]:
        b = ref.RefBatch(sa, sb, str(z["alphabet"]), z["matrix"], z["gaps"], gap_model=str(z["gap_model"]),
                         mode=str(z["mode"]), grad_mode="hard", n_threads=1)
        b.score_and_grad()
        b.set_params(z["matrix1"], z["gaps1"])
        assert b.banded_grad(int(bw), False) == float(z[f"band{bw}_total"])
        np.testing.assert_array_equal(b.scores(), z[f"band{bw}_scores"])
        np.testing.assert_array_equal(b.guides(z["off_a"]), z[f"band{bw}_guides"])


BANDSOFT = sorted(glob.glob(os.path.join(HERE, "bandsoft_*.npz")))


@pytest.mark.skipif(not BANDSOFT, reason="no soft banded fixtures")
@pytest.mark.parametrize("path", BANDSOFT, ids=lambda p: os.path.basename(p)[:-4])
def test_bandsoft_guides_are_the_hard_banded_pass(path):
    """A soft batch's banded pass leaves the guides and paths of the banded Viterbi around the
    same (eager) guides: the reference engine's hard banded pass gives them bit for bit."""
    import oracle
    if not oracle.have_ref():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    from oracle import ref
    z = np.load(path)
    sa, sb = _seqs(z)
    for bw in z["bands"]:
        b = ref.RefBatch(sa, sb, str(z["alphabet"]), z["matrix"], z["gaps"], gap_model=str(z["gap_model"]),
                         mode=str(z["mode"]), grad_mode="hard", n_threads=1)
        b.score_and_grad()
        b.set_params(z["matrix1"], z["gaps1"])
        b.banded_grad(int(bw), False)
        np.testing.assert_array_equal(b.guides(z["off_a"]), z[f"band{bw}_guides"])


@pytest.mark.skipif(not BANDSOFT, reason="no soft banded fixtures")
@pytest.mark.parametrize("path", BANDSOFT, ids=lambda p: os.path.basename(p)[:-4])
def test_bandsoft_fixture_against_brute_force(path):
    """The reference's banded soft results against an independent brute-force forward-backward
    over the band (tests/banded_soft_model.py), on the fixture's smaller pairs; the model itself
    is checked against the oracle's log-space full pass first."""
    import oracle
    from tests.banded_soft_model import banded_soft
    z = np.load(path)
    gm, mode = str(z["gap_model"]), str(z["mode"])
    oa, ob = z["off_a"], z["off_b"]
    sizes = (oa[1:] - oa[:-1]) * (ob[1:] - ob[:-1])
    pairs = [int(p) for p in np.argsort(sizes, kind="stable")[:4]]
    for p in pairs:
        a, b = z["codes_a"][oa[p]:oa[p + 1]], z["codes_b"][ob[p]:ob[p + 1]]
        lz, cnt = banded_soft(a, b, z["matrix"], z["gaps"], mode, None, None, gap_model=gm)
        s_o, g_o = oracle.soft_log(z["matrix"], z["gaps"], a, np.array([0, len(a)]), b, np.array([0, len(b)]),
                                   gap_model=gm, mode=mode)
        assert abs(lz - s_o[0]) <= 1e-9 * max(1.0, abs(s_o[0]))
        np.testing.assert_allclose(cnt, g_o[0], rtol=1e-8, atol=1e-9)
        g0 = z["guides0"][oa[p] + p:oa[p + 1] + p + 1]
        for bw in z["bands"]:
            lz, cnt = banded_soft(a, b, z["matrix1"], z["gaps1"], mode, g0, int(bw), gap_model=gm)
            want_s, want_g = float(z[f"band{bw}_scores"][p]), z[f"band{bw}_grads"][p]
            assert abs(lz - want_s) <= 2e-6 * max(1.0, abs(want_s))   # the reference runs float32
            scale = np.max(np.abs(want_g))
            assert np.all(np.abs(cnt - want_g) <= 1e-5 * (np.abs(want_g) + 1e-7 * scale) + 1e-12), (p, bw)
