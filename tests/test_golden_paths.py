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
