"""The same inputs give bit-identical results from run to run, on the same batch and on a fresh one.

The soft passes sum their histograms in a fixed order (pair order within a block, then block
order), so a result that changes between runs is a race or a register left stale across a
divergent region, whether or not it stays inside the parity tolerance.  The cases are the ones
such faults hit first: Local mode, pairs whose last column panel leaves lanes idle (513 and
2,049 columns), tracked float64 Local tiles run whole (plain fill), and the banded pass.
"""
import numpy as np
import pytest

import nwgrad_amd as nw
from nwgrad_amd import workloads as wl
from tests.gpu_util import batch, grads_rows, params_of, say

pytestmark = pytest.mark.gpu

RUNS = 3


def _pairs(name, alphabet_size, lens_b):
    r = wl.rng(name, "seqs")
    B = [r.integers(0, alphabet_size, size=n).astype(np.uint8) for n in lens_b]
    A = [wl.mutate(r, b_, alphabet_size, 0.2, 0.02, 2.0) for b_ in B]
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    return r, ca, oa, cb, ob


def _same(runs, label):
    s0, g0 = runs[0]
    for k, (s, g) in enumerate(runs[1:], 1):
        ds = int(np.count_nonzero(s != s0))
        dg = int(np.count_nonzero(np.any(g != g0, axis=1)))
        say(f"  {label} run {k}: {ds} scores, {dg} gradient rows differ from run 0")
        np.testing.assert_array_equal(s, s0)
        np.testing.assert_array_equal(g, g0)


LENS_B = [40, 300, 513, 700, 1100, 2049, 90, 1, 2, 64, 65]


@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("fill", ["interpair", "striped"])
def test_soft_repeat(gpu, mode, precision, fill):
    r, ca, oa, cb, ob = _pairs(f"repeat_soft_{mode}", 20, LENS_B)
    mat = wl.random_matrix(r)
    gaps = np.array([10.0, 1.0, 9.0, 1.5])
    say("case soft repeat", mode, precision, fill)
    b = batch(wl.AA, mat, gaps, ca, oa, cb, ob, precision=precision, mode=mode, grad_mode="soft", fill=fill)
    runs = []
    for _ in range(RUNS):
        b.score_and_grad()
        runs.append((b.scores(), grads_rows(b)))
    fresh = batch(wl.AA, mat, gaps, ca, oa, cb, ob, precision=precision, mode=mode, grad_mode="soft", fill=fill)
    fresh.score_and_grad()
    runs.append((fresh.scores(), grads_rows(fresh)))
    _same(runs, "soft")


@pytest.mark.parametrize("mode", ["local", "global"])
@pytest.mark.parametrize("fill", ["interpair", "striped"])
def test_soft_tracked_repeat(gpu, mode, fill):
    r, ca, oa, cb, ob = _pairs(f"repeat_soft_trk_{mode}", 20, LENS_B)
    K = 4
    mats = np.stack([wl.random_matrix(r) for _ in range(K)])
    gaps = np.array([[10.5, 0.75, 9.25, 1.5], [8.0, 1.25, 11.0, 0.5], [12.0, 2.0, 6.5, 1.5], [9.0, 1.0, 9.0, 1.0]])
    tracks = np.repeat(r.integers(0, K, size=int(oa[-1]) // 7 + 1), 7)[:int(oa[-1])].astype(np.uint16)
    gtr = r.integers(0, K, size=int(oa[-1]) + len(oa) - 1).astype(np.uint16)
    params = nw.AlignParams.from_arrays(mats, gaps, wl.AA)
    say("case soft tracked repeat float64", mode, fill)
    runs = []
    for _ in range(2):   # two batches, two passes each
        b = nw.SeqPairBatchDouble.from_codes(ca, oa, cb, ob, params, gap_model="affine", mode=mode,
                                             grad_mode="soft", tracks=tracks, gap_tracks=gtr)
        b.fill = fill
        for _ in range(2):
            b.score_and_grad()
            mats_g, gaps_g = b.grads()
            runs.append((b.scores(), np.concatenate([mats_g.reshape(len(b), -1), gaps_g.reshape(len(b), -1)], axis=1)))
    _same(runs, "soft tracked")


@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
def test_banded_soft_repeat(gpu, mode):
    r, ca, oa, cb, ob = _pairs(f"repeat_band_{mode}", 20, [40, 300, 513, 700, 1100, 90, 65])
    mat = wl.random_matrix(r)
    gaps = np.array([11.0, 1.0, 11.0, 1.0])
    b = batch(wl.AA, mat, gaps, ca, oa, cb, ob, mode=mode, grad_mode="soft")
    for bw in (24, 2000):
        say("case banded soft repeat", mode, "band", bw)
        runs = []
        for _ in range(RUNS):
            b.set_params(params_of(wl.AA, mat, gaps))
            b.score_and_grad()   # the same guides each time
            b.banded_grad(bw)
            runs.append((b.scores(), grads_rows(b)))
        _same(runs, f"band {bw}")


@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
@pytest.mark.parametrize("gap_model", ["affine", "linear"])
def test_hard_repeat(gpu, mode, gap_model):
    r, ca, oa, cb, ob = _pairs(f"repeat_hard_{mode}_{gap_model}", 20, LENS_B + [4096, 5000])
    mat = wl.random_matrix(r)
    gaps = np.array([10.0, 1.0, 9.0, 1.5])
    say("case hard repeat", mode, gap_model)
    outs = []
    for _ in range(2):
        b = batch(wl.AA, mat, gaps, ca, oa, cb, ob, gap_model=gap_model, mode=mode, grad_mode="hard")
        for _ in range(2):
            b.score_and_grad(keep_paths=True)
            guides = np.concatenate([np.asarray(b[i].guide_j, np.int32) for i in range(len(b))])
            outs.append((b.scores(), grads_rows(b), guides, b.cigars()))
    for s, g, gd, cg in outs[1:]:
        np.testing.assert_array_equal(s, outs[0][0])
        np.testing.assert_array_equal(g, outs[0][1])
        np.testing.assert_array_equal(gd, outs[0][2])
        assert cg == outs[0][3]
