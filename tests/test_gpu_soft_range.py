"""The soft pass's precision policy (the reference's, aligner.hpp:6554-6586): a pair whose float32
pass refuses (a weight or a row out of range, or lost mass beyond 2^-20) reruns in double; a
pair that double refuses too fails the call with ValueError (tests/Python/test_soft_scaled.py
out_of_range_case); an asynchronous (captured) step never applies a refused pair's float32
gradient, and finish() reports it.

The forcing inputs are the reference's: a diagonal score so large that exp(score / T) does not
fit the working format (100 nats: past float32, within double; 800 nats: past double).
"""
import numpy as np
import pytest

import oracle
from nwgrad_amd import workloads as wl
from tests.gpu_util import assert_soft_close, batch, grads_rows, say

pytestmark = pytest.mark.gpu

GAPS = np.array([2.0, 0.5, 2.0, 0.5])


def _matrix(diag, seed):
    r = wl.rng("soft_range", f"matrix{seed}")
    m = np.clip(r.normal(-1.0, 0.5, (20, 20)), -3.0, 1.0)
    np.fill_diagonal(m, diag)
    return m


def _mixed_pairs(name, n):
    """n pairs: even ones related (B a mutated A: the diagonal scores appear), odd ones over
    disjoint letters (A from 0..9, B from 10..19: no diagonal score is ever used)."""
    r = wl.rng(name, "seqs")
    A, B = [], []
    for k in range(n):
        m = int(r.integers(20, 160))
        if k % 2 == 0:
            a = r.integers(0, 20, size=m).astype(np.uint8)
            b = wl.mutate(r, a, 20, 0.3, 0.02, 2.0)
        else:
            a = r.integers(0, 10, size=m).astype(np.uint8)
            b = r.integers(10, 20, size=int(r.integers(20, 160))).astype(np.uint8)
        A.append(a)
        B.append(b)
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    return ca, oa, cb, ob


@pytest.mark.parametrize("mode", ["global", "local"])
def test_float32_refusal_reruns_in_double(gpu, mode):
    ca, oa, cb, ob = _mixed_pairs(f"soft_refuse_{mode}", 24)
    mat = _matrix(100.0, 1)
    say("case float32 refusal", mode)
    s_ref, g_ref = oracle.soft_log(mat, GAPS, ca, oa, cb, ob, mode=mode)
    for fill in ("interpair", "striped"):
        b = batch(wl.AA, mat, GAPS, ca, oa, cb, ob, mode=mode, grad_mode="soft", fill=fill)
        b.score_and_grad()
        say(f"  {fill}: {b.soft_reruns} of {len(b)} pairs rerun in double")
        assert b.soft_reruns == 12   # exactly the related pairs
        s, g = b.scores(), grads_rows(b)
        related = np.arange(0, 24, 2)
        # the reruns are the double pass's results
        assert_soft_close(s[related], g[related], s_ref[related], g_ref[related], rtol=1e-10, floor=1e-12,
                          label="reruns vs oracle: ")
        assert_soft_close(s, g, s_ref, g_ref, label="all pairs vs oracle: ")


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_double_refusal_raises(gpu, precision):
    """The reference's out_of_range_case: a local self-pair of 50 residues under a diagonal of
    800 nats; exp(800) does not fit a double."""
    r = wl.rng("soft_out_of_range", "seq")
    a = r.integers(0, 20, size=50).astype(np.uint8)
    ca, oa = wl._pack([a])
    mat = np.full((20, 20), -1.0)
    np.fill_diagonal(mat, 800.0)
    b = batch(wl.AA, mat, GAPS, ca, oa, ca, oa, precision=precision, mode="local", grad_mode="soft")
    assert b.soft_impl == "scaled"
    with pytest.raises(ValueError, match="scaled"):
        b.score_and_grad()
    with pytest.raises(RuntimeError):   # no valid results after the failure
        b.scores()


def test_soft_impl_values(gpu):
    ca, oa, cb, ob = _mixed_pairs("soft_impl", 4)
    b = batch(wl.AA, _matrix(3.0, 2), GAPS, ca, oa, cb, ob, grad_mode="soft")
    b.soft_impl = "scaled_or_log"
    assert b.soft_impl == "scaled_or_log"
    with pytest.raises(ValueError, match="log"):
        b.soft_impl = "log"
    with pytest.raises(ValueError):
        b.soft_impl = "nope"


def test_async_refusal_masked(gpu):
    """A step through the autograd function on a batch whose float32 pass refuses half its pairs:
    the backward's weighted sum leaves the refused rows out (it equals the oracle's sum over the
    accepted pairs), finish() raises naming the pass, reruns the refusals, and the batch's
    results are then complete."""
    import torch
    from nwgrad_amd import autograd as ag
    ca, oa, cb, ob = _mixed_pairs("soft_async_refuse", 16)
    mat = _matrix(100.0, 3)
    s_ref, g_ref = oracle.soft_log(mat, GAPS, ca, oa, cb, ob)
    b = batch(wl.AA, mat, GAPS, ca, oa, cb, ob, grad_mode="soft")
    theta = ag.params_vector(_params(mat)).requires_grad_(True)
    s = ag.soft_scores(theta, b)
    (g,) = torch.autograd.grad(s.sum(), theta)
    torch.cuda.synchronize()
    accepted = np.arange(1, 16, 2)
    want = oracle.weighted_sum(g_ref[accepted])
    got = g.detach().cpu().numpy()
    say(f"  masked step gradient vs oracle over the accepted pairs: {np.max(np.abs(got - want)):.3e}")
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-7 * np.max(np.abs(want)))
    refused, first, passes = b._soft_refusals(False)
    assert (refused, first, passes) == (8, 1, 1)
    with pytest.raises(ag.SoftRefusalError, match="pass 1"):
        ag.finish(b)
    assert b._soft_refusals(False)[:2] == (0, 0)   # finish() cleared the record
    assert_soft_close(b.scores(), grads_rows(b), s_ref, g_ref, label="after finish: ")


def _params(mat):
    import nwgrad_amd as nw
    return nw.AlignParams(mat, alphabet=wl.AA, gap_open_a=GAPS[0], gap_extend_a=GAPS[1], gap_open_b=GAPS[2],
                          gap_extend_b=GAPS[3])
