"""GPU parity of the hard and score-only paths: bit-exact against the reference's golden
vectors and against the oracle, through the C ABI (via the engine and its module)."""
import glob
import os

import numpy as np
import pytest

import oracle
from nwgrad_amd import workloads as wl
from tests.gpu_util import batch, from_workload, grads_rows, say

pytestmark = pytest.mark.gpu

GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if ("/hard_" in p or "/score_" in p) and "_tracks_" not in p)


def _golden_batch(z, fill):
    return batch(str(z["alphabet"]), z["matrix"], z["gaps"], z["codes_a"], z["off_a"], z["codes_b"], z["off_b"],
                 precision=str(z["precision"]), gap_model=str(z["gap_model"]), mode=str(z["mode"]),
                 grad_mode=str(z["grad_mode"]), fill=fill)


@pytest.mark.parametrize("fill", ["interpair", "striped"])
@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: os.path.basename(p)[:-4])
def test_golden_bit_exact(gpu, path, fill):
    z = np.load(path)
    say("case", os.path.basename(path), fill)
    b = _golden_batch(z, fill)
    total = b.score_and_grad()
    np.testing.assert_array_equal(b.scores(), z["scores"])
    assert total == float(z["total"])
    if str(z["grad_mode"]) == "hard":
        np.testing.assert_array_equal(grads_rows(b), z["grads"])
        cg = b.compute_grad().to_dict()
        flat = np.concatenate([cg["matrix"].ravel(), [cg["gap_open_a"], cg["gap_extend_a"], cg["gap_open_b"],
                                                       cg["gap_extend_b"]]])
        np.testing.assert_array_equal(flat, z["compute_grad"])
        wg = b.weighted_grad(np.ascontiguousarray(z["weights"])).to_dict()
        flat = np.concatenate([wg["matrix"].ravel(), [wg["gap_open_a"], wg["gap_extend_a"], wg["gap_open_b"],
                                                       wg["gap_extend_b"]]])
        np.testing.assert_array_equal(flat, z["weighted_grad"])


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_protein_global_hard_config(gpu, precision):
    """4,096 protein pairs, lengths 100-300, global affine 11/1: every score and count."""
    w = wl.protein_global_hard()
    say("case protein_global_hard", precision, w.n_pairs, "pairs", w.cells, "cells")
    s_ref, g_ref = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision=precision)
    for fill in ("interpair", "striped"):
        b = from_workload(w, precision=precision, fill=fill)
        total = b.score_and_grad()
        np.testing.assert_array_equal(b.scores(), s_ref)
        np.testing.assert_array_equal(grads_rows(b), g_ref)
        assert total == float(np.cumsum(s_ref)[-1])


def test_dna_local_short_config(gpu):
    """1,048,576 DNA pairs (half with a planted reverse complement), local linear: hard and
    score-only, persistent and per-bucket launches, every pair against the oracle."""
    w = wl.dna_local_short()
    say("case dna_local_short", w.n_pairs, "pairs", w.cells, "cells")
    s_ref, g_ref = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, gap_model=w.gap_model,
                               mode=w.mode)
    for fill in ("interpair", "striped"):
        b = from_workload(w, fill=fill)
        b.score_and_grad()
        np.testing.assert_array_equal(b.scores(), s_ref)
        np.testing.assert_array_equal(grads_rows(b), g_ref)
        np.testing.assert_array_equal(b.compute_grad().to_dict()["matrix"].ravel(),
                                      oracle.weighted_sum(g_ref)[:16])
        del b
        b = from_workload(w, fill=fill, grad_mode="none")
        b.score_and_grad()
        np.testing.assert_array_equal(b.scores(), s_ref)


@pytest.mark.parametrize("gm", ["affine", "linear"])
@pytest.mark.parametrize("mode", ["global", "local", "glocal"])
def test_ragged_lane_groups(gpu, gm, mode):
    """Every lane-group width (len(B) from 0 to 1024) and ragged len(A) in one batch."""
    r = wl.rng(f"ragged_{gm}_{mode}", "lens")
    nb = np.concatenate([np.arange(0, 40), r.integers(0, 1025, size=200), [1023, 1024, 16, 17, 32, 33]])
    na_ = r.integers(0, 300, size=len(nb))
    A = [r.integers(0, 20, size=int(x)).astype(np.uint8) for x in na_]
    B = [r.integers(0, 20, size=int(x)).astype(np.uint8) for x in nb]
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    mat = wl.random_matrix(r)
    gaps = np.array([4.5, 0.5, 3.25, 1.0])
    say("case ragged", gm, mode, len(nb), "pairs")
    for precision in ("float32", "float64"):
        s_ref, g_ref = oracle.hard(mat, gaps, ca, oa, cb, ob, gap_model=gm, mode=mode, precision=precision)
        b = batch(wl.AA, mat, gaps, ca, oa, cb, ob, precision=precision, gap_model=gm, mode=mode)
        b.score_and_grad()
        np.testing.assert_array_equal(b.scores(), s_ref)
        np.testing.assert_array_equal(grads_rows(b), g_ref)


def test_set_params_and_errors(gpu):
    w = wl.random_pairs("set_params", 50, "ACGT", (5, 30), (5, 60), gaps=(2, 1, 2, 1))
    b = from_workload(w)
    with pytest.raises(RuntimeError):
        b.scores()
    b.score_and_grad()
    with pytest.raises(ValueError):
        b.weighted_grad(np.ones(3))
    from tests.gpu_util import params_of
    mat2 = w.matrix * 2 - 0.5
    b.set_params(params_of(w.alphabet, mat2, [3, 0.5, 1, 2]))
    with pytest.raises(RuntimeError):
        b.compute_grad()
    b.score_and_grad()
    s_ref, g_ref = oracle.hard(mat2, [3, 0.5, 1, 2], w.codes_a, w.off_a, w.codes_b, w.off_b)
    np.testing.assert_array_equal(b.scores(), s_ref)
    np.testing.assert_array_equal(grads_rows(b), g_ref)


def _total_is_exact_in_any_order(s):
    """The certificate the library applies before it takes the device's sum as the total: all
    scores finite, multiples of 2^q (q: the lowest set-bit exponent over the nonzero ones),
    and the sum of magnitudes, with rounding margin, below 2^(53 + q)."""
    s = np.asarray(s, np.float64)
    if not np.all(np.isfinite(s)):
        return False
    nz = s[s != 0]
    if nz.size == 0:
        return True
    m, e = np.frexp(np.abs(nz))
    ints = (m * 2.0 ** 53).astype(np.int64)
    low = np.log2((ints & -ints).astype(np.float64)).astype(np.int64)
    q = int(np.min(low + e - 53))
    bound = float(np.sum(np.abs(nz))) * (1.0 + (len(s) + 4) * 2.0 ** -52)
    return bound < 2.0 ** (53 + q)


def test_total_paths(gpu):
    """score_and_grad()'s total is the pair-order sum of the scores (the reference's
    left-to-right loop) whichever way it is formed: the device's sum where every order is
    exact, the host's pair-order loop elsewhere.  Cases on both sides of the certificate,
    checked against np.cumsum (sequential), and the path each took."""
    w = wl.random_pairs("total_paths", 3000, wl.AA, (20, 200), (20, 200), related=0.5)
    r = wl.rng("total_paths", "matrix")
    frac = w.matrix + r.uniform(-0.5, 0.5, size=w.matrix.shape)   # arbitrary doubles
    frac = (frac + frac.T) / 2
    cases = [
        ("integer matrix, float32", w.matrix, (11, 1, 11, 1), "float32", "global", True),
        ("integer matrix, quarter gaps, float64", w.matrix, (4.5, 0.25, 3.75, 1.0), "float64", "global", True),
        ("fractional matrix, float64", frac, (11.1, 0.9, 10.7, 1.3), "float64", "global", False),
        ("all scores zero (local, negative matrix)", -np.abs(w.matrix) - 1, (11, 1, 11, 1), "float64", "local",
         True),
        # 2^-40 in one gap: every global score carries it; thousands of pairs of ~100 push the
        # sum of magnitudes past 2^13
        ("gap with a 2^-40 part, float64", w.matrix, (11, 1 + 2.0 ** -40, 11, 1), "float64", "global", False),
    ]
    for name, mat, gaps, precision, mode, exact in cases:
        say("case total", name)
        b = batch(wl.AA, mat, gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, precision=precision, mode=mode)
        f0, o0 = b.total_info
        total = b.score_and_grad()
        s = b.scores()
        assert _total_is_exact_in_any_order(s) == exact, name
        f1, o1 = b.total_info
        assert (f1 - f0, o1 - o0) == ((1, 0) if exact else (0, 1)), name
        assert total == float(np.cumsum(s)[-1]), name
        assert np.signbit(total) == np.signbit(np.cumsum(s)[-1]), name


@pytest.mark.parametrize("fill", ["interpair", "striped"])
def test_table_budget(gpu, fill):
    """A pointer-code budget too small for a slot per tile: the default fill turns into one
    persistent launch whose waves (one slot each) fit the budget, pulling tiles from the device
    queue; the per-bucket fill splits its buckets into launches of 1,024 tiles.  Every score and
    count unchanged, and the oracle's."""
    w = wl.random_pairs("table_budget", 100000, wl.DNA, (10, 40), (10, 60), gaps=(3, 1, 3, 1), related=0.5)
    s_ref, g_ref = oracle.hard(w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b)
    b = from_workload(w, fill=fill)
    launches, persistent, _ = b.hard_launches
    assert persistent == 0 and launches >= 1 and (launches == 1 or fill == "striped")
    budget = 1 << 20
    b.set_table_budget(budget)
    launches2, persistent2, _ = b.hard_launches
    if fill == "interpair":
        assert (launches2, persistent2) == (1, 1)
    else:
        assert persistent2 == 0 and launches2 >= -(-b.tiles // 1024) and launches2 > launches
    total = b.score_and_grad()
    grid = b.hard_launches[2]
    say("case table budget", fill, b.tiles, "tiles", launches, "->", launches2, "launches, persistent grid", grid)
    if fill == "interpair":
        # slots of at least (10 + 1 - 1) rows of 1 KiB each: the budget holds at most 102 of them
        assert 1 <= grid <= budget // (10 * 1024)
    else:
        assert grid == 0
    np.testing.assert_array_equal(b.scores(), s_ref)
    np.testing.assert_array_equal(grads_rows(b), g_ref)
    assert total == float(np.cumsum(s_ref)[-1])
    b.set_table_budget(0)
    assert b.hard_launches[:2] == (launches, 0)
    b.score_and_grad()
    assert b.hard_launches[2] == 0
    np.testing.assert_array_equal(grads_rows(b), g_ref)


class _DevF64:
    """A device buffer of n doubles at `ptr`, for torch.as_tensor (no copy)."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<f8", "data": (ptr, False), "version": 3}


@pytest.mark.parametrize("budget", [0, 1 << 20])
def test_hard_pass_captured(gpu, budget):
    """The hard pass captured in a graph: short tiles in one plain launch (budget 0) or in a
    persistent launch whose queue counter a kernel clears (1 MiB budget), and column-panel pairs
    whose progress words and queue counter kernels clear.  Before every replay the scores are
    overwritten on the device, with eager work and synchronisation around it; every replay
    gives the eager (and the oracle's) scores and counts again."""
    import torch
    r = wl.rng("hard_captured", "seqs")
    A = [r.integers(0, 20, size=int(r.integers(0, 200))).astype(np.uint8) for _ in range(2000)]
    B = [r.integers(0, 20, size=int(r.integers(0, 300))).astype(np.uint8) for _ in range(2000)]
    A += [r.integers(0, 20, size=int(r.integers(100, 400))).astype(np.uint8) for _ in range(6)]
    B += [r.integers(0, 20, size=int(r.integers(1500, 2600))).astype(np.uint8) for _ in range(6)]
    ca, oa = wl._pack(A)
    cb, ob = wl._pack(B)
    mat = wl.random_matrix(wl.rng("hard_captured", "matrix"))
    gaps = np.array([11.0, 1.0, 11.0, 1.0])
    s_ref, g_ref = oracle.hard(mat, gaps, ca, oa, cb, ob)
    b = batch(wl.AA, mat, gaps, ca, oa, cb, ob)
    b.set_table_budget(budget)
    stream = torch.cuda.Stream()
    b._set_stream(stream.cuda_stream)
    b.score_and_grad()   # sizes every device buffer outside the capture
    assert b.hard_launches[1] == (1 if budget else 0)
    np.testing.assert_array_equal(b.scores(), s_ref)
    np.testing.assert_array_equal(grads_rows(b), g_ref)
    say("case hard captured, budget", budget, "persistent grid", b.hard_launches[2])
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        b._run_async()
    scores_dev = torch.as_tensor(_DevF64(b._scores_dev_ptr(), len(b)), device="cuda")
    for k in range(3):
        scores_dev.fill_(-7.0)
        _ = (torch.ones(8, device="cuda") * 2).sum().item()   # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(b.scores(), s_ref)
        np.testing.assert_array_equal(grads_rows(b), g_ref)
