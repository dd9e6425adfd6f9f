"""The serving loop (nwgrad_amd.serving): jobs coalesced into shared launches, several launches in
flight, give each job the result it has alone -- hard and score-only bit for bit, soft within
1e-5 relative -- and admission respects the HBM budget."""
import numpy as np
import pytest

from nwgrad_amd import workloads as wl
from tests.gpu_util import batch, grads_rows, params_of, say

pytestmark = pytest.mark.gpu


def _jobs(name, n_jobs, max_pairs, lens, kinds=("score", "hard", "soft")):
    r = wl.rng(name, "jobs")
    jobs = []
    for k in range(n_jobs):
        kind = kinds[k % len(kinds)]
        n = int(r.integers(1, max_pairs + 1))
        A, B = [], []
        for _ in range(n):
            a = r.integers(0, 20, size=int(r.integers(lens[0], lens[1] + 1))).astype(np.uint8)
            if kind == "soft":
                b = wl.mutate(r, a, 20, 0.3, 0.02, 3.0) if len(a) else a
            else:
                b = r.integers(0, 20, size=int(r.integers(lens[0], lens[1] + 1))).astype(np.uint8)
            A.append(a)
            B.append(b)
        ca, oa = wl._pack(A)
        cb, ob = wl._pack(B)
        jobs.append((kind, ca, oa, cb, ob))
    return jobs


def _alone(kind, ca, oa, cb, ob, mat, gaps):
    b = batch(wl.AA, mat, gaps, ca, oa, cb, ob, grad_mode="none" if kind == "score" else kind)
    b.score_and_grad()
    return b.scores(), (None if kind == "score" else grads_rows(b))


def _check(kind, res, alone):
    s, g = alone
    assert res["state"] == 2, res["error"]
    if kind == "soft":
        np.testing.assert_allclose(res["scores"], s, rtol=1e-5, atol=1e-9)
        np.testing.assert_allclose(res["grads"], g, rtol=1e-5, atol=1e-7)
    else:
        np.testing.assert_array_equal(res["scores"], s)
        if g is not None:
            np.testing.assert_array_equal(res["grads"], g)


def test_same_alone_and_in_the_mix(gpu):
    """36 jobs (score / hard / soft, 1 to 40 pairs, lengths up to 1,400 so some hard and score
    pairs take column panels), submitted in three waves while earlier launches run; every
    job's result against the same pairs run alone."""
    from nwgrad_amd.serving import Server
    mat = wl.random_matrix(wl.rng("serving_mix", "matrix"))
    gaps = np.array([11.0, 1.0, 11.0, 1.0])
    jobs = _jobs("serving_mix", 36, 40, (20, 1400))
    srv = Server(params_of(wl.AA, mat, gaps), slots_per_kind=2, max_cells_per_launch=20_000_000)
    ids = []
    for wave in range(3):
        for j in jobs[wave * 12:(wave + 1) * 12]:
            ids.append(srv.submit(*j))
        for _ in range(3):
            srv.poll()
        say(f"  wave {wave}: {srv.running_launches} launches running, {srv.waiting} jobs waiting")
    srv.drain(timeout=300)
    say(f"  {srv.launches} launches for {len(jobs)} jobs")
    assert srv.launches < len(jobs)          # jobs were coalesced
    for jid, j in zip(ids, jobs):
        _check(j[0], srv.result(jid), _alone(*j, mat, gaps))


def test_hbm_budget_admission(gpu):
    """A budget that holds about one job: jobs run one launch at a time, all complete, results
    unchanged, and the bytes in flight stay within the budget."""
    from nwgrad_amd.serving import Server
    mat = wl.random_matrix(wl.rng("serving_budget", "matrix"))
    gaps = np.array([11.0, 1.0, 11.0, 1.0])
    jobs = _jobs("serving_budget", 8, 20, (50, 300), kinds=("hard",))
    budget = 20 * (404 * 8 + 24 + 600) * 1.05
    srv = Server(params_of(wl.AA, mat, gaps), hbm_budget_bytes=budget, slots_per_kind=4)
    ids = [srv.submit(*j) for j in jobs]
    peak = 0.0
    while srv.waiting or srv.running_launches:
        srv.poll()
        peak = max(peak, srv.bytes_in_flight)
    assert peak <= budget
    for jid, j in zip(ids, jobs):
        _check(j[0], srv.result(jid), _alone(*j, mat, gaps))


def test_poisson_replay(gpu):
    """Two seconds of the serving workload at a reduced scale (100 requests/s, up to 64 pairs of
    up to 1,500 residues): every job completes, latency percentiles are reported."""
    from nwgrad_amd.serving import Server, replay
    load = wl.ServingLoad(rate=100, duration=2.0, max_pairs=64, lo=30, hi=1500, name="serving_smoke")
    reqs = [load.request(k) for k in range(len(load))]
    srv = Server(params_of(wl.AA, load.matrix, load.gaps))
    out = replay(srv, reqs, timeout=120)
    say(f"  {out}")
    assert out["failed"] == 0 and out["jobs"] == len(reqs)
    assert out["p99_ms"] >= out["p50_ms"] > 0


def test_tracked_k4_alone_and_in_the_mix(gpu):
    """The serving configuration's K = 4 arm: params of four sets, every job's A carrying a
    per-residue track in uniform-random contiguous segments (workloads.segment_tracks).  18 jobs
    of the three kinds in two waves; every job against the same pairs run alone in a tracked
    batch, and one hard job against the oracle's tracked restatement."""
    import nwgrad_amd as nw
    import oracle
    from nwgrad_amd.serving import Server
    load = wl.ServingLoad(rate=100, duration=1.0, sets=4, name="serving_k4")
    pk = nw.AlignParams.from_arrays(load.matrices, load.set_gaps, alphabet=wl.AA)
    jobs = _jobs("serving_k4", 18, 30, (20, 1300))
    r = wl.rng("serving_k4", "tracks")
    tracks = [wl.segment_tracks(r, j[2], 4) for j in jobs]
    srv = Server(pk, slots_per_kind=2, max_cells_per_launch=10_000_000)
    assert srv.num_sets == 4
    ids = []
    for wave in range(2):
        for j, t in list(zip(jobs, tracks))[wave * 9:(wave + 1) * 9]:
            ids.append(srv.submit(*j, tracks=t))
        for _ in range(3):
            srv.poll()
        say(f"  wave {wave}: {srv.running_launches} launches running, {srv.waiting} jobs waiting")
    srv.drain(timeout=300)
    say(f"  {srv.launches} launches for {len(jobs)} jobs")
    assert srv.launches < len(jobs)
    checked_oracle = False
    for jid, j, t in zip(ids, jobs, tracks):
        kind, ca, oa, cb, ob = j
        b = nw.SeqPairBatch.from_codes(ca, oa, cb, ob, pk, grad_mode="none" if kind == "score" else kind, tracks=t)
        b.score_and_grad()
        alone = (b.scores(), None if kind == "score" else grads_rows(b))
        res = srv.result(jid)
        _check(kind, res, alone)
        if kind == "hard" and not checked_oracle:
            s_ref, g_ref = oracle.hard_tracked(load.matrices, load.set_gaps, ca, oa, cb, ob, t)
            np.testing.assert_array_equal(res["scores"], s_ref)
            np.testing.assert_array_equal(res["grads"], g_ref)
            checked_oracle = True
    assert checked_oracle
