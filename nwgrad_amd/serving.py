"""Serving: continuous batching of alignment jobs on one GPU.

``Server`` wraps the native ``AlignServer`` (``cpp/nwgrad_amd/serving.hpp``): jobs of pairs --
score only, hard gradient or soft gradient, under one parameter set and problem type -- are
queued by ``submit``; ``poll`` finishes the launches whose streams are idle and coalesces
waiting jobs into new persistent launches on free slots (several slots per kind, each on its own
HIP stream, so launches overlap), within an HBM budget.  Encoded sequences stay on the device in
each slot's batch, which is reloaded, not reallocated, for every launch.

``replay`` drives a server with timed arrivals (the serving workload,
``workloads.ServingLoad``) and reports per-job latency (submit to completion) as p50 / p99 with
the cells/s the server delivered.

The reference has no serving loop: it is a synchronous library.  The per-pair results are the
batch classes' own (hard and score-only bit for bit whichever launch a pair lands in).
"""
from __future__ import annotations

import time

import numpy as np

from ._nwgrad_amd import AlignServer

KINDS = ("score", "hard", "soft")


class Server:
    """A continuous-batching server over ``params`` (an ``AlignParams``, one or K parameter sets).

    ``double``: float64 (the ``SeqPairBatchDouble`` numerics) instead of float32.
    ``hbm_budget_bytes``: device memory the jobs in flight may hold (codes, results and hard
    long-pair pointer tables); a job larger than the budget runs alone.
    ``slots_per_kind``: concurrent launches per kind.  ``max_cells_per_launch`` /
    ``max_pairs_per_launch``: how much waiting work one launch coalesces."""

    def __init__(self, params, *, gap_model="affine", mode="global", double=False, hbm_budget_bytes=200e9,
                 slots_per_kind=2, max_cells_per_launch=4_000_000_000, max_pairs_per_launch=1 << 20,
                 soft_temperature=1.0, device=0):
        self.alphabet = params.alphabet
        self._s = AlignServer(params, gap_model=gap_model, mode=mode, double=double,
                              hbm_budget_bytes=float(hbm_budget_bytes), slots_per_kind=int(slots_per_kind),
                              max_cells_per_launch=int(max_cells_per_launch),
                              max_pairs_per_launch=int(max_pairs_per_launch), soft_temperature=float(soft_temperature),
                              device=int(device))

    def submit(self, kind, codes_a, off_a, codes_b, off_b, tracks=None, gap_tracks=None) -> int:
        """Queue a job of encoded pairs (``uint8`` codes, ``int64`` offsets from 0); its id.

        With params of K > 1 parameter sets the server is tracked: ``tracks`` is one set per
        residue of A (flat, ``off_a[-1]`` entries) and ``gap_tracks`` one per DP row (flat,
        ``off_a[-1] + n`` entries, pair p's at ``off_a[p] + p``); either may be None for the
        reference's defaults.  Gradient rows then hold K matrices, then K rows of gaps."""
        if kind not in KINDS:
            raise ValueError(f"nwgrad_amd: unknown job kind {kind!r} (expected one of {KINDS})")
        return self._s.submit(kind, np.ascontiguousarray(codes_a, np.uint8), np.ascontiguousarray(off_a, np.int64),
                              np.ascontiguousarray(codes_b, np.uint8), np.ascontiguousarray(off_b, np.int64),
                              tracks=None if tracks is None else np.ascontiguousarray(tracks, np.int64),
                              gap_tracks=None if gap_tracks is None else np.ascontiguousarray(gap_tracks, np.int64))

    def submit_strings(self, kind, seqs_a, seqs_b) -> int:
        """Queue a job given as two lists of strings over the params' alphabet."""
        if len(seqs_a) != len(seqs_b):
            raise ValueError("nwgrad: seqs_a and seqs_b must have the same length")
        table = np.full(256, 255, np.uint8)
        for i, ch in enumerate(self.alphabet):
            table[ord(ch)] = i

        def enc(seqs):
            off = np.zeros(len(seqs) + 1, np.int64)
            np.cumsum([len(s) for s in seqs], out=off[1:])
            raw = np.frombuffer("".join(seqs).encode("latin-1"), np.uint8)
            codes = table[raw]
            if np.any(codes == 255):
                bad = chr(int(raw[np.argmax(codes == 255)]))
                raise ValueError(f"nwgrad: character '{bad}' is not in alphabet \"{self.alphabet}\"")
            return codes, off

        ca, oa = enc(seqs_a)
        cb, ob = enc(seqs_b)
        return self.submit(kind, ca, oa, cb, ob)

    def poll(self) -> int:
        """One scheduler pass; the number of jobs that completed in it."""
        return self._s.poll()

    def state(self, job_id) -> str:
        return self._s.state(job_id)

    def result(self, job_id) -> dict:
        """``scores`` (n,), ``grads`` (n, n_alpha**2 + 4; hard and soft) or None, timestamps on
        the server clock (``t_submit``, ``t_launch``, ``t_done``), ``error`` ('' when done)."""
        return self._s.result(job_id)

    def release(self, job_id):
        self._s.release(job_id)

    def now(self) -> float:
        return self._s.now()

    @property
    def num_sets(self) -> int:
        return self._s.num_sets

    @property
    def waiting(self) -> int:
        return self._s.waiting

    @property
    def running_launches(self) -> int:
        return self._s.running_launches

    @property
    def launches(self) -> int:
        return self._s.launches

    @property
    def bytes_in_flight(self) -> float:
        return self._s.bytes_in_flight

    def drain(self, timeout=600.0, idle_sleep=2e-4):
        """Poll until nothing waits or runs."""
        t_end = time.perf_counter() + timeout
        while self.waiting or self.running_launches:
            if self.poll() == 0:
                time.sleep(idle_sleep)
            if time.perf_counter() > t_end:
                raise TimeoutError("nwgrad_amd: the server did not drain in time")


def replay(server: Server, requests, *, idle_sleep=2e-4, timeout=900.0, progress=None, progress_every=10.0):
    """Submit ``requests`` (objects with ``t`` (s from start), ``kind`` and encoded pairs) at
    their arrival times, polling the server in between; wait for all of them.  Returns a dict:
    per-job latency (submit to completion) p50 / p99 / max in ms, jobs, failed jobs, cells, the
    span from the first submit to the last completion, GCUPS over it, launches.
    ``progress``: called every ``progress_every`` seconds with a dict of the replay's state
    (seconds since start, jobs submitted and completed, jobs waiting, launches running)."""
    order = sorted(range(len(requests)), key=lambda k: requests[k].t)
    ids = []
    cells = 0
    done = 0
    t0 = time.perf_counter()
    t_end = t0 + timeout
    t_report = progress_every
    i = 0
    while i < len(order) or server.waiting or server.running_launches:
        now = time.perf_counter() - t0
        while i < len(order) and requests[order[i]].t <= now:
            q = requests[order[i]]
            ids.append(server.submit(q.kind, q.codes_a, q.off_a, q.codes_b, q.off_b, tracks=getattr(q, "tracks", None)))
            cells += q.cells
            i += 1
        if progress is not None and now >= t_report:
            progress({"t_s": now, "submitted": i, "of": len(order), "completed": done, "waiting": server.waiting,
                      "running": server.running_launches})
            t_report += progress_every
        finished = server.poll()
        done += finished
        if finished == 0:
            nxt = requests[order[i]].t - (time.perf_counter() - t0) if i < len(order) else idle_sleep
            time.sleep(max(0.0, min(idle_sleep, nxt)))
        if time.perf_counter() > t_end:
            raise TimeoutError("nwgrad_amd: replay did not finish in time")
    lat, failed, t_first, t_last = [], 0, None, None
    for jid in ids:
        r = server.result(jid)
        if r["state"] != 2:
            failed += 1
        lat.append(r["t_done"] - r["t_submit"])
        t_first = r["t_submit"] if t_first is None else min(t_first, r["t_submit"])
        t_last = r["t_done"] if t_last is None else max(t_last, r["t_done"])
        server.release(jid)
    lat = np.array(lat) * 1e3
    span = (t_last - t_first) if ids else 0.0
    return {"jobs": len(ids), "failed": failed, "cells": cells, "span_s": span,
            "gcups": cells / span / 1e9 if span > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "max_ms": float(lat.max()) if len(lat) else None, "launches": server.launches}
