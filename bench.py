"""Benchmark of the batched alignment hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--fill interpair|striped]

A step is one score_and_grad() pass over the whole batch, inputs already resident in HBM.
Prints ONE JSON line: GCUPS (cells = sum of len(A)*len(B), the reference's convention), the
roofline of the pass's kernels (measured with HIP events on the stream they run on), the same
pass with plain per-bucket launches for comparison, and the CPU baseline: the reference engine
(oracle/_ref, when built) or the repository's C restatement, timed on this host's cores over a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic lane-operations per cell (one max / add / sub / mul / FMA = 1), counted from the
# reference recurrences: affine hard / score 13, linear 5 (local 6), soft affine 32 (forward 11,
# backward with the fused gradient 21), soft linear 14.
OPS_PER_CELL = {("affine", "hard"): 13, ("affine", "none"): 13, ("linear", "hard"): 5, ("linear", "none"): 5,
                ("affine", "soft"): 32, ("linear", "soft"): 14}
FP32_PEAK_TFLOPS = 157.3        # MI355X published FP32 vector peak (packed FMA = 4 flops / lane / clock)
VALU_ISSUE_PEAK_TOPS = 39.3     # 256 CUs x 64 lanes x 2.4 GHz non-packed VALU lane-ops
HBM_PEAK_GBS = 8000.0           # MI355X HBM3E
# Algorithmic HBM bytes per cell of the soft pass: 3 states x 4 B written by the forward and
# read back by the backward (24 B/cell without checkpointing).
SOFT_BYTES_PER_CELL = 24

WORKLOADS = ("protein_soft_homologs", "dna_local_short", "dna_local_short_score", "protein_global_hard",
             "long_hard_36000", "long_hard_36000_double", "long_hard_36000_hb", "long_hard_36000_hb_double",
             "long_hard_4096", "long_soft_8000", "serving", "protein_soft_train")


def measured_traffic(workload, fill):
    """HBM bytes per launch of the workload's dominant kernel and where the figure comes from: the
    committed rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what
    the counter reads for a 1 GiB device copy) -- counters are not read in this run -- or
    (None, None) when no summary exists for this workload and fill."""
    rel = os.path.join("profiles", f"traffic_{workload}_{fill}.json")
    path = os.path.join(ROOT, rel)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    return int((t["fetch_kib"] * t["fetch_factor"] + t["write_kib"] * t["write_factor"]) * 1024), rel


def make_workload(name):
    from nwgrad_amd import workloads as wl
    if name == "protein_global_hard":
        return wl.protein_global_hard(), "hard"
    if name == "dna_local_short":
        return wl.dna_local_short(), "hard"
    if name == "dna_local_short_score":
        return wl.dna_local_short(), "none"
    if name == "protein_soft_homologs":
        return wl.protein_soft_homologs(), "soft"
    # long pairs: B a mutated copy of A (substitution 0.1, indels 0.01, mean length 5), global affine 11/1
    if name in ("long_hard_36000", "long_hard_36000_double", "long_hard_36000_hb", "long_hard_36000_hb_double"):
        return wl.long_pairs(16, 36000), "hard"
    if name == "long_hard_4096":
        return wl.long_pairs(256, 4096), "hard"
    if name == "long_soft_8000":
        return wl.long_pairs(4, 8000, grad_mode="soft"), "soft"
    raise SystemExit(f"unknown workload {name}")


def traceback_of(workload):
    """The long-pair configuration's linear-memory arm runs the Hirschberg recursion (base case
    of 512 rows); every other workload the default traceback."""
    return "hirschberg" if "_hb" in workload else "auto"


def make_batch(w, grad_mode, fill, precision="float32", traceback="auto"):
    import nwgrad_amd as nw
    p = nw.AlignParams(w.matrix, alphabet=w.alphabet, gap_open_a=w.gaps[0], gap_extend_a=w.gaps[1],
                       gap_open_b=w.gaps[2], gap_extend_b=w.gaps[3])
    cls = nw.SeqPairBatch if precision == "float32" else nw.SeqPairBatchDouble
    b = cls.from_codes(w.codes_a, w.off_a, w.codes_b, w.off_b, p, gap_model=w.gap_model, mode=w.mode,
                       grad_mode=grad_mode, traceback=traceback)
    b.fill = fill
    return b


def time_steps(b, steps, warmup, sync):
    for _ in range(warmup):
        b.score_and_grad()
    sync()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        b.score_and_grad()
        kms.append(b.last_kernel_ms)
    sync()
    dt = time.perf_counter() - t0
    return dt, float(np.mean(kms))


def cpu_baseline(w, grad_mode, precision="float32", budget_s=8.0, traceback="auto"):
    """The reference engine (oracle/_ref) on all host cores, else the C oracle on one: a
    bounded sample of the workload's pairs, best of 2 timed passes after a warm-up."""
    import oracle
    cores = min(len(os.sched_getaffinity(0)), 16)
    rng = np.random.default_rng(0)
    if oracle.have_ref():
        from oracle import ref
        kind = "reference"

        def run(sub):
            sa, sb = sub.strings()
            rb = ref.RefBatch(sa, sb, sub.alphabet, sub.matrix, sub.gaps, gap_model=sub.gap_model, mode=sub.mode,
                              grad_mode=grad_mode, n_threads=cores, double=precision == "float64",
                              traceback=traceback)
            if grad_mode == "soft":
                rb.set_soft(precision="float32", impl="scaled_or_log")
            rb.score_and_grad()  # warm-up
            best = 1e30
            for _ in range(2):
                t0 = time.perf_counter()
                rb.score_and_grad()
                best = min(best, time.perf_counter() - t0)
            return best
    else:
        kind = "port"
        cores = 1

        def run(sub):
            t0 = time.perf_counter()
            if grad_mode == "soft":
                oracle.soft_log(sub.matrix, sub.gaps, sub.codes_a, sub.off_a, sub.codes_b, sub.off_b,
                                gap_model=sub.gap_model, mode=sub.mode)
            else:
                oracle.hard(sub.matrix, sub.gaps, sub.codes_a, sub.off_a, sub.codes_b, sub.off_b,
                            gap_model=sub.gap_model, mode=sub.mode, with_grad=grad_mode == "hard",
                            precision=precision)
            return time.perf_counter() - t0
    n = w.n_pairs
    # a first sample of about 2e8 cells, and at least one pair per core (the reference
    # parallelises over pairs: fewer pairs than cores would leave cores idle)
    k = min(n, max(cores if kind == "reference" else 1, min(256, int(2e8 * n / max(w.cells, 1)))))
    idx = np.sort(rng.choice(n, size=k, replace=False))
    sub = w.subset(idx)
    t = run(sub)
    per_cell = t / max(sub.cells, 1)
    target = int(min(n, max(k, k * (budget_s / 3.0) / max(t, 1e-6))))
    if target > k:
        idx = np.sort(rng.choice(n, size=target, replace=False))
        sub = w.subset(idx)
        t = run(sub)
        per_cell = t / max(sub.cells, 1)
    return {"value": round(1.0 / per_cell / 1e9, 4), "unit": "GCUPS", "cores": cores, "kind": kind,
            "sample": f"{sub.n_pairs} of {n} pairs ({sub.cells} cells), uniform without replacement, "
                      f"best of 2 passes after a warm-up"}


def serving_leg(duration, rate=200.0, warm=3.0):
    """The serving workload (Poisson arrivals at `rate`/s; pairs per request log-U[1, 4096];
    lengths log-U[30, 5000]; 70 / 20 / 10 % score / hard / soft; global affine 11/1, float32;
    HBM budget 200 GB) replayed against nwgrad_amd.serving.Server in real time.  The first
    `warm` seconds of arrivals run first, unmeasured (slot creation, buffer growth); requests
    are generated before the replay, so the client's own work is not in the latencies."""
    import nwgrad_amd as nw
    from nwgrad_amd import workloads as wl
    from nwgrad_amd.serving import Server, replay
    load = wl.ServingLoad(rate=rate, duration=duration + warm)
    reqs = [load.request(k) for k in range(len(load))]
    for q in reqs:
        q.t -= warm
    warm_reqs = [q for q in reqs if q.t < 0]
    for q in warm_reqs:
        q.t += warm
    timed = [q for q in reqs if q.t >= 0]
    p = nw.AlignParams(load.matrix, alphabet=wl.AA, gap_open_a=11.0, gap_extend_a=1.0, gap_open_b=11.0,
                       gap_extend_b=1.0)
    srv = Server(p, hbm_budget_bytes=200e9)
    print(f"bench: serving warm-up, {len(warm_reqs)} requests", file=sys.stderr, flush=True)
    replay(srv, warm_reqs)
    print(f"bench: serving, {len(timed)} requests over {duration:.0f} s", file=sys.stderr, flush=True)

    def report(st):
        print(f"bench: serving t={st['t_s']:.0f} s: {st['submitted']}/{st['of']} submitted, {st['completed']} done, "
              f"{st['waiting']} waiting, {st['running']} launches running", file=sys.stderr, flush=True)

    out = replay(srv, timed, progress=report)
    offered = sum(q.cells for q in timed) / duration / 1e9
    res = {"workload": "serving", "rate_per_s": rate, "duration_s": duration, "requests": out["jobs"],
           "failed": out["failed"], "p50_ms": round(out["p50_ms"], 3), "p99_ms": round(out["p99_ms"], 3),
           "max_ms": round(out["max_ms"], 3), "gcups_delivered": round(out["gcups"], 3),
           "gcups_offered": round(offered, 3), "launches": out["launches"],
           "mix": "70/20/10 score/hard/soft", "pairs_per_request": "log-U[1,4096]", "lengths": "log-U[30,5000]",
           "warmup_s": warm}
    print(f"bench: serving p50 {res['p50_ms']} ms, p99 {res['p99_ms']} ms, {res['gcups_delivered']} GCUPS "
          f"delivered of {res['gcups_offered']} offered", file=sys.stderr, flush=True)
    return res


def training_leg(steps, warmup, lr=1e-3, margin=20.0):
    """The contrastive training loop on protein_soft_homologs: per step the soft pass
    (torch.autograd.Function over the HIP kernel), the margin loss of homolog against decoy
    scores, the adjoint (weighted-gradient reduction) and an SGD update of the parameter
    vector, captured once as a HIP graph and replayed `steps` times.  Returns (seconds,
    cells per step, final loss, pairs rerun in double)."""
    import torch
    import nwgrad_amd as nw
    from nwgrad_amd import autograd as ag
    from nwgrad_amd import workloads as wl
    w = wl.protein_soft_homologs()
    b = make_batch(w, "soft", "interpair")
    n_pos = w.n_pairs // 2
    p = nw.AlignParams(w.matrix, alphabet=w.alphabet, gap_open_a=w.gaps[0], gap_extend_a=w.gaps[1],
                       gap_open_b=w.gaps[2], gap_extend_b=w.gaps[3])
    theta = ag.params_vector(p).requires_grad_(True)
    loss_buf = torch.zeros((), dtype=torch.float64, device="cuda")

    def step():
        s = ag.soft_scores(theta, b)
        loss = ag.margin_loss(s, n_pos, margin)
        (g,) = torch.autograd.grad(loss, theta)
        with torch.no_grad():
            theta.sub_(lr * g)
            loss_buf.copy_(loss)

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(max(1, warmup)):
            step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    ag.finish(b)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        graph.replay()
        if (k + 1) % 25 == 0:
            torch.cuda.synchronize()
            print(f"bench: training step {k + 1}, loss {float(loss_buf):.5f}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    reruns = ag.finish(b)
    return dt, b.cells, float(loss_buf), reruns


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=os.environ.get("NWG_BENCH_WORKLOAD", "protein_soft_homologs"),
                    choices=WORKLOADS)
    ap.add_argument("--fill", default="interpair", choices=("interpair", "striped"))
    ap.add_argument("--columns", type=int, default=0, choices=(0, 8, 16),
                    help="columns a lane of the hard / score-only tiles (0: the lane-width table's choice)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-launch comparison leg")
    ap.add_argument("--serve-seconds", type=float, default=None,
                    help="seconds of serving arrivals (default: 10 beside the main workload, 60 for "
                         "--workload serving; 0: no serving leg)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only time a 1 GiB device-to-device copy (the known-size reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: one GPU only; --gpus {args.gpus} is not supported (across GPUs the pairs "
              "would shard by cells, each GPU running its share, with one all-reduce of the gradient "
              "per step)", file=sys.stderr)
        return 2
    import torch
    sync = torch.cuda.synchronize
    if args.copy_calibration:
        nbytes = 1 << 30
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        sync()
        t0 = time.perf_counter()
        for _ in range(5):
            dst.copy_(src)
        sync()
        dt = (time.perf_counter() - t0) / 5
        print(json.dumps({"copy_bytes": nbytes, "ms": round(dt * 1e3, 4), "GBps_read_plus_write": round(2 * nbytes / dt / 1e9, 1)}),
              flush=True)
        return 0
    print(f"bench: workload {args.workload}", file=sys.stderr, flush=True)
    if args.workload == "protein_soft_train":
        steps = 100 if args.steps == 10 else args.steps
        dt, cells, loss, reruns = training_leg(steps, args.warmup)
        gcups = cells * steps / dt / 1e9
        line = {"metric": "GCUPS soft-grad training steps (soft pass, margin loss, adjoint, SGD) in one HIP graph",
                "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
                "ms_per_step": round(dt / steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                "config": {"workload": "protein_soft_train", "pairs": 65536, "cells": cells, "lr": 1e-3,
                           "margin": 20.0, "final_loss": loss, "pairs_rerun_in_double": reruns}}
        print(json.dumps(line), flush=True)
        return 0
    if args.workload == "serving":
        res = serving_leg(60.0 if args.serve_seconds is None else args.serve_seconds)
        line = {"metric": "GCUPS delivered under Poisson arrivals; p50 / p99 latency", "value": res["gcups_delivered"],
                "unit": "GCUPS", "n_gpus": 1, "steps": res["requests"], "warmup": 0, "ms_per_step": res["p50_ms"],
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                "data": "synthetic", "config": {"workload": "serving"}, "serving": res}
        print(json.dumps(line), flush=True)
        return 0
    w, grad_mode = make_workload(args.workload)
    precision = "float64" if args.workload.endswith("_double") else "float32"
    tb = traceback_of(args.workload)
    b = make_batch(w, grad_mode, args.fill, precision, tb)
    if args.columns and grad_mode != "soft":
        b.columns_per_lane = args.columns
    cells = b.cells
    print(f"bench: {w.n_pairs} pairs, {cells} cells, {b.tiles} tiles; timing {args.steps} steps", file=sys.stderr,
          flush=True)
    dt, kms = time_steps(b, args.steps, args.warmup, sync)
    gcups = cells * args.steps / dt / 1e9
    reruns = b.soft_reruns if grad_mode == "soft" else 0
    launch = b.launch_info if grad_mode == "soft" else None
    wide = b.wide_info if grad_mode == "soft" else None
    print(f"bench: {args.fill}: {dt / args.steps * 1e3:.3f} ms/step, {gcups:.2f} GCUPS, kernels {kms:.3f} ms"
          + (f", {reruns} pairs rerun in double, launch (grid, wide blocks, per CU) {launch}, wide tiles "
             f"(pipelined, pooled, pool rows, lag) {wide}" if grad_mode == "soft" else ""), file=sys.stderr, flush=True)
    # the other launch mode, same batch shape
    other = "striped" if args.fill == "interpair" else "interpair"
    gcups_other = kms2 = None
    if not args.no_plain:
        b.fill = other
        dt2, kms2 = time_steps(b, max(2, args.steps // 2), 1, sync)
        gcups_other = cells * max(2, args.steps // 2) / dt2 / 1e9
        print(f"bench: {other}: {dt2 / max(2, args.steps // 2) * 1e3:.3f} ms/step, {gcups_other:.2f} GCUPS",
              file=sys.stderr, flush=True)
    b.fill = args.fill
    ops = OPS_PER_CELL[(w.gap_model, grad_mode)] + (1 if (w.mode == "local" and w.gap_model == "linear") else 0)
    # the recursion fills about twice the pairs' cells (each level re-sweeps its half): the
    # roofline counts the cells filled, GCUPS the pairs' cells
    work_cells = int(b.hb_info[4]) if tb != "auto" else cells
    achieved = ops * work_cells / (kms * 1e-3) / 1e12
    valu = {"achieved": round(achieved, 3), "unit": "T lane-ops/s", "ops_per_cell": ops,
            "issue_peak": VALU_ISSUE_PEAK_TOPS, "issue_frac": round(achieved / VALU_ISSUE_PEAK_TOPS, 4),
            "fp32_peak_tflops": FP32_PEAK_TFLOPS, "fp32_frac": round(achieved / FP32_PEAK_TFLOPS, 4)}
    traffic, traffic_from = measured_traffic(args.workload, args.fill)
    # Every kernel of the path is bound by VALU issue (max-plus, or sum-product with per-cell
    # scalar weights: no MFMA, and the counters show HBM well below its peak): the roof is the
    # FP32 vector peak, against which `frac` is quoted; `valu` adds the non-packed issue fraction.
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic, "traffic_from": traffic_from,
            "ops_per_cell": ops, "cells": work_cells, "kernel_ms": round(kms, 4), "valu": valu}
    if grad_mode == "soft":
        # secondary: the table traffic the pass would have without recomputation (3 states x 4 B
        # written, then read back), as an HBM rate -- a model, not a measurement
        gbs = SOFT_BYTES_PER_CELL * cells / (kms * 1e-3) / 1e9
        roof["hbm_model"] = {"bytes_per_cell": SOFT_BYTES_PER_CELL, "achieved": round(gbs, 1), "unit": "GB/s",
                             "peak": HBM_PEAK_GBS, "frac": round(gbs / HBM_PEAK_GBS, 4)}
    if traffic:
        # the committed counters' bytes per step over this run's kernel time per step
        roof["hbm_measured"] = {"achieved": round(traffic / (kms * 1e-3) / 1e9, 1), "unit": "GB/s",
                                "peak": HBM_PEAK_GBS, "frac": round(traffic / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if not args.no_cpu:
        print("bench: CPU baseline (bounded sample)", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline(w, grad_mode, precision, traceback=tb)
    serve_s = 10.0 if args.serve_seconds is None else args.serve_seconds
    serving = serving_leg(serve_s) if serve_s > 0 else None
    dtype = "f32" if precision == "float32" else "f64"
    line = {"metric": "GCUPS soft-grad, batched vs plain GPU fill @ bit-exact/1e-5; p99 latency",
            "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(dt / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic",
            "config": {"workload": args.workload, "pairs": w.n_pairs, "cells": cells, "grad_mode": grad_mode,
                       "gap_model": w.gap_model, "mode": w.mode, "fill": args.fill, "traceback": tb,
                       "columns_per_lane": int(b.columns_per_lane)},
            "kernel_gcups": round(cells / (kms * 1e-3) / 1e9, 3),
            "plain_vs_batched": None if gcups_other is None else
                {"fill": other, "gcups": round(gcups_other, 3), "kernel_gcups": round(cells / (kms2 * 1e-3) / 1e9, 3),
                 "batched_over_plain": round(gcups / gcups_other, 3)},
            "roofline": roof, "cpu_baseline": cpu, "serving": serving}
    if tb != "auto":
        split, levels, blocks, scratch, filled = b.hb_info
        line["hirschberg"] = {"pairs_split": int(split), "levels": int(levels), "base_blocks": int(blocks),
                              "scratch_bytes": int(scratch), "hb_cutoff": int(b.hb_cutoff), "cells_filled": int(filled),
                              "pointer_table_bytes": int(cells)}
    if grad_mode == "soft":
        line["soft_pairs_rerun_in_double"] = int(reruns)
        line["launch"] = {"grid": int(launch[0]), "wide_queue_blocks": int(launch[1]), "resident_per_cu": int(launch[2])}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
