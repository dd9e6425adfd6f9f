"""Per-length-bucket lane-width tuning of the hard and score-only fills.

The hard and score-only fills cut their tiles at 8 or 16 columns a lane.  A pair of len(B) = n
then takes the power of two of ceil(n / columns) lanes, so 64 / lanes pairs share a wavefront,
and a tile runs max len(A) + lanes - 1 row steps.  Which width is faster depends on two things:
len(B), which sets the lanes per pair and with them the skew a tile pays; and the batch's tile
count per CU, since a fill with fewer than about two waves a SIMD has nothing to hide a row
step's latency behind, and 8 columns a lane gives twice the waves.

``sweep()`` measures both widths on the GPU over a grid of (len(B), tiles per CU) buckets --
square protein pairs, global affine, hard gradient, float32 -- and returns the measurements and
the rules that pick the faster width per bucket.  ``save()`` writes them to
``nwgrad_amd/lane_table.json``; ``install()`` loads a table into the library (the package does
so at import when that file exists).  A batch whose median len(B) and tile count fall in a
bucket is cut at the bucket's width; outside the table the built-in rule holds.  Results never
depend on the width (the fills are bit-exact at both), only the time does.

The reference's counterpart is its runtime gates and tuning table (``simd_levels.hpp:704-751``,
``src/nwgrad/_tuning.py``), which choose the inter-pair group width per ISA.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lane_table.json")

# The sweep's grid.  Length buckets are the ranges of len(B) that take one lane count at 16
# columns a lane (and twice that at 8): (32, 64], (64, 128], ..., each measured at three
# quarters of its top (len(A) = len(B)).  Tile counts are per CU, at 16 columns a lane.
BUCKETS = ((32, 24), (64, 48), (128, 96), (256, 192), (512, 384))   # (len(B) at most, measured at)
TILES_PER_CU = (0.25, 0.5, 1.0, 2.0, 4.0, 8.0, 32.0)
WIDTHS = (16, 8)


def _lanes(n: int, columns: int) -> int:
    g = max(1, -(-n // columns))
    p = 1
    while p < g:
        p <<= 1
    return min(p, 64)


def _batch(nw, n: int, n_pairs: int, seed: int):
    """n_pairs square protein pairs of len(A) = len(B) = n, residues uniform, under a random
    symmetric matrix (diagonal 4..11, off-diagonal -4..3) and gaps 11 / 1."""
    r = np.random.default_rng(seed)
    ca = r.integers(0, 20, size=n_pairs * n, dtype=np.uint8)
    cb = r.integers(0, 20, size=n_pairs * n, dtype=np.uint8)
    off = np.arange(0, (n_pairs + 1) * n, n, dtype=np.int64)
    mat = r.integers(-4, 4, size=(20, 20)).astype(np.float64)
    mat = np.triu(mat) + np.triu(mat, 1).T
    np.fill_diagonal(mat, r.integers(4, 12, size=20))
    p = nw.AlignParams(mat, alphabet=nw.AA_ORDER, gap_open_a=11.0, gap_extend_a=1.0, gap_open_b=11.0,
                       gap_extend_b=1.0)
    return nw.SeqPairBatch.from_codes(ca, off, cb, off, p, gap_model="affine", mode="global", grad_mode="hard")


def _time(b, sync, reps: int) -> float:
    b.score_and_grad()
    sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.score_and_grad()
    sync()
    return (time.perf_counter() - t0) / reps


def sweep(n_cu: int | None = None, reps: int = 5, log=print) -> dict:
    """Measure 16 and 8 columns a lane over the grid on cuda:0; returns {"rules", "measured",
    "device", "problem"}.  Each point times `reps` passes after one warm-up pass."""
    import torch

    import nwgrad_amd as nw
    sync = torch.cuda.synchronize
    props = torch.cuda.get_device_properties(0)
    n_cu = n_cu or props.multi_processor_count
    measured = []
    for n_max, n in BUCKETS:
        per_tile = 64 // _lanes(n, 16)
        for tpc in TILES_PER_CU:
            n_pairs = max(per_tile, int(round(tpc * n_cu)) * per_tile)
            b = _batch(nw, n, n_pairs, seed=n * 1000 + int(tpc * 4))
            row = {"n_max": n_max, "n": n, "tiles_per_cu": tpc, "pairs": n_pairs, "cells": n_pairs * n * n}
            ref = None
            for w in WIDTHS:
                b.columns_per_lane = w
                assert b.columns_per_lane == w
                dt = _time(b, sync, reps)
                row[f"gcups_{w}"] = round(n_pairs * n * n / dt / 1e9, 3)
                s = b.scores()
                if ref is None:
                    ref = s
                elif not np.array_equal(s, ref):
                    raise RuntimeError(f"lane widths disagree at n={n}, {n_pairs} pairs")
            row["best"] = max(WIDTHS, key=lambda w: row[f"gcups_{w}"])
            measured.append(row)
            log(f"tuning: len(B) {n:4d}, {tpc:5.2f} tiles/CU ({n_pairs} pairs): "
                + ", ".join(f"{w} columns {row[f'gcups_{w}']:.1f} GCUPS" for w in WIDTHS) + f" -> {row['best']}")
            del b
    return {"rules": rules_from(measured), "measured": measured, "device": props.name,
            "problem": "square protein pairs, global affine 11/1, hard, float32"}


def rules_from(measured) -> list:
    """The rules (n_max, tiles16_per_cu_max, columns), first match wins, tile counts in
    sixteenths of a tile per CU: per length bucket, ascending, the width that was faster at
    each measured tile count, each count's rule reaching up to the geometric midpoint to the
    next count (the last one without bound)."""
    rules = []
    for n in sorted({r["n_max"] for r in measured}):
        rows = sorted((r for r in measured if r["n_max"] == n), key=lambda r: r["tiles_per_cu"])
        for k, r in enumerate(rows):
            if k + 1 < len(rows):
                cap = int(np.floor(16 * np.sqrt(r["tiles_per_cu"] * rows[k + 1]["tiles_per_cu"])))
            else:
                cap = 1 << 30
            if rules and rules[-1][0] == n and rules[-1][2] == r["best"]:
                rules[-1][1] = cap   # the same width as the previous count: one rule
            else:
                rules.append([int(n), cap, int(r["best"])])
    return rules


def save(table: dict, path: str = TABLE) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1)


def install(table=None, path: str = TABLE) -> list:
    """Load the rules of `table` (a dict as sweep() returns, or None: the file at `path`) into
    the library; returns them.  No file: nothing is installed and the built-in rule holds."""
    from ._nwgrad_amd import _set_lane_table
    if table is None:
        if not os.path.exists(path):
            return []
        with open(path) as f:
            table = json.load(f)
    rules = [tuple(int(x) for x in r) for r in table["rules"]]
    _set_lane_table(rules)
    return rules


def clear() -> None:
    """Remove the table: every batch planned afterwards takes the built-in rule."""
    from ._nwgrad_amd import _set_lane_table
    _set_lane_table([])


if __name__ == "__main__":
    t = sweep()
    save(t)
    print(json.dumps(t["rules"]))
