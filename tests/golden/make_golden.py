"""Writes the golden fixtures of tests/golden/ from the reference engine built from its own
sources (oracle/_ref/libnwgref.so; see oracle/Makefile.ref).  Run in the build container:

    make -f oracle/Makefile.ref && python tests/golden/make_golden.py        # all sets
    python tests/golden/make_golden.py long                                  # the long-pair sets only
    python tests/golden/make_golden.py tracks                                # the tracked sets only
    python tests/golden/make_golden.py soft_tracks                           # the tracked soft sets only
    python tests/golden/make_golden.py paths                                 # guides, paths, banded, align()
    python tests/golden/make_golden.py hirschberg                            # traceback="hirschberg*" sets

Each fixture is an .npz (no pickles) holding the inputs -- alphabet, matrix, gaps, problem
type, encoded pairs -- and the reference's outputs: per-pair scores and gradients, and for the
hard sets also compute_grad() and weighted_grad(w) for seeded weights.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from nwgrad_amd import workloads as wl  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def run_ref(w, precision, grad_mode, temperature=1.0, soft_precision="float32"):
    sa, sb = w.strings()
    b = ref.RefBatch(sa, sb, w.alphabet, w.matrix, w.gaps, double=(precision == "float64"),
                     gap_model=w.gap_model, mode=w.mode, grad_mode=grad_mode, n_threads=1)
    if grad_mode == "soft":
        b.set_soft(temperature=temperature, precision=soft_precision)
    total = b.score_and_grad()
    scores = b.scores()
    out = {"total": np.array(total), "scores": scores}
    if grad_mode != "none":
        mats, gaps = b.grads()
        out["grads"] = np.concatenate([mats.reshape(len(scores), -1), gaps], axis=1)
        out["compute_grad"] = b.compute_grad()
        wts = wl.rng(w.name, "weights").normal(size=len(scores))
        out["weights"] = wts
        out["weighted_grad"] = b.weighted_grad(wts)
    return out


def save(name, w, precision, grad_mode, res, **extra):
    path = os.path.join(OUT, name + ".npz")
    np.savez_compressed(path, alphabet=np.array(w.alphabet), matrix=w.matrix, gaps=w.gaps,
                        gap_model=np.array(w.gap_model), mode=np.array(w.mode), precision=np.array(precision),
                        grad_mode=np.array(grad_mode), codes_a=w.codes_a, off_a=w.off_a, codes_b=w.codes_b,
                        off_b=w.off_b, **res, **{k: np.array(v) for k, v in extra.items()})
    print("wrote", path, len(w.off_a) - 1, "pairs")


def edge_pairs(w):
    """Append the reference tests' edge cases: empty A, empty B, both empty, length 1."""
    extra_a = [[], [1], [], [0], [2, 2, 2]]
    extra_b = [[3, 1], [], [], [0], [2]]
    ca = [w.codes_a[w.off_a[i]:w.off_a[i + 1]] for i in range(w.n_pairs)]
    cb = [w.codes_b[w.off_b[i]:w.off_b[i + 1]] for i in range(w.n_pairs)]
    ca += [np.array(x, np.uint8) for x in extra_a]
    cb += [np.array(x, np.uint8) for x in extra_b]
    pa, oa = wl._pack(ca)
    pb, ob = wl._pack(cb)
    w.codes_a, w.off_a, w.codes_b, w.off_b = pa, oa, pb, ob
    return w


def long_sets():
    """Pairs wider than one wavefront's columns (len(B) 1,025 to 2,600: two or three column
    panels on the GPU), every problem type, hard and score-only, plus one soft set."""
    for gm in ("affine", "linear"):
        for mode in ("global", "local", "glocal"):
            w = wl.random_pairs(f"golden_long_{gm}_{mode}", 5, wl.AA, (150, 400), (1025, 2600),
                                gaps=(10.5, 0.75, 9.25, 1.5) if gm == "affine" else (0.0, 1.5, 0.0, 2.0),
                                gap_model=gm, mode=mode, related=0.0)
            for prec in ("float32", "float64"):
                save(f"hard_long_{gm}_{mode}_{prec}", w, prec, "hard", run_ref(w, prec, "hard"))
            save(f"score_long_{gm}_{mode}_float32", w, "float32", "none", run_ref(w, "float32", "none"))
    w = wl.random_pairs("golden_long_soft", 4, wl.AA, (1100, 1400), (1100, 1400), gaps=(11.0, 1.0, 11.0, 1.0),
                        gap_model="affine", mode="global", related=1.0)
    save("soft_long_affine_global_float32", w, "float32", "soft",
         run_ref(w, "float32", "soft", soft_precision="float32"))


def track_sets():
    """Tracked batches: K = 3 parameter sets, one per residue of A (segment tracks: runs of
    one set), explicit or default gap tracks, every problem type, hard and score-only.  Sets
    0 and 2 share gap_extend_b, so the column-0 border's runs of equal extends cross sets."""
    K = 3
    for alpha, name, lens in ((wl.AA, "prot", ((0, 90), (0, 140))), ("ACGT", "dna", ((0, 30), (0, 60)))):
        n = len(alpha)
        r = wl.rng(f"golden_tracks_{name}", "params")
        if n > 4:
            mats = np.stack([wl.random_matrix(r, n, -3, 2, 2, 6) for _ in range(K)])
        else:
            mats = np.stack([np.where(np.eye(4) > 0, 2.0, -1.0), np.where(np.eye(4) > 0, 1.0, -2.0),
                             np.where(np.eye(4) > 0, 3.0, -1.0)])
        gaps_aff = np.array([[10.5, 0.75, 9.25, 1.5], [8.0, 1.25, 11.0, 0.5], [12.0, 2.0, 6.5, 1.5]])
        if n == 4:
            gaps_aff = np.array([[2.0, 1.0, 3.0, 1.0], [1.0, 2.0, 2.0, 0.5], [3.0, 0.5, 1.0, 1.0]])
        for gm in ("affine", "linear"):
            gaps = gaps_aff if gm == "affine" else gaps_aff * np.array([0.0, 1.0, 0.0, 1.0])
            for mode in ("global", "local", "glocal"):
                w = wl.random_pairs(f"golden_tracks_{name}_{gm}_{mode}", 24, alpha, lens[0], lens[1],
                                    gap_model=gm, mode=mode, related=0.5)
                rt = wl.rng(w.name, "tracks")
                tracks = []
                for q in range(w.n_pairs):   # segment tracks: runs of one set
                    m = int(w.off_a[q + 1] - w.off_a[q])
                    t = np.zeros(m, np.uint16)
                    pos = 0
                    while pos < m:
                        ln = int(rt.integers(1, 12))
                        t[pos:pos + ln] = rt.integers(0, K)
                        pos += ln
                    tracks.append(t)
                tracks = np.concatenate(tracks) if tracks else np.zeros(0, np.uint16)
                explicit = mode != "glocal"
                gtracks = (rt.integers(0, K, size=int(w.off_a[-1]) + w.n_pairs).astype(np.uint16)
                           if explicit else None)
                for prec in ("float32", "float64"):
                    for gd in ("hard", "none"):
                        if gd == "none" and prec == "float64":
                            continue
                        res = run_ref_tracked(w, mats, gaps, tracks, gtracks, prec, gd)
                        extra = {"matrices": mats, "gaps_k": gaps, "tracks": tracks}
                        if gtracks is not None:
                            extra["gap_tracks"] = gtracks
                        tag = "hard" if gd == "hard" else "score"
                        save_tracked(f"{tag}_tracks_{name}_{gm}_{mode}_{prec}", w, prec, gd, res, extra)


def soft_track_sets():
    """Tracked soft batches: K = 3 parameter sets over segment tracks (explicit gap tracks in
    Global and Local, the default in Glocal), protein and DNA, both gap models, every mode,
    float32 and float64.  A is a mutated copy of B (the scaled pass needs related pairs)."""
    K = 3
    for alpha, name, lens in ((wl.AA, "prot", (0, 120)), ("ACGT", "dna", (0, 60))):
        n = len(alpha)
        r = wl.rng(f"golden_soft_tracks_{name}", "params")
        if n > 4:
            mats = np.stack([wl.random_matrix(r, n, -3, 2, 2, 6) for _ in range(K)])
        else:
            mats = np.stack([np.where(np.eye(4) > 0, 2.0, -1.0), np.where(np.eye(4) > 0, 1.0, -2.0),
                             np.where(np.eye(4) > 0, 3.0, -1.0)])
        gaps_aff = np.array([[10.5, 0.75, 9.25, 1.5], [8.0, 1.25, 11.0, 0.5], [12.0, 2.0, 6.5, 1.5]])
        if n == 4:
            gaps_aff = np.array([[2.0, 1.0, 3.0, 1.0], [1.0, 2.0, 2.0, 0.5], [3.0, 0.5, 1.0, 1.0]])
        for gm in ("affine", "linear"):
            gaps = gaps_aff if gm == "affine" else gaps_aff * np.array([0.0, 1.0, 0.0, 1.0])
            for mode in ("global", "local", "glocal"):
                rs = wl.rng(f"golden_soft_tracks_{name}_{gm}_{mode}", "seqs")
                A, B = [], []
                for _ in range(24):
                    b_ = rs.integers(0, n, size=int(rs.integers(lens[0], lens[1] + 1))).astype(np.uint8)
                    a_ = wl.mutate(rs, b_, n, 0.2, 0.02, 2.0) if len(b_) else b_
                    A.append(a_)
                    B.append(b_)
                ca, oa = wl._pack(A)
                cb, ob = wl._pack(B)
                w = wl.Workload(f"golden_soft_tracks_{name}_{gm}_{mode}", alpha, mats[0], gaps[0], gm, mode, "soft",
                                ca, oa, cb, ob)
                rt = wl.rng(w.name, "tracks")
                tracks = []
                for q in range(w.n_pairs):   # segment tracks: runs of one set
                    m = int(w.off_a[q + 1] - w.off_a[q])
                    t = np.zeros(m, np.uint16)
                    pos = 0
                    while pos < m:
                        ln = int(rt.integers(1, 12))
                        t[pos:pos + ln] = rt.integers(0, K)
                        pos += ln
                    tracks.append(t)
                tracks = np.concatenate(tracks) if tracks else np.zeros(0, np.uint16)
                gtracks = (rt.integers(0, K, size=int(w.off_a[-1]) + w.n_pairs).astype(np.uint16)
                           if mode != "glocal" else None)
                for prec in ("float32", "float64"):
                    res = run_ref_tracked(w, mats, gaps, tracks, gtracks, prec, "soft")
                    extra = {"matrices": mats, "gaps_k": gaps, "tracks": tracks}
                    if gtracks is not None:
                        extra["gap_tracks"] = gtracks
                    save_tracked(f"soft_tracks_{name}_{gm}_{mode}_{prec}", w, prec, "soft", res, extra)


def run_ref_tracked(w, mats, gaps, tracks, gtracks, precision, grad_mode):
    sa, sb = w.strings()
    b = ref.RefBatch(sa, sb, w.alphabet, mats, gaps, double=(precision == "float64"), gap_model=w.gap_model,
                     mode=w.mode, grad_mode=grad_mode, n_threads=1, tracks=tracks, gap_tracks=gtracks)
    if grad_mode == "soft":
        b.set_soft(temperature=1.0, precision=precision)
    total = b.score_and_grad()
    scores = b.scores()
    out = {"total": np.array(total), "scores": scores}
    if grad_mode != "none":
        m, g = b.grads()
        out["grads"] = np.concatenate([m.reshape(len(scores), -1), g.reshape(len(scores), -1)], axis=1)
        out["compute_grad"] = b.compute_grad()
        wts = wl.rng(w.name, "weights").normal(size=len(scores))
        out["weights"] = wts
        out["weighted_grad"] = b.weighted_grad(wts)
    return out


def save_tracked(name, w, precision, grad_mode, res, extra):
    path = os.path.join(OUT, name + ".npz")
    np.savez_compressed(path, alphabet=np.array(w.alphabet), matrix=extra["matrices"][0], gaps=extra["gaps_k"][0],
                        gap_model=np.array(w.gap_model), mode=np.array(w.mode), precision=np.array(precision),
                        grad_mode=np.array(grad_mode), codes_a=w.codes_a, off_a=w.off_a, codes_b=w.codes_b,
                        off_b=w.off_b, **res, **extra)
    print("wrote", path, len(w.off_a) - 1, "pairs")


def _perturbed(w, seed):
    """A second parameter point near the first, as a training step would move it: matrix entries
    and gaps shifted by multiples of 0.25."""
    r = wl.rng(w.name, seed)
    mat = w.matrix + 0.25 * r.integers(-3, 4, size=w.matrix.shape)
    mat = (mat + mat.T) / 2
    gaps = np.maximum(w.gaps + 0.25 * r.integers(-2, 3, size=4), 0.0)
    if w.gap_model == "linear":
        gaps[0] = gaps[2] = 0.0
    return mat, gaps


def _paths_of(b, w):
    """Every pair's stored path as the reference renders it."""
    al_a, al_b, cig, cig_x, coords, coff = [], [], [], [], [], [0]
    for i in range(w.n_pairs):
        cap = int(w.off_a[i + 1] - w.off_a[i] + w.off_b[i + 1] - w.off_b[i])
        a, bb, cc, c0, c1 = b.path(i, cap)
        al_a.append(a)
        al_b.append(bb)
        cig.append(c0)
        cig_x.append(c1)
        coords.append(cc)
        coff.append(coff[-1] + cc.shape[1])
    return {"aligned_a": np.array(al_a), "aligned_b": np.array(al_b), "cigars": np.array(cig),
            "cigars_ext": np.array(cig_x), "coords": np.concatenate(coords, axis=1), "coords_off": np.array(coff)}


def _hard_results(b, n):
    scores = b.scores()
    mats, gaps = b.grads()
    return scores, np.concatenate([mats.reshape(n, -1), gaps.reshape(n, -1)], axis=1)


def path_set(name, w, prec, bands=(3,), traceback="auto", hb_cutoff=None):
    """One hard batch: score_and_grad(keep_paths=True) under (matrix, gaps) -- scores, gradients,
    guides and the stored paths -- then, for each bandwidth, set_params(the perturbed point) and
    banded_grad(bw, keep_paths=True) from those guides: scores, gradients, new guides, cigars.

    traceback: the full pass's.  The long sets name "pointers": their pairs of more than 512
    rows would take Grid under the reference's `auto` on a SIMD build -- Pointers' result by
    the reference's contract, guides included (aligner.hpp:89-100, guide_j_grid_) -- but this
    scalar build has no Grid kernels, and its fallback writes guides in Hirschberg's convention
    (the column a path enters a row at, and n at row m)."""
    sa, sb = w.strings()
    mat1, gaps1 = _perturbed(w, "p1")
    out = {"matrix1": mat1, "gaps1": gaps1, "bands": np.array(bands)}
    if hb_cutoff is not None:
        out["hb_cutoff"] = np.array(hb_cutoff)
        out["traceback"] = np.array(traceback)
    b = ref.RefBatch(sa, sb, w.alphabet, w.matrix, w.gaps, double=(prec == "float64"), gap_model=w.gap_model,
                     mode=w.mode, grad_mode="hard", n_threads=1, traceback=traceback)
    if hb_cutoff is not None:
        b.set_hb_cutoff(hb_cutoff)
    out["total"] = np.array(b.score_and_grad_keep(True))
    out["scores"], out["grads"] = _hard_results(b, w.n_pairs)
    out["guides"] = b.guides(w.off_a)
    out.update(_paths_of(b, w))
    for bw in bands:
        b = ref.RefBatch(sa, sb, w.alphabet, w.matrix, w.gaps, double=(prec == "float64"), gap_model=w.gap_model,
                         mode=w.mode, grad_mode="hard", n_threads=1, traceback=traceback)
        if hb_cutoff is not None:
            b.set_hb_cutoff(hb_cutoff)
        b.score_and_grad_keep(False)
        b.set_params(mat1, gaps1)
        out[f"band{bw}_total"] = np.array(b.banded_grad(bw, True))
        out[f"band{bw}_scores"], out[f"band{bw}_grads"] = _hard_results(b, w.n_pairs)
        out[f"band{bw}_guides"] = b.guides(w.off_a)
        out[f"band{bw}_cigars"] = _paths_of(b, w)["cigars"]
    save(name, w, prec, "hard", out)


def bandscore_sets():
    """Banded passes of score-only batches: the banded Viterbi's scores, guides and paths around
    the eager guides of a full pass (batch_engine.hpp:2188-2205)."""
    for gm in ("affine", "linear"):
        for mode in ("global", "local", "glocal"):
            w = wl.random_pairs(f"golden_bandscore_{gm}_{mode}", 24, wl.AA, (0, 150), (0, 220),
                                gaps=(10.5, 0.75, 9.25, 1.5) if gm == "affine" else (0.0, 1.25, 0.0, 1.75),
                                gap_model=gm, mode=mode, related=0.6)
            sa, sb = w.strings()
            mat1, gaps1 = _perturbed(w, "p1")
            out = {"matrix1": mat1, "gaps1": gaps1, "bands": np.array((2, 16))}
            for bw in (2, 16):
                b = ref.RefBatch(sa, sb, w.alphabet, w.matrix, w.gaps, gap_model=gm, mode=mode, grad_mode="none",
                                 n_threads=1)
                b.score_and_grad()
                b.set_params(mat1, gaps1)
                out[f"band{bw}_total"] = np.array(b.banded_grad(bw, True))
                out[f"band{bw}_scores"] = b.scores()
                out[f"band{bw}_guides"] = b.guides(w.off_a)
                out[f"band{bw}_cigars"] = _paths_of(b, w)["cigars"]
            save(f"bandscore_prot_{gm}_{mode}_float32", w, "float32", "none", out)


def bandsoft_sets():
    """Banded passes of soft batches (scaled_or_log, float32): after a full pass,
    set_params(the perturbed point) and banded_grad(bw, keep_paths=True) -- the banded
    forward-backward's scores and gradients around the eager guides, and the banded Viterbi's
    new guides and paths (batch_engine.hpp:2188-2205, 2125-2134) -- and streaming align() with a
    band around the diagonal (scores and the summed gradient)."""
    for gm in ("affine", "linear"):
        for mode in ("global", "local", "glocal"):
            w = wl.random_pairs(f"golden_bandsoft_{gm}_{mode}", 20, wl.AA, (1, 120), (1, 160),
                                gaps=(11.0, 1.0, 11.0, 1.0) if gm == "affine" else (0.0, 1.5, 0.0, 1.5),
                                gap_model=gm, mode=mode, related=0.6)
            sa, sb = w.strings()
            mat1, gaps1 = _perturbed(w, "p1")
            out = {"matrix1": mat1, "gaps1": gaps1, "bands": np.array((2, 16))}
            for bw in (2, 16):
                b = ref.RefBatch(sa, sb, w.alphabet, w.matrix, w.gaps, gap_model=gm, mode=mode, grad_mode="soft",
                                 n_threads=1)
                b.set_soft(precision="float32", impl="scaled_or_log")
                b.score_and_grad()
                out["guides0"] = b.guides(w.off_a)   # the eager guides the band follows
                b.set_params(mat1, gaps1)
                out[f"band{bw}_total"] = np.array(b.banded_grad(bw, True))
                out[f"band{bw}_scores"], out[f"band{bw}_grads"] = _hard_results(b, w.n_pairs)
                out[f"band{bw}_guides"] = b.guides(w.off_a)
                out[f"band{bw}_cigars"] = _paths_of(b, w)["cigars"]
            b = ref.RefBatch(sa[:2], sb[:2], w.alphabet, w.matrix, w.gaps, gap_model=gm, mode=mode, grad_mode="soft",
                             n_threads=1)
            b.set_soft(precision="float32", impl="scaled_or_log")
            out["align8_scores"], out["align8_grad"] = b.align(sa, sb, mat1, gaps1, band=8)
            save(f"bandsoft_prot_{gm}_{mode}_float32", w, "float32", "soft", out)


def path_sets():
    for gm in ("affine", "linear"):
        for mode in ("global", "local", "glocal"):
            for prec in ("float32", "float64"):
                w = wl.random_pairs(f"golden_paths_dna_{gm}_{mode}", 40, "ACGT", (0, 40), (0, 80),
                                    gaps=(2.0, 1.0, 3.0, 1.0) if gm == "affine" else (0.0, 1.0, 0.0, 2.0),
                                    gap_model=gm, mode=mode, related=0.6)
                path_set(f"paths_dna_{gm}_{mode}_{prec}", edge_pairs(w), prec, bands=(1, 4))
                w = wl.random_pairs(f"golden_paths_prot_{gm}_{mode}", 24, wl.AA, (1, 150), (1, 220),
                                    gaps=(10.5, 0.75, 9.25, 1.5) if gm == "affine" else (0.0, 1.25, 0.0, 1.75),
                                    gap_model=gm, mode=mode, related=0.6)
                path_set(f"paths_prot_{gm}_{mode}_{prec}", w, prec, bands=(2, 16))
            # long pairs (column panels on the GPU)
            w = wl.random_pairs(f"golden_paths_long_{gm}_{mode}", 4, wl.AA, (300, 900), (1025, 2400),
                                gaps=(10.5, 0.75, 9.25, 1.5) if gm == "affine" else (0.0, 1.25, 0.0, 1.75),
                                gap_model=gm, mode=mode, related=0.0)
            path_set(f"paths_long_{gm}_{mode}_float32", w, "float32", bands=(8,), traceback="pointers")
    bandscore_sets()
    bandsoft_sets()
    # the guides of soft batches (the eager guide: the hard fill under the params scored)
    for mode in ("global", "local", "glocal"):
        w = wl.random_pairs(f"golden_softguide_{mode}", 16, wl.AA, (1, 90), (1, 120), gaps=(11.0, 1.0, 11.0, 1.0),
                            mode=mode, related=0.5)
        sa, sb = w.strings()
        b = ref.RefBatch(sa, sb, w.alphabet, w.matrix, w.gaps, gap_model="affine", mode=mode, grad_mode="soft",
                         n_threads=1)
        b.set_soft(precision="float32", impl="scaled_or_log")
        total = b.score_and_grad()
        save(f"softguide_prot_affine_{mode}_float32", w, "float32", "soft",
             {"total": np.array(total), "scores": b.scores(), "guides": b.guides(w.off_a)})
    # streaming align(): every pair's score and the summed gradient, nothing kept
    for gm in ("affine", "linear"):
        for mode in ("global", "local", "glocal"):
            w = wl.random_pairs(f"golden_align_{gm}_{mode}", 20, wl.AA, (1, 120), (1, 160),
                                gaps=(10.5, 0.75, 9.25, 1.5) if gm == "affine" else (0.0, 1.25, 0.0, 1.75),
                                gap_model=gm, mode=mode, related=0.5)
            sa, sb = w.strings()
            mat1, gaps1 = _perturbed(w, "align")
            res = {"matrix1": mat1, "gaps1": gaps1}
            for gd in ("hard", "none", "soft"):
                if gd == "soft" and gm == "linear":
                    continue
                b = ref.RefBatch(sa[:2], sb[:2], w.alphabet, w.matrix, w.gaps, gap_model=gm, mode=mode, grad_mode=gd,
                                 n_threads=1)
                if gd == "soft":
                    b.set_soft(precision="float32", impl="scaled_or_log")
                sc, gr = b.align(sa, sb, mat1, gaps1)
                res[f"{gd}_scores"], res[f"{gd}_grad"] = sc, gr
            save(f"align_prot_{gm}_{mode}_float32", w, "float32", "hard", res)


def hb_sets():
    """Global affine hard batches under traceback="hirschberg" (the exact carry) and
    "hirschberg_pmax" (the prefix-max carry) at small base-case cutoffs, so that pairs of a
    few hundred rows recurse several levels (aligner.hpp:4456-4518) while the shortest run as
    Pointers: scores, gradients, the guides in Hirschberg's convention, stored paths, and a
    banded pass that follows those guides.  Integer gaps (11/1) make ties common, where the
    recursion's path differs from Pointers'; fractional gaps make the two carries round
    differently."""
    for prec in ("float32", "float64"):
        for tb, cutoff in (("hirschberg", 5), ("hirschberg_pmax", 24)):
            for tag, gaps in (("int", (11.0, 1.0, 11.0, 1.0)), ("frac", (10.5, 0.7, 9.3, 1.1))):
                w = wl.random_pairs(f"golden_hb_{tb}_{tag}", 18, wl.AA, (1, 420), (1, 460), gaps=gaps,
                                    related=0.6)
                w = edge_pairs(w)
                path_set(f"hb_{tb}_{tag}_{prec}", w, prec, bands=(4,), traceback=tb, hb_cutoff=cutoff)


def main():
    if sys.argv[1:] == ["hirschberg"]:
        hb_sets()
        return
    if sys.argv[1:] == ["bandsoft"]:
        bandsoft_sets()
        return
    if sys.argv[1:] == ["bandscore"]:
        bandscore_sets()
        return
    if sys.argv[1:] == ["paths"]:
        path_sets()
        return
    if sys.argv[1:] == ["long"]:
        long_sets()
        return
    if sys.argv[1:] == ["tracks"]:
        track_sets()
        return
    if sys.argv[1:] == ["soft_tracks"]:
        soft_track_sets()
        return
    for gm in ("affine", "linear"):
        for mode in ("global", "local", "glocal"):
            for prec in ("float32", "float64"):
                # DNA, +2/-1, small gaps: tie-heavy (the decisive case for the tie rules)
                w = wl.random_pairs(f"golden_dna_{gm}_{mode}", 48, "ACGT", (0, 30), (0, 70),
                                    gaps=(2.0, 1.0, 3.0, 1.0) if gm == "affine" else (0.0, 1.0, 0.0, 2.0),
                                    gap_model=gm, mode=mode, related=0.5)
                w = edge_pairs(w)
                save(f"hard_dna_{gm}_{mode}_{prec}", w, prec, "hard", run_ref(w, prec, "hard"))
                # protein, random integer matrix, non-integral gaps
                w = wl.random_pairs(f"golden_prot_{gm}_{mode}", 32, wl.AA, (1, 120), (1, 200),
                                    gaps=(10.5, 0.75, 9.25, 1.5), gap_model=gm, mode=mode, related=0.5)
                save(f"hard_prot_{gm}_{mode}_{prec}", w, prec, "hard", run_ref(w, prec, "hard"))
            # score-only
            w = wl.random_pairs(f"golden_score_{gm}_{mode}", 32, "ACGT", (1, 30), (1, 70),
                                gaps=(2.0, 1.0, 2.0, 1.0), gap_model=gm, mode=mode, related=0.5)
            save(f"score_dna_{gm}_{mode}_float32", w, "float32", "none", run_ref(w, "float32", "none"))
            # soft (float32 scaled pass and double)
            w = wl.random_pairs(f"golden_soft_{gm}_{mode}", 24, wl.AA, (1, 90), (1, 120),
                                gaps=(11.0, 1.0, 11.0, 1.0), gap_model=gm, mode=mode, related=0.5)
            save(f"soft_prot_{gm}_{mode}_float32", w, "float32", "soft",
                 run_ref(w, "float32", "soft", soft_precision="float32"))
            save(f"soft_prot_{gm}_{mode}_float64", w, "float64", "soft",
                 run_ref(w, "float64", "soft", soft_precision="double"))
    # temperature != 1
    w = wl.random_pairs("golden_soft_temp", 16, "ACGT", (1, 40), (1, 60), gaps=(3.0, 1.0, 3.0, 1.0),
                        gap_model="affine", mode="global", related=0.5)
    save("soft_dna_affine_global_float64_T2", w, "float64", "soft",
         run_ref(w, "float64", "soft", temperature=2.0, soft_precision="double"), temperature=2.0)


if __name__ == "__main__":
    main()
