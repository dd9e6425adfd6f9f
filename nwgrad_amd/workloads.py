"""Synthetic inputs of the benchmark configurations, seeded and reproducible.

Seeds follow the reference tests' convention (tests/Python/test_soft_float32.py:35-37):
``zlib.crc32(f"{config}|{stream}")`` into numpy's PCG64.  Pairs come back encoded: ``uint8``
codes plus ``int64`` offsets, pair i being ``codes[off[i]:off[i+1]]``.

Configurations (cells = sum of len(A) * len(B)):

* ``protein_global_hard``: 4,096 protein pairs, lengths U{100..300}, random symmetric integer
  matrix (off-diagonal U{-4..3}, diagonal U{4..11}), global affine 11/1.
* ``dna_local_short``: 1,048,576 DNA pairs, len(A) U{18..25}, len(B) U{40..60}; half carry
  reverse-complement(A) planted in B with 10 % point substitutions; +2/-1, local linear ge 2.
* ``protein_soft_homologs``: 32,768 homolog + 32,768 decoy pairs, len(A) log-normal (median 250,
  sigma 0.6, clipped to [50, 1500]); homolog B = A with substitution p 0.3 and indels p 0.02
  (geometric length, mean 3); global affine 11/1, soft.
* ``long_pairs``: protein pairs of one length, B = A mutated (substitution 0.1, indel 0.01,
  geometric mean 5).
"""
from __future__ import annotations

import dataclasses
import zlib

import numpy as np

AA = "ACDEFGHIKLMNPQRSTVWY"
DNA = "ACGT"


def rng(config: str, stream: str) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(zlib.crc32(f"{config}|{stream}".encode())))


@dataclasses.dataclass
class Workload:
    name: str
    alphabet: str
    matrix: np.ndarray          # (n, n) float64
    gaps: np.ndarray            # go_a, ge_a, go_b, ge_b
    gap_model: str
    mode: str
    grad_mode: str
    codes_a: np.ndarray
    off_a: np.ndarray
    codes_b: np.ndarray
    off_b: np.ndarray
    tracks: np.ndarray | None = None   # K > 1 sets: one set per residue of A, flat

    @property
    def n_pairs(self) -> int:
        return len(self.off_a) - 1

    @property
    def cells(self) -> int:
        la = np.diff(self.off_a)
        lb = np.diff(self.off_b)
        return int(np.sum(la.astype(np.int64) * lb.astype(np.int64)))

    def strings(self, idx=None):
        """Pairs as (list of str, list of str) (for the string constructor / reference)."""
        idx = range(self.n_pairs) if idx is None else idx
        al = np.frombuffer(self.alphabet.encode(), np.uint8)
        sa = [al[self.codes_a[self.off_a[i]:self.off_a[i + 1]]].tobytes().decode() for i in idx]
        sb = [al[self.codes_b[self.off_b[i]:self.off_b[i + 1]]].tobytes().decode() for i in idx]
        return sa, sb

    def subset(self, idx) -> "Workload":
        idx = np.asarray(idx)
        ca, oa = _gather(self.codes_a, self.off_a, idx)
        cb, ob = _gather(self.codes_b, self.off_b, idx)
        return dataclasses.replace(self, codes_a=ca, off_a=oa, codes_b=cb, off_b=ob)


def _gather(codes, off, idx):
    lens = (off[1:] - off[:-1])[idx]
    o = np.zeros(len(idx) + 1, np.int64)
    o[1:] = np.cumsum(lens)
    c = np.empty(int(o[-1]), np.uint8)
    for k, i in enumerate(idx):
        c[o[k]:o[k + 1]] = codes[off[i]:off[i + 1]]
    return c, o


def _pack(seqs):
    off = np.zeros(len(seqs) + 1, np.int64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    codes = np.concatenate(seqs).astype(np.uint8) if seqs else np.zeros(0, np.uint8)
    return codes, off


def random_matrix(r: np.random.Generator, n: int = 20, lo=-4, hi=3, dlo=4, dhi=11) -> np.ndarray:
    """Symmetric integer matrix: off-diagonal U{lo..hi}, diagonal U{dlo..dhi}."""
    m = r.integers(lo, hi + 1, size=(n, n)).astype(np.float64)
    m = np.triu(m, 1)
    m = m + m.T
    m[np.diag_indices(n)] = r.integers(dlo, dhi + 1, size=n)
    return m


def mutate(r: np.random.Generator, a: np.ndarray, n_alpha: int, sub_p: float, indel_p: float,
           indel_mean: float) -> np.ndarray:
    """Point substitutions (uniform replacement by a different letter) and indel events (50/50
    insertion / deletion, geometric length with the given mean) per residue."""
    L = len(a)
    b = a.astype(np.int64)
    subs = r.random(L) < sub_p
    b[subs] = (b[subs] + r.integers(1, n_alpha, size=int(subs.sum()))) % n_alpha
    ev = np.nonzero(r.random(L) < indel_p)[0]
    if len(ev) == 0:
        return b.astype(np.uint8)
    lens = r.geometric(1.0 / indel_mean, size=len(ev))
    ins = r.random(len(ev)) < 0.5
    keep = np.ones(L, bool)
    for pos, ln, is_ins in zip(ev, lens, ins):
        if not is_ins:
            keep[pos:pos + ln] = False
    pieces = []
    prev = 0
    for pos, ln, is_ins in zip(ev, lens, ins):
        if is_ins:
            pieces.append(b[prev:pos][keep[prev:pos]])
            pieces.append(r.integers(0, n_alpha, size=int(ln)))
            prev = pos
    pieces.append(b[prev:][keep[prev:]])
    return np.concatenate(pieces).astype(np.uint8)


def protein_global_hard(n_pairs: int = 4096) -> Workload:
    name = "protein_global_hard"
    r = rng(name, "matrix")
    mat = random_matrix(r)
    rs = rng(name, "seqs")
    la = rs.integers(100, 301, size=n_pairs)
    lb = rs.integers(100, 301, size=n_pairs)
    A = [rs.integers(0, 20, size=int(x)).astype(np.uint8) for x in la]
    B = [rs.integers(0, 20, size=int(x)).astype(np.uint8) for x in lb]
    ca, oa = _pack(A)
    cb, ob = _pack(B)
    return Workload(name, AA, mat, np.array([11.0, 1.0, 11.0, 1.0]), "affine", "global", "hard",
                    ca, oa, cb, ob)


def dna_local_short(n_pairs: int = 1 << 20) -> Workload:
    name = "dna_local_short"
    r = rng(name, "seqs")
    la = r.integers(18, 26, size=n_pairs)
    lb = r.integers(40, 61, size=n_pairs)
    total_a = int(la.sum())
    total_b = int(lb.sum())
    ca = r.integers(0, 4, size=total_a).astype(np.uint8)
    cb = r.integers(0, 4, size=total_b).astype(np.uint8)
    oa = np.zeros(n_pairs + 1, np.int64)
    ob = np.zeros(n_pairs + 1, np.int64)
    oa[1:] = np.cumsum(la)
    ob[1:] = np.cumsum(lb)
    planted = np.nonzero(r.random(n_pairs) < 0.5)[0]
    for i in planted:
        a = ca[oa[i]:oa[i + 1]]
        rc = (3 - a[::-1]).astype(np.uint8)
        subs = r.random(len(rc)) < 0.1
        rc[subs] = (rc[subs] + r.integers(1, 4, size=int(subs.sum()))) % 4
        nb = lb[i]
        off = int(r.integers(0, nb - len(rc) + 1))
        cb[ob[i] + off: ob[i] + off + len(rc)] = rc
    mat = np.full((4, 4), -1.0)
    np.fill_diagonal(mat, 2.0)
    return Workload(name, DNA, mat, np.array([0.0, 2.0, 0.0, 2.0]), "linear", "local", "hard",
                    ca, oa, cb, ob)


def protein_soft_homologs(n_homologs: int = 32768, n_decoys: int = 32768) -> Workload:
    name = "protein_soft_homologs"
    r = rng(name, "matrix")
    mat = random_matrix(r)
    rs = rng(name, "seqs")

    def length():
        return int(np.clip(np.round(rs.lognormal(np.log(250.0), 0.6)), 50, 1500))

    A, B = [], []
    for _ in range(n_homologs):
        a = rs.integers(0, 20, size=length()).astype(np.uint8)
        A.append(a)
        B.append(mutate(rs, a, 20, 0.3, 0.02, 3.0))
    for _ in range(n_decoys):
        A.append(rs.integers(0, 20, size=length()).astype(np.uint8))
        B.append(rs.integers(0, 20, size=length()).astype(np.uint8))
    ca, oa = _pack(A)
    cb, ob = _pack(B)
    return Workload(name, AA, mat, np.array([11.0, 1.0, 11.0, 1.0]), "affine", "global", "soft",
                    ca, oa, cb, ob)


def long_pairs(n_pairs: int, length: int, *, grad_mode: str = "hard") -> Workload:
    name = f"long_pairs_{length}"
    r = rng(name, "matrix")
    mat = random_matrix(r)
    rs = rng(name, "seqs")
    A, B = [], []
    for _ in range(n_pairs):
        a = rs.integers(0, 20, size=length).astype(np.uint8)
        A.append(a)
        B.append(mutate(rs, a, 20, 0.1, 0.01, 5.0))
    ca, oa = _pack(A)
    cb, ob = _pack(B)
    return Workload(name, AA, mat, np.array([11.0, 1.0, 11.0, 1.0]), "affine", "global", grad_mode,
                    ca, oa, cb, ob)


def random_pairs(name: str, n_pairs: int, alphabet: str, len_a, len_b, *, matrix=None, gaps=(11.0, 1.0, 11.0, 1.0),
                 gap_model="affine", mode="global", grad_mode="hard", related: float = 0.0) -> Workload:
    """Small seeded test batches: lengths U{len_a} / U{len_b} (inclusive ranges); with
    probability `related` B is a mutated copy of A instead (substitution 0.2, indels 0.05)."""
    r = rng(name, "seqs")
    n = len(alphabet)
    if matrix is None:
        matrix = random_matrix(rng(name, "matrix"), n, -3, 2, 2, 6) if n > 4 else \
            np.where(np.eye(n) > 0, 2.0, -1.0)
    A, B = [], []
    for _ in range(n_pairs):
        a = r.integers(0, n, size=int(r.integers(len_a[0], len_a[1] + 1))).astype(np.uint8)
        if r.random() < related and len(a):
            b = mutate(r, a, n, 0.2, 0.05, 2.0)
        else:
            b = r.integers(0, n, size=int(r.integers(len_b[0], len_b[1] + 1))).astype(np.uint8)
        A.append(a)
        B.append(b)
    ca, oa = _pack(A)
    cb, ob = _pack(B)
    return Workload(name, alphabet, np.asarray(matrix, np.float64), np.asarray(gaps, np.float64), gap_model, mode,
                    grad_mode, ca, oa, cb, ob)


@dataclasses.dataclass
class Request:
    """One alignment job of the serving workload: pairs of one kind, arriving at time t (s)."""
    t: float
    kind: str                   # "score", "hard" or "soft"
    codes_a: np.ndarray
    off_a: np.ndarray
    codes_b: np.ndarray
    off_b: np.ndarray
    tracks: np.ndarray | None = None   # K > 1 sets: one set per residue of A, flat

    @property
    def n_pairs(self) -> int:
        return len(self.off_a) - 1

    @property
    def cells(self) -> int:
        return int(np.sum(np.diff(self.off_a) * np.diff(self.off_b)))


def _log_uniform_int(r: np.random.Generator, lo: int, hi: int, size) -> np.ndarray:
    """Integers log-uniform on [lo, hi]: floor(exp(U(ln lo, ln(hi + 1))))."""
    x = np.floor(np.exp(r.uniform(np.log(lo), np.log(hi + 1), size=size))).astype(np.int64)
    return np.clip(x, lo, hi)


class ServingLoad:
    """The serving configuration: Poisson arrivals at ``rate`` requests/s for ``duration`` s;
    pairs per request log-uniform on [1, max_pairs] (mean about 492 at 4,096); len(A) and len(B)
    log-uniform on [lo, hi], independently (mean about 972 at [30, 5000]); kinds score / hard /
    soft in the proportions ``mix``.  Residues are i.i.d. uniform protein letters: a request's
    A (and B) codes are one contiguous slice of a seeded pool, cut into its pairs, so making a
    request costs O(pairs); a soft request's B is its A with 30 % point substitutions (a
    homolog: the scaled soft pass needs the dynamic range of related sequences, as the
    reference's does).  Request k is a deterministic function of (name, k)."""

    def __init__(self, rate: float = 200.0, duration: float = 60.0, max_pairs: int = 4096, lo: int = 30,
                 hi: int = 5000, mix=(0.7, 0.2, 0.1), name: str = "serving", sets: int = 1,
                 soft_homologs: bool = True):
        self.name, self.max_pairs, self.lo, self.hi = name, max_pairs, lo, hi
        self.sets, self.soft_homologs = int(sets), bool(soft_homologs)
        r = rng(name, "arrivals")
        t, times = 0.0, []
        while True:
            t += r.exponential(1.0 / rate)
            if t >= duration:
                break
            times.append(t)
        self.times = np.array(times)
        self.kinds = r.choice(np.array(["score", "hard", "soft"]), size=len(times), p=np.asarray(mix, float))
        self.pool = rng(name, "pool").integers(0, 20, size=(1 << 25) + max_pairs * hi).astype(np.uint8)
        self.matrix = random_matrix(rng(name, "matrix"))
        self.gaps = np.array([11.0, 1.0, 11.0, 1.0])
        # K parameter sets: set 0 the matrix above; sets 1.. random matrices of the same law with
        # gaps drawn around 11 / 1
        rs = rng(name, "sets")
        self.matrices = np.stack([self.matrix] + [random_matrix(rs) for _ in range(self.sets - 1)])
        self.set_gaps = np.stack([self.gaps] + [np.array([rs.integers(8, 14), rs.uniform(0.5, 1.5),
                                                          rs.integers(8, 14), rs.uniform(0.5, 1.5)])
                                                for _ in range(self.sets - 1)])

    def __len__(self) -> int:
        return len(self.times)

    def _slices(self, r, lens):
        off = np.zeros(len(lens) + 1, np.int64)
        np.cumsum(lens, out=off[1:])
        s0 = int(r.integers(0, len(self.pool) - off[-1]))
        return self.pool[s0:s0 + off[-1]], off

    def request(self, k: int) -> Request:
        r = rng(self.name, f"request{k}")
        n = int(_log_uniform_int(r, 1, self.max_pairs, 1)[0])
        la = _log_uniform_int(r, self.lo, self.hi, n)
        kind = str(self.kinds[k])
        ca, oa = self._slices(r, la)
        if kind == "soft" and self.soft_homologs:
            cb = ca.copy()
            subs = r.random(len(cb)) < 0.3
            cb[subs] = (cb[subs] + r.integers(1, 20, size=int(subs.sum()))) % 20
            ob = oa.copy()
        else:
            lb = _log_uniform_int(r, self.lo, self.hi, n)
            cb, ob = self._slices(r, lb)
        tracks = segment_tracks(r, oa, self.sets) if self.sets > 1 else None
        return Request(float(self.times[k]), kind, ca, oa, cb.astype(np.uint8), ob, tracks)


def segment_tracks(r: np.random.Generator, off_a: np.ndarray, sets: int, max_segments: int = 8) -> np.ndarray:
    """Per-residue tracks of A in uniform-random contiguous segments: pair p's A (m residues) is
    cut at U{1..min(max_segments, m)} - 1 uniform points in [1, m) (coinciding cuts merge), and
    each segment takes a set drawn uniformly from [0, sets).  Flat, off_a[-1] entries."""
    off = np.asarray(off_a, np.int64)
    la = int(off[-1])
    if la == 0:
        return np.zeros(0, np.uint16)
    m = np.diff(off)
    starts = np.zeros(la, np.int64)          # 1 where a segment starts
    nz = m > 0
    starts[off[:-1][nz]] = 1
    nseg = np.where(nz, r.integers(1, np.minimum(max_segments, np.maximum(m, 1)) + 1), 0)
    ncut = np.maximum(nseg - 1, 0)
    owner = np.repeat(np.arange(len(m)), ncut)
    if owner.size:
        pos = off[owner] + 1 + np.floor(r.random(owner.size) * (m[owner] - 1)).astype(np.int64)
        starts[pos] = 1
    seg = np.cumsum(starts) - 1
    return r.integers(0, sets, size=int(seg[-1]) + 1).astype(np.uint16)[seg]
