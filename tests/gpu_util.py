"""Helpers of the GPU tests: batches from fixtures and workloads."""
import sys

import numpy as np

import nwgrad_amd as nw


def params_of(alphabet, matrix, gaps):
    g = [float(x) for x in gaps]
    return nw.AlignParams(np.asarray(matrix, np.float64), alphabet=str(alphabet), gap_open_a=g[0],
                          gap_extend_a=g[1], gap_open_b=g[2], gap_extend_b=g[3])


def batch(alphabet, matrix, gaps, codes_a, off_a, codes_b, off_b, *, precision="float32", gap_model="affine",
          mode="global", grad_mode="hard", fill="interpair"):
    cls = nw.SeqPairBatch if precision == "float32" else nw.SeqPairBatchDouble
    b = cls.from_codes(codes_a, off_a, codes_b, off_b, params_of(alphabet, matrix, gaps), gap_model=gap_model,
                       mode=mode, grad_mode=grad_mode)
    b.fill = fill
    return b


def from_workload(w, **kw):
    kw.setdefault("gap_model", w.gap_model)
    kw.setdefault("mode", w.mode)
    kw.setdefault("grad_mode", w.grad_mode)
    return batch(w.alphabet, w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, **kw)


def grads_rows(b):
    mats, gaps = b.grads()
    return np.concatenate([mats.reshape(len(b), -1), gaps.reshape(len(b), -1)], axis=1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# The soft tolerance: scores |s - s_ref| <= rtol * max(1, |s_ref|); every gradient entry
# |g - g_ref| <= rtol * |g_ref| + floor * max_k |g_ref[pair, k]|, element by element (the floor,
# relative to the pair's largest expected count, covers counts that are rounding noise).
SOFT_RTOL = 1e-5
SOFT_FLOOR = 1e-7


def soft_errors(s, g, s_ref, g_ref, floor=SOFT_FLOOR):
    """(score error, worst element-wise gradient error in units of the tolerance's scale,
    worst norm-relative gradient error): the element-wise figure is max over entries of
    |g - g_ref| / (|g_ref| + floor * max_k |g_ref[pair]|)."""
    s, g, s_ref, g_ref = (np.asarray(x, np.float64) for x in (s, g, s_ref, g_ref))
    if len(s) == 0:
        return 0.0, 0.0, 0.0
    es = float(np.max(np.abs(s - s_ref) / np.maximum(1.0, np.abs(s_ref))))
    scale = np.max(np.abs(g_ref), axis=1, keepdims=True)
    denom = np.abs(g_ref) + floor * scale
    d = np.abs(g - g_ref)
    ee = float(np.max(np.where(denom > 0, d / np.where(denom > 0, denom, 1.0), np.where(d > 0, np.inf, 0.0))))
    en = float(np.max(np.max(d, axis=1) / np.maximum(scale[:, 0], 1e-300)))
    return es, ee, en


def assert_soft_close(s, g, s_ref, g_ref, rtol=SOFT_RTOL, floor=SOFT_FLOOR, label=""):
    es, ee, en = soft_errors(s, g, s_ref, g_ref, floor)
    say(f"  {label}max errors: score {es:.3e}, gradient element-wise {ee:.3e}, norm-relative {en:.3e} "
        f"(rtol {rtol:g}, floor {floor:g})")
    assert es <= rtol and ee <= rtol, (es, ee, en)
    return es, ee, en
