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
    return np.concatenate([mats.reshape(len(b), -1), gaps], axis=1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()
