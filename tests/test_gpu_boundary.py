"""The drop-in boundary as a caller of the reference sees it.

* The reference's constructor, SeqPairBatch(seqs_a, seqs_b, params, ...), over the strings of
  the golden fixtures the reference engine wrote: results equal the fixtures (hard and
  score-only bit for bit, soft within 1e-5 element by element).
* Pair views, batch[i] (py_exports.cpp:895-905): score, grad, sequences, tracks; negative
  indices; IndexError; a view keeps its batch alive.
* The C ABI bound with ctypes exactly as INTEGRATION.md's stub binds it, against the oracle.
"""
import ctypes
import gc
import glob
import os

import numpy as np
import pytest

import nwgrad_amd as nw
import oracle
from nwgrad_amd import workloads as wl
from tests.gpu_util import assert_soft_close, batch, grads_rows, params_of, say

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
GOLDEN = sorted(p for p in glob.glob(os.path.join(HERE, "golden", "*.npz"))
                if os.path.basename(p).split("_")[0] in ("hard", "score", "soft") and "_tracks_" not in p)


def _strings(alphabet, codes, off):
    al = np.frombuffer(str(alphabet).encode(), np.uint8)
    return [al[codes[off[i]:off[i + 1]]].tobytes().decode() for i in range(len(off) - 1)]


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: os.path.basename(p)[:-4])
def test_string_constructor_golden(gpu, path):
    z = np.load(path)
    kind = os.path.basename(path).split("_")[0]
    sa = _strings(z["alphabet"], z["codes_a"], z["off_a"])
    sb = _strings(z["alphabet"], z["codes_b"], z["off_b"])
    cls = nw.SeqPairBatch if str(z["precision"]) == "float32" else nw.SeqPairBatchDouble
    grad_mode = {"hard": "hard", "score": "none", "soft": "soft"}[kind]
    b = cls(sa, sb, params_of(z["alphabet"], z["matrix"], z["gaps"]), gap_model=str(z["gap_model"]),
            mode=str(z["mode"]), grad_mode=grad_mode)
    if "temperature" in z.files:
        b.soft_temperature = float(z["temperature"])
    total = b.score_and_grad()
    if kind == "soft":
        assert_soft_close(b.scores(), grads_rows(b), z["scores"], z["grads"],
                          rtol=1e-5 if str(z["precision"]) == "float32" else 1e-9)
        return
    np.testing.assert_array_equal(b.scores(), z["scores"])
    assert total == float(z["total"])
    if kind == "hard":
        np.testing.assert_array_equal(grads_rows(b), z["grads"])


def test_pair_views(gpu):
    w = wl.random_pairs("views", 40, wl.AA, (0, 90), (0, 120), related=0.5)
    sa, sb = w.strings()
    p = params_of(w.alphabet, w.matrix, w.gaps)
    b = nw.SeqPairBatch(sa, sb, p)
    v = b[3]
    assert v.score is None and v.grad is None and not v.score_valid
    assert v.seq_a == sa[3] and v.seq_b == sb[3] and v.track is None
    b.score_and_grad()
    s, g = b.scores(), grads_rows(b)
    for i in (0, 3, 17, 39, -1, -40):
        vi = b[i]
        k = i % len(b)
        assert vi.index == k and vi.score == s[k]
        gd = vi.grad.to_dict()
        row = np.concatenate([gd["matrix"].ravel(), [gd["gap_open_a"], gd["gap_extend_a"], gd["gap_open_b"],
                                                     gd["gap_extend_b"]]])
        np.testing.assert_array_equal(row, g[k])
        assert vi.seq_a == sa[k] and vi.seq_b == sb[k]
    with pytest.raises(IndexError):
        b[40]
    with pytest.raises(IndexError):
        b[-41]
    del b
    gc.collect()
    assert v.score == s[3]   # the view kept its batch
    # a tracked batch's views carry their tracks
    K = 2
    mats = np.stack([w.matrix, w.matrix + 1.0])
    pk = nw.AlignParams.from_arrays(mats, np.array([w.gaps, w.gaps + 0.5]), alphabet=w.alphabet)
    trk = [np.arange(len(a)) % K for a in sa]
    bt = nw.SeqPairBatch(sa, sb, pk, tracks=trk)
    assert list(bt[5].track) == list(trk[5]) and len(bt[5].gap_track) == len(sa[5]) + 1
    bt.score_and_grad()
    assert bt[5].grad.num_sets == K


def test_ctypes_stub(gpu):
    """INTEGRATION.md's ctypes stub, line for line, then the scores and the blocked sum against the
    oracle, and the status / message convention of a failing call."""
    w = wl.random_pairs("ctypes_stub", 300, wl.AA, (20, 300), (20, 300), related=0.5)
    codes_a, off_a, codes_b, off_b = w.codes_a, w.off_a, w.codes_b, w.off_b
    matrix = np.ascontiguousarray(w.matrix)
    gaps4 = np.ascontiguousarray(w.gaps, np.float64)
    n_pairs = w.n_pairs

    lib = ctypes.CDLL(nw.lib_path())

    class Problem(ctypes.Structure):
        _fields_ = [("precision", ctypes.c_int32), ("gap_model", ctypes.c_int32),
                    ("mode", ctypes.c_int32), ("grad_mode", ctypes.c_int32)]

    ctx = ctypes.c_void_p(); lib.nwg_init(0, ctypes.byref(ctx))
    P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))  # noqa: E731
    batch_h = ctypes.c_void_p()
    rc = lib.nwg_batch_create(ctx, ctypes.byref(Problem(0, 1, 0, 1)), 20, ctypes.c_int64(n_pairs),
                              P(codes_a, ctypes.c_uint8), P(off_a, ctypes.c_int64),
                              P(codes_b, ctypes.c_uint8), P(off_b, ctypes.c_int64), ctypes.byref(batch_h))
    assert rc == 0
    lib.nwg_batch_set_params(batch_h, P(matrix, ctypes.c_double), P(gaps4, ctypes.c_double))
    total = ctypes.c_double(); rc = lib.nwg_batch_run(batch_h, ctypes.byref(total))
    assert rc == 0

    s_ref, g_ref = oracle.hard(w.matrix, w.gaps, codes_a, off_a, codes_b, off_b)
    assert total.value == float(np.cumsum(s_ref)[-1])
    scores = np.zeros(n_pairs)
    assert lib.nwg_batch_scores(batch_h, P(scores, ctypes.c_double)) == 0
    np.testing.assert_array_equal(scores, s_ref)
    out = np.zeros(404)
    assert lib.nwg_batch_weighted_grad(batch_h, None, ctypes.c_int64(0), P(out, ctypes.c_double)) == 0
    np.testing.assert_array_equal(out, oracle.weighted_sum(g_ref))
    # a failing call: a status, and the message through nwg_last_error
    lib.nwg_last_error.restype = ctypes.c_char_p
    one = np.ones(n_pairs - 1)
    rc = lib.nwg_batch_weighted_grad(batch_h, P(one, ctypes.c_double), ctypes.c_int64(n_pairs - 1),
                                     P(out, ctypes.c_double))
    assert rc == 1 and b"one weight per pair" in lib.nwg_last_error(ctx)
    lib.nwg_batch_free(batch_h)
    lib.nwg_destroy(ctx)
    say("  ctypes stub: total", total.value)
