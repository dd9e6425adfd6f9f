"""nwg_hard_noalloc, the entry point that allocates nothing, through the C ABI: its scores and
gradient rows equal nwg_batch_run's on the same pairs (bit for bit), over two calls in a row,
the second with a smaller input (the library's scratch keeps the first call's contents), for
hard and score-only passes of every gap model and mode in float32 and float64.  The caller's
buffers are torch tensors on the device."""
import ctypes

import numpy as np
import pytest

import nwgrad_amd as nw
import oracle
from nwgrad_amd import workloads as wl

pytestmark = pytest.mark.gpu

c_p = ctypes.c_void_p
P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))  # noqa: E731


class Problem(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("gap_model", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("grad_mode", ctypes.c_int32)]


def _lib():
    lib = ctypes.CDLL(nw.lib_path())
    lib.nwg_last_error.restype = ctypes.c_char_p
    lib.nwg_hard_noalloc.argtypes = [c_p, ctypes.POINTER(Problem), ctypes.c_int, ctypes.c_int64, c_p, c_p, c_p, c_p,
                                     ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), c_p, c_p]
    return lib


def _batch_path(lib, ctx, prob, na, w):
    b = ctypes.c_void_p()
    assert lib.nwg_batch_create(ctx, ctypes.byref(prob), na, ctypes.c_int64(w.n_pairs), P(w.codes_a, ctypes.c_uint8),
                                P(w.off_a, ctypes.c_int64), P(w.codes_b, ctypes.c_uint8), P(w.off_b, ctypes.c_int64),
                                ctypes.byref(b)) == 0, lib.nwg_last_error(ctx)
    mat = np.ascontiguousarray(w.matrix, np.float64)
    gaps = np.ascontiguousarray(w.gaps, np.float64)
    assert lib.nwg_batch_set_params(b, P(mat, ctypes.c_double), P(gaps, ctypes.c_double)) == 0
    assert lib.nwg_batch_run(b, None) == 0, lib.nwg_last_error(ctx)
    s = np.zeros(w.n_pairs)
    assert lib.nwg_batch_scores(b, P(s, ctypes.c_double)) == 0
    g = None
    if prob.grad_mode == 1:
        mats = np.zeros((w.n_pairs, na * na))
        gp = np.zeros((w.n_pairs, 4))
        assert lib.nwg_batch_grads(b, P(mats, ctypes.c_double), P(gp, ctypes.c_double)) == 0
        g = np.concatenate([mats, gp], axis=1)
    lib.nwg_batch_free(b)
    return s, g


def _noalloc(lib, ctx, prob, na, w):
    import torch
    dev = torch.device("cuda")
    ca = torch.from_numpy(np.ascontiguousarray(w.codes_a)).to(dev)
    cb = torch.from_numpy(np.ascontiguousarray(w.codes_b)).to(dev)
    oa = torch.from_numpy(np.ascontiguousarray(w.off_a)).to(dev)
    ob = torch.from_numpy(np.ascontiguousarray(w.off_b)).to(dev)
    scores = torch.full((w.n_pairs,), float("nan"), dtype=torch.float64, device=dev)
    grads = torch.full((w.n_pairs, na * na + 4), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    mat = np.ascontiguousarray(w.matrix, np.float64)
    gaps = np.ascontiguousarray(w.gaps, np.float64)
    oa_h = np.ascontiguousarray(w.off_a, np.int64)
    ob_h = np.ascontiguousarray(w.off_b, np.int64)
    rc = lib.nwg_hard_noalloc(ctx, ctypes.byref(prob), na, w.n_pairs, ca.data_ptr(), oa.data_ptr(), cb.data_ptr(),
                              ob.data_ptr(), P(oa_h, ctypes.c_int64), P(ob_h, ctypes.c_int64),
                              P(mat, ctypes.c_double), P(gaps, ctypes.c_double), scores.data_ptr(),
                              grads.data_ptr() if prob.grad_mode == 1 else None)
    assert rc == 0, lib.nwg_last_error(ctx)
    return scores.cpu().numpy(), (grads.cpu().numpy() if prob.grad_mode == 1 else None)


CASES = [(gm, mode, prec, grad) for gm in ("affine", "linear") for mode in ("global", "local", "glocal")
         for prec in ("float32", "float64") for grad in ("hard", "none")]


@pytest.mark.parametrize("gm,mode,prec,grad", CASES)
def test_noalloc_matches_batch_path(gpu, gm, mode, prec, grad):
    lib = _lib()
    ctx = ctypes.c_void_p()
    assert lib.nwg_init(0, ctypes.byref(ctx)) == 0
    prob = Problem(0 if prec == "float32" else 1, 1 if gm == "affine" else 0, {"global": 0, "local": 1, "glocal": 2}[mode],
                   1 if grad == "hard" else 0)
    big = wl.random_pairs(f"noalloc_big_{gm}_{mode}", 3000, wl.AA, (0, 400), (0, 600), gap_model=gm, mode=mode,
                          related=0.5)
    small = wl.random_pairs(f"noalloc_small_{gm}_{mode}", 300, wl.AA, (0, 120), (0, 1024), gap_model=gm, mode=mode,
                            related=0.5)
    print(f"case noalloc {gm} {mode} {prec} {grad}: {big.n_pairs} then {small.n_pairs} pairs", flush=True)
    for w in (big, small):
        s_b, g_b = _batch_path(lib, ctx, prob, 20, w)
        s_n, g_n = _noalloc(lib, ctx, prob, 20, w)
        np.testing.assert_array_equal(s_n, s_b)
        if grad == "hard":
            np.testing.assert_array_equal(g_n, g_b)
    s_ref, g_ref = oracle.hard(small.matrix, small.gaps, small.codes_a, small.off_a, small.codes_b, small.off_b,
                               gap_model=gm, mode=mode, precision=prec, with_grad=grad == "hard")
    np.testing.assert_array_equal(s_n, s_ref)
    if grad == "hard":
        np.testing.assert_array_equal(g_n, g_ref)
    lib.nwg_destroy(ctx)


def test_noalloc_limits(gpu):
    """Sizes past the scratch's limits are refused with a message, and nothing runs."""
    lib = _lib()
    ctx = ctypes.c_void_p()
    assert lib.nwg_init(0, ctypes.byref(ctx)) == 0
    prob = Problem(0, 1, 0, 1)
    wide = wl.random_pairs("noalloc_wide", 2, wl.AA, (10, 20), (1025, 1030))
    z = ctypes.c_void_p(1)
    oa_h = np.ascontiguousarray(wide.off_a)
    ob_h = np.ascontiguousarray(wide.off_b)
    mat = np.ascontiguousarray(wide.matrix)
    gaps = np.ascontiguousarray(wide.gaps)
    rc = lib.nwg_hard_noalloc(ctx, ctypes.byref(prob), 20, 2, z, z, z, z, P(oa_h, ctypes.c_int64),
                              P(ob_h, ctypes.c_int64), P(mat, ctypes.c_double), P(gaps, ctypes.c_double), z, z)
    assert rc == 1 and b"at most 1024 columns" in lib.nwg_last_error(ctx)
    rc = lib.nwg_hard_noalloc(ctx, ctypes.byref(prob), 20, (1 << 20) + 1, z, z, z, z, P(oa_h, ctypes.c_int64),
                              P(ob_h, ctypes.c_int64), P(mat, ctypes.c_double), P(gaps, ctypes.c_double), z, z)
    assert rc == 1 and b"at most" in lib.nwg_last_error(ctx)
    lib.nwg_destroy(ctx)
