"""The C ABI library and the Python module load and expose what they declare.  CPU only:
no compute call is made without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import nwgrad_amd
from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "nwgrad_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(nwg_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_abi():
    names = declared_functions()
    for required in ("nwg_init", "nwg_batch_create", "nwg_batch_run", "nwg_batch_weighted_grad",
                     "nwg_batch_grads", "nwg_batch_scores"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(nwgrad_amd.lib_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_surface():
    for name in ("SeqPairBatch", "SeqPairBatchDouble", "AlignParams", "SubstMatrix"):
        assert hasattr(nwgrad_amd, name)
    for meth in ("score_and_grad", "compute_grad", "scores", "grads", "weighted_grad", "set_params"):
        assert hasattr(nwgrad_amd.SeqPairBatch, meth)
        assert hasattr(nwgrad_amd.SeqPairBatchDouble, meth)
    assert isinstance(nwgrad_amd.SeqPairBatch.fill, property) or hasattr(nwgrad_amd.SeqPairBatch, "fill")


def test_align_params_arithmetic():
    m = np.arange(16, dtype=float).reshape(4, 4)
    p = nwgrad_amd.AlignParams(m, alphabet="ACGT", gap_open_a=1, gap_extend_a=2, gap_open_b=3, gap_extend_b=4)
    q = p + p * 2.0
    d = q.to_dict()
    np.testing.assert_array_equal(d["matrix"], 3 * m)
    assert (d["gap_open_a"], d["gap_extend_a"], d["gap_open_b"], d["gap_extend_b"]) == (3, 6, 9, 12)
    r = -(q - p)
    assert r.to_dict()["gap_extend_b"] == -8
    with pytest.raises(ValueError):
        p + nwgrad_amd.AlignParams(np.zeros((20, 20)))
    with pytest.raises(ValueError):
        nwgrad_amd.AlignParams(np.zeros((3, 3)), alphabet="ACGT")
    assert p.matrix.score("A", "C") == 1.0
    with pytest.raises(ValueError):
        p.matrix.score("A", "Z")


def test_constructor_validation_before_device():
    p = nwgrad_amd.AlignParams(np.eye(4), alphabet="ACGT")
    with pytest.raises(ValueError):
        nwgrad_amd.SeqPairBatch(["ACGX"], ["ACG"], p)          # bad character
    with pytest.raises(ValueError):
        nwgrad_amd.SeqPairBatch(["A"], ["A"], p, mode="sideways")
    with pytest.raises(ValueError):
        nwgrad_amd.SeqPairBatch(["A"], ["A"], p, gap_model="cubic")
    with pytest.raises(ValueError):
        nwgrad_amd.SeqPairBatch(["A"], ["A"], p, n_threads=-1)


@pytest.mark.skipif(nwgrad_amd.device_available(), reason="a GPU is present")
def test_no_cpu_fallback_without_gpu():
    p = nwgrad_amd.AlignParams(np.eye(4), alphabet="ACGT")
    with pytest.raises(RuntimeError, match="gfx950"):
        nwgrad_amd.SeqPairBatch(["ACGT"], ["ACG"], p)
