"""AlignParams with K parameter sets (no GPU needed): construction, set access, stacking,
arithmetic and the dict form, against the reference's API (py_exports.cpp:1400-1535)."""
import numpy as np
import pytest

import nwgrad_amd as nw
from nwgrad_amd import workloads as wl


def _sets(K=3):
    r = wl.rng("params_sets", "m")
    mats = np.stack([wl.random_matrix(r) for _ in range(K)])
    gaps = np.arange(4 * K, dtype=np.float64).reshape(K, 4) + 0.5
    return mats, gaps


def test_from_arrays_and_set_access():
    mats, gaps = _sets()
    p = nw.AlignParams.from_arrays(mats, gaps, wl.AA)
    assert p.num_sets == 3
    np.testing.assert_array_equal(p.matrices, mats)
    np.testing.assert_array_equal(p.gaps, gaps)
    np.testing.assert_array_equal(p.matrix.to_matrix(), mats[0])     # plain fields: set 0
    assert (p.gap_open_a, p.gap_extend_a, p.gap_open_b, p.gap_extend_b) == tuple(gaps[0])
    s2 = p[2]
    assert s2.num_sets == 1
    np.testing.assert_array_equal(s2.matrix.to_matrix(), mats[2])
    assert s2.gap_extend_b == gaps[2, 3] and p[-1].gap_open_a == gaps[2, 0]
    with pytest.raises(IndexError):
        p[3]
    p[1] = p[0]
    np.testing.assert_array_equal(p.matrices[1], mats[0])
    d = p.to_dict()
    assert d["num_sets"] == 3 and d["matrices"].shape == (3, 20, 20) and d["gaps"].shape == (3, 4)
    with pytest.raises(ValueError):
        nw.AlignParams.from_arrays(mats, gaps[:2], wl.AA)
    with pytest.raises(ValueError):
        nw.AlignParams.from_arrays(mats[:, :3, :3], gaps, wl.AA)


def test_stack_and_arithmetic():
    mats, gaps = _sets()
    parts = [nw.AlignParams(mats[k], alphabet=wl.AA, gap_open_a=gaps[k, 0], gap_extend_a=gaps[k, 1],
                            gap_open_b=gaps[k, 2], gap_extend_b=gaps[k, 3]) for k in range(3)]
    p = nw.AlignParams.stack(parts)
    np.testing.assert_array_equal(p.matrices, mats)
    np.testing.assert_array_equal(p.gaps, gaps)
    q = p + p * 2.0
    np.testing.assert_array_equal(q.matrices, 3 * mats)
    np.testing.assert_array_equal(q.gaps, 3 * gaps)
    np.testing.assert_array_equal((-p).gaps, -gaps)
    with pytest.raises(ValueError):       # different set counts do not combine
        p + parts[0]
    p.gaps = gaps * 0
    assert not p.gaps.any()
