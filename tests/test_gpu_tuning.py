"""Lane widths of the hard and score-only fills: results identical at 8 and 16 columns a lane,
and the lane-width table's rules followed when a batch is planned."""
import numpy as np
import pytest

from nwgrad_amd import tuning
from nwgrad_amd import workloads as wl
from tests.gpu_util import batch, grads_rows, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("grad_mode", ["hard", "none"])
@pytest.mark.parametrize("mode", ["global", "local"])
def test_widths_bit_exact(gpu, grad_mode, mode):
    w = wl.random_pairs(f"lanes_{grad_mode}_{mode}", 600, wl.AA, (0, 300), (0, 500), mode=mode, related=0.5)
    b = batch(w.alphabet, w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b, mode=mode, grad_mode=grad_mode)
    out = {}
    for c in (16, 8, 0):
        b.columns_per_lane = c
        say(f"  columns {c}: cut at {b.columns_per_lane}")
        assert c == 0 or b.columns_per_lane == c
        t = b.score_and_grad()
        out[c] = (t, b.scores(), grads_rows(b) if grad_mode == "hard" else None)
    for c in (8, 0):
        assert out[c][0] == out[16][0]
        np.testing.assert_array_equal(out[c][1], out[16][1])
        if grad_mode == "hard":
            np.testing.assert_array_equal(out[c][2], out[16][2])


def test_table_rules_followed(gpu):
    w = wl.random_pairs("lanes_table", 64, wl.AA, (50, 100), (50, 100))
    mk = lambda: batch(w.alphabet, w.matrix, w.gaps, w.codes_a, w.off_a, w.codes_b, w.off_b)
    try:
        tuning.install({"rules": [[512, 1 << 30, 16]]})
        assert mk().columns_per_lane == 16
        tuning.install({"rules": [[512, 1 << 30, 8]]})
        assert mk().columns_per_lane == 8
        tuning.install({"rules": [[32, 1 << 30, 16]]})   # no rule for len(B) ~ 75: the built-in rule
        assert mk().columns_per_lane == 8                  # (64 pairs: far fewer tiles than SIMDs)
        b = mk()
        b.columns_per_lane = 16                            # a batch's own setting wins
        assert b.columns_per_lane == 16
        with pytest.raises(ValueError, match="columns per lane"):
            b.columns_per_lane = 12
    finally:
        tuning.clear()
        tuning.install()
