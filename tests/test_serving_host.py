"""Host side of the serving loop, no GPU needed: job validation, the no-device failure (a job
fails loudly, nothing runs on the CPU), and the serving workload's generator against the
configuration's closed forms."""
import numpy as np
import pytest

from nwgrad_amd import workloads as wl
from tests.gpu_util import params_of


def _server(**kw):
    from nwgrad_amd.serving import Server
    return Server(params_of(wl.AA, wl.random_matrix(wl.rng("serving_host", "m")), [11, 1, 11, 1]), **kw)


def test_submit_validates_jobs():
    s = _server()
    ok = (np.array([1, 2, 3], np.uint8), np.array([0, 3]), np.array([4, 5], np.uint8), np.array([0, 2]))
    with pytest.raises(ValueError):
        s.submit("viterbi", *ok)
    with pytest.raises(ValueError):   # code outside the 20-letter alphabet
        s.submit("hard", np.array([1, 25, 3], np.uint8), ok[1], ok[2], ok[3])
    with pytest.raises(ValueError):   # offsets must start at 0
        s.submit("hard", ok[0], np.array([1, 3]), ok[2], ok[3])
    with pytest.raises(ValueError):   # offsets past the codes
        s.submit("hard", ok[0], np.array([0, 9]), ok[2], ok[3])
    with pytest.raises(ValueError):
        s.submit_strings("hard", ["ACD"], ["AXB"])
    with pytest.raises(ValueError):
        _server(slots_per_kind=0)
    assert s.submit("hard", *ok) == 0 and s.waiting == 1


def test_no_device_fails_loudly():
    import nwgrad_amd as nw
    if nw.device_available():
        pytest.skip("a GPU is present")
    s = _server()
    j = s.submit_strings("score", ["ACDE"], ["ACDF"])
    s.poll()
    r = s.result(j)
    assert s.state(j) == "failed" and "HIP device" in r["error"]
    s.release(j)
    with pytest.raises(IndexError):
        s.result(j)


def test_serving_load_matches_its_configuration():
    """200 requests/s for 60 s; pairs per request log-U[1, 4096] (mean (4097 - 1) / ln 4097 =
    492.3); lengths log-U[30, 5000] (mean (5001 - 30) / ln(5001 / 30) = 972.5); kinds 70 / 20
    / 10 % -- checked on the generator's draws with sampling tolerances."""
    load = wl.ServingLoad()
    n = len(load)
    assert abs(n - 12000) < 4 * np.sqrt(12000)
    gaps = np.diff(load.times)
    assert abs(gaps.mean() - 1 / 200) < 4 * (1 / 200) / np.sqrt(n)
    kinds = {k: float(np.mean(load.kinds == k)) for k in ("score", "hard", "soft")}
    for k, p in (("score", 0.7), ("hard", 0.2), ("soft", 0.1)):
        assert abs(kinds[k] - p) < 4 * np.sqrt(p * (1 - p) / n)
    r = wl.rng("serving_host", "draws")
    pairs = wl._log_uniform_int(r, 1, 4096, 200000)
    lens = wl._log_uniform_int(r, 30, 5000, 200000)
    assert abs(pairs.mean() - 4096 / np.log(4097)) < 0.02 * 492
    assert abs(lens.mean() - 4971 / np.log(5001 / 30)) < 0.01 * 972
    assert pairs.min() >= 1 and pairs.max() <= 4096 and lens.min() >= 30 and lens.max() <= 5000
    q = load.request(3)
    assert q.kind == load.kinds[3] and q.off_a[0] == 0 and q.off_a[-1] == len(q.codes_a)
    assert q.codes_a.max() < 20 and q.codes_b.max() < 20
    q2 = load.request(3)
    np.testing.assert_array_equal(q.codes_b, q2.codes_b)
