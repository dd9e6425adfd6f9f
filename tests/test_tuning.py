"""The lane-width tuner's host side: rules from measurements, and the library's table setter."""
import json

import numpy as np
import pytest

import nwgrad_amd as nw
from nwgrad_amd import tuning


def test_rules_from_measurements():
    m = [{"n_max": 32, "tiles_per_cu": t, "best": 16} for t in (0.25, 0.5, 1.0, 2.0)]
    m += [{"n_max": 128, "tiles_per_cu": t, "best": b} for t, b in ((0.25, 8), (0.5, 8), (1.0, 8), (2.0, 16), (4.0, 16))]
    r = tuning.rules_from(m)
    # one rule per run of equal widths; the 8-column run reaches the midpoint between 1 and 2 tiles
    assert r == [[32, 1 << 30, 16], [128, int(np.floor(16 * np.sqrt(2.0))), 8], [128, 1 << 30, 16]]


def test_install_and_clear(tmp_path):
    path = tmp_path / "table.json"
    path.write_text(json.dumps({"rules": [[64, 16, 8], [512, 1 << 30, 16]]}))
    assert tuning.install(path=str(path)) == [(64, 16, 8), (512, 1 << 30, 16)]
    with pytest.raises(ValueError, match="lane-width"):
        tuning.install({"rules": [[64, 16, 12]]})
    tuning.clear()
    assert tuning.install(path=str(tmp_path / "absent.json")) == []
    tuning.install()   # back to the package's own table, as at import


def test_committed_table_is_well_formed():
    if not __import__("os").path.exists(tuning.TABLE):
        pytest.skip("no measured table committed")
    t = json.load(open(tuning.TABLE))
    assert t["rules"] and all(len(r) == 3 and r[2] in (8, 16) for r in t["rules"])
    assert [r[0] for r in t["rules"]] == sorted(r[0] for r in t["rules"])
    # every rule is the measured best of its points
    for row in t["measured"]:
        assert row["best"] == max((16, 8), key=lambda w: row[f"gcups_{w}"])
    assert tuning.rules_from(t["measured"]) == t["rules"]
