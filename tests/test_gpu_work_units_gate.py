"""Work-units regression gate on the device, the counterpart of the
reference's strict PDLP gate (tests/perf/check_regression.py,
run_pdlp_lp_regression_gate.py:67-71: no regression allowed): the seeded
solves of lp_cases.work_unit_workload() through the single-LP, packed
batch, bounds-variant batch and serving paths must end with the committed
baseline's status and at most its work units (the baseline,
tests/golden/device_work_units.json, is written on the device by
tests/golden/make_device_work_units.py).  Every reduction on the device
runs in a fixed order, so an unchanged tree reproduces the baseline
exactly; a change that needs more iterations anywhere fails here."""
import json
import os

import pytest

from conftest import TIGHT
from lp_cases import GOLDEN, run_work_unit_workload
from mipx_amd import native

pytestmark = pytest.mark.gpu


def test_device_work_units_do_not_regress():
    with open(os.path.join(GOLDEN, "device_work_units.json")) as f:
        base = json.load(f)["runs"]
    got = run_work_unit_workload(native, TIGHT)
    assert set(got) == set(base), sorted(set(got) ^ set(base))
    worse, better = [], []
    for name, b in sorted(base.items()):
        g = got[name]
        assert g["status"] == b["status"], (name, g, b)
        if g["work_units"] > b["work_units"]:
            worse.append((name, b["iterations"], g["iterations"]))
        elif g["work_units"] < b["work_units"]:
            better.append((name, b["iterations"], g["iterations"]))
    print(f"{len(base)} solves; fewer work units: {better}")
    assert not worse, worse
