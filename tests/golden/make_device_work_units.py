"""Writes the baseline of the device work-units gate
(tests/test_gpu_work_units_gate.py): status, iterations and work units
(iterations x (4 nnz + 2 (m + n)), the reference's measure,
src/lp/pdlp.cpp:1716-1719) of the seeded solves of
lp_cases.work_unit_workload() on the device, through the single-LP, packed
batch, bounds-variant batch and serving paths.  The reference gates its own
PDLP work units the same way, with no regression allowed
(tests/perf/run_pdlp_lp_regression_gate.py:67-71).  Run on an MI355X from
the repository root after building; the file goes to the path given (by
default device_work_units.json next to this script)."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))
from conftest import TIGHT  # noqa: E402
from lp_cases import run_work_unit_workload  # noqa: E402
from mipx_amd import native  # noqa: E402


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "device_work_units.json")
    runs = run_work_unit_workload(native, TIGHT)
    with open(path, "w") as f:
        json.dump({"_source": "tests/golden/make_device_work_units.py on one MI355X",
                   "runs": runs}, f, indent=1, sort_keys=True)
    print(f"{len(runs)} solves written to {path}")


if __name__ == "__main__":
    main()
