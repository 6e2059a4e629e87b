"""The reference's GPU gate (src/lp/pdlp.cpp:862-869) in the C++ mirror and
the CLI.  No device work: every LP here is rejected by the gate before the
library is called, or only asked whether it would pass."""
import os
import re
import subprocess

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mipx_amd", "bin")


def test_mirror_gate_checks():
    r = subprocess.run([os.path.join(BIN, "host_tests"), "--gate-only"], capture_output=True,
                       text=True, timeout=60)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and " 0 failed" in r.stdout


def _lines(out):
    return dict(re.findall(r"^(PDLP backend|Status|Objective|Iterations|Work units|Error detail): "
                           r"(.*)$", out, flags=re.M))


def test_cli_below_gate_reports_cpu_backend_and_error():
    """300 rows is below the reference CLI's default --gpu-min-rows 512; with
    --no-gpu any LP is.  The reference would solve these on its CPU loop,
    which this library does not contain: status Error, backend CPU."""
    base = [os.path.join(BIN, "mipx_solve"), "--gen", "c0", "--seed", "101", "--rows", "300",
            "--cols", "750", "--pdlp", "--print-backend", "--quiet"]
    for extra in ([], ["--no-gpu", "--gpu-min-rows", "0", "--gpu-min-nnz", "0"],
                  ["--gpu-min-rows", "0", "--gpu-min-nnz", "100000"]):
        r = subprocess.run(base + extra, capture_output=True, text=True, timeout=60)
        print(r.stdout)
        lines = _lines(r.stdout)
        assert r.returncode == 0
        assert lines["PDLP backend"] == "CPU" and lines["Status"] == "Error"
        assert "GPU gate" in lines["Error detail"]
        assert float(lines["Objective"]) == 0.0 and int(lines["Iterations"]) == 0


def test_cli_presolve_proves_infeasibility_without_device(tmp_path):
    """An LP that presolve proves infeasible (a singleton row against a
    column bound) is reported Infeasible before any device work."""
    import numpy as np
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from lp_cases import write_free_mps
    from mipx_amd.lp import from_dense
    lp = from_dense(np.array([[1.0, 0.0], [1.0, 1.0]]), [1.0, 1.0], [0.0, 0.0], [5.0, 5.0],
                    [6.0, -np.inf], [np.inf, 4.0])
    path = str(tmp_path / "inf.mps")
    write_free_mps(lp, path)
    r = subprocess.run([os.path.join(BIN, "mipx_solve"), path, "--pdlp", "--quiet"],
                       capture_output=True, text=True, timeout=60)
    print(r.stdout, r.stderr)
    lines = _lines(r.stdout)
    assert r.returncode == 0 and lines["Status"] == "Infeasible"
    assert "Presolve: infeasible" in r.stdout
