"""GPU tests of the C++ host mirror (mipx::PdlpSolver) and the CLI."""
import os
import re
import subprocess

import pytest

from mipx_amd import lpgen
from oracle import binding as ob

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mipx_amd", "bin")
# The test LPs are below the reference CLI's default GPU gate (512 rows,
# 10,000 nonzeros, src/cli/main.cpp:79-80).
OPEN_GATE = ["--gpu-min-rows", "0", "--gpu-min-nnz", "0"]


def test_cpp_host_tests():
    r = subprocess.run([os.path.join(BIN, "host_tests")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "0 failed" in r.stdout


def test_cli_stable_lines_match_oracle():
    r = subprocess.run([os.path.join(BIN, "mipx_solve"), "--gen", "c0", "--seed", "101",
                        "--rows", "300", "--cols", "750", "--pdlp", "--print-backend", "--quiet"] + OPEN_GATE,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0
    lines = dict(re.findall(r"^(PDLP backend|Status|Objective|Iterations|Work units|Time): (.*)$",
                            r.stdout, flags=re.M))
    assert lines["PDLP backend"] == "GPU"
    assert lines["Status"] == "Optimal"
    lp = lpgen.c0_plain(seed=101, m=300, n=750).lp
    ro, _, _ = ob.solve(lp, primal_tol=1e-6, dual_tol=1e-6, optimality_tol=1e-6, verbose=0)
    obj = float(lines["Objective"])
    assert abs(obj - ro["objective"]) <= 1e-6 * (1 + abs(ro["objective"]))
    it = int(lines["Iterations"])
    assert float(lines["Work units"]) == pytest.approx(it * (4.0 * lp.nnz + 2.0 * (300 + 750)))


def _stable_lines(out):
    return dict(re.findall(r"^(PDLP backend|Status|Objective|Iterations|Work units): (.*)$", out,
                           flags=re.M))


def test_cli_mps_round_trip(tmp_path):
    """A generated LP written as MPS (--write-mps) and solved from the file
    gives the same stable lines as solving it directly."""
    mps = str(tmp_path / "c0.mps")
    base = [os.path.join(BIN, "mipx_solve"), "--pdlp", "--print-backend", "--quiet"] + OPEN_GATE
    r1 = subprocess.run(base + ["--gen", "c0", "--seed", "101", "--rows", "300", "--cols", "750",
                                "--write-mps", mps], capture_output=True, text=True, timeout=120)
    r2 = subprocess.run(base + [mps], capture_output=True, text=True, timeout=120)
    print(r1.stdout, r2.stdout)
    assert r1.returncode == 0 and r2.returncode == 0
    a, b = _stable_lines(r1.stdout), _stable_lines(r2.stdout)
    assert a["Status"] == "Optimal" and a == b


def test_cli_reads_reference_tiny_mps():
    """tests/data/tiny.mps of the reference read as its LP relaxation (the
    fields the reference's MPS test expects): x2 + x3 = 7 cannot hold with
    x2 <= 1 and x3 <= 5; neither the oracle nor the GPU path finds a
    certificate within 20,000 iterations, and both stop at the limit."""
    import numpy as np
    from mipx_amd.lp import from_dense
    from lp_cases import GOLDEN
    r = subprocess.run([os.path.join(BIN, "mipx_solve"), os.path.join(GOLDEN, "tiny.mps"),
                        "--pdlp", "--print-backend", "--quiet", "--max-iter", "20000"] + OPEN_GATE,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0
    lines = _stable_lines(r.stdout)
    inf = float("inf")
    lp = from_dense(np.array([[2.0, 1.0, 0.0], [1.0, 0.0, 1.0], [0.0, 1.0, 1.0]]),
                    [1.0, 3.0, 2.0], [0.0, 0.0, 1.0], [8.0, 1.0, 5.0],
                    [-inf, 5.0, 7.0], [10.0, inf, 7.0])
    ro, _, _ = ob.solve(lp, verbose=0, max_iter=20000, primal_tol=1e-6, dual_tol=1e-6,
                        optimality_tol=1e-6)
    names = {0: "Optimal", 1: "Infeasible", 2: "Unbounded", 3: "Iteration limit"}
    assert lines["Status"] == names[ro["status"]] == "Iteration limit"
    assert int(lines["Iterations"]) == ro["iterations"] == 20000


def test_cli_presolve_default_vs_no_presolve(tmp_path):
    """The reference CLI presolves by default (src/cli/main.cpp:455-467) and
    reports the objective of the postsolved columns (560-568).  On an LP with
    fixed columns, singleton and empty rows and empty columns: with presolve
    the GPU solve is Optimal and within 1e-6 of the exact optimum once solved
    to 1e-8; without it the fixed columns stall the reference's dual-residual
    test (pdlp.cpp:1127-1142) and the oracle stops at the iteration limit
    just as the device path does."""
    from lp_cases import highs_objective, planted_presolve_lp, write_free_mps
    lp, _, _ = planted_presolve_lp()
    path = str(tmp_path / "planted.mps")
    write_free_mps(lp, path)
    base = [os.path.join(BIN, "mipx_solve"), path, "--pdlp", "--print-backend", "--quiet",
            "--max-iter", "40000"] + OPEN_GATE
    r1 = subprocess.run(base + ["--tol", "1e-8", "--max-iter", "2000000"], capture_output=True,
                        text=True, timeout=300)
    r2 = subprocess.run(base + ["--no-presolve"], capture_output=True, text=True, timeout=300)
    print(r1.stdout, r2.stdout)
    a, b = _stable_lines(r1.stdout), _stable_lines(r2.stdout)
    assert "Presolve: removed 19 rows, 24 cols" in r1.stdout
    assert a["Status"] == "Optimal" and a["PDLP backend"] == "GPU"
    opt = highs_objective(lp)
    assert abs(float(a["Objective"]) - opt) <= 1e-6 * (1 + abs(opt))
    ro, _, _ = ob.solve(lp, max_iter=40000, primal_tol=1e-6, dual_tol=1e-6, optimality_tol=1e-6,
                        verbose=0)
    assert b["Status"] == "Iteration limit" and ro["status"] == 3


def test_presolved_gpu_solve_postsolved_and_certified():
    """The reduced LP solved on the GPU (FP64 and mixed precision), primal
    and dual postsolved: certified on the original LP, objective within 1e-6
    of the oracle's solve of the same reduced LP (FP64)."""
    from conftest import TIGHT
    from lp_cases import planted_presolve_lp
    from mipx_amd import native
    from mipx_amd.presolve import Presolved
    lp, _, _ = planted_presolve_lp()
    pre = Presolved(lp)
    ro, _, _ = ob.solve(pre.reduced, **TIGHT)
    for prec in (0, 1):
        r, xr, yr = native.solve(pre.reduced, precision=prec, **TIGHT)
        assert r["status"] == 0
        x, y = pre.postsolve(xr, yr)
        k = ob.kkt_check(lp, x, y)
        print(prec, r["objective"], ro["objective"], k)
        assert k["rel_primal"] <= 1.01e-6 and k["rel_dual"] <= 1.01e-6 and k["rel_gap"] <= 1.01e-6
        assert k["dual_sign_viol"] == 0.0 and k["bound_viol"] == 0.0
        assert abs(lp.obj @ x - r["objective"]) <= 1e-9 * (1 + abs(r["objective"]))
        bar = 1e-6 if prec == 0 else 5e-6
        assert abs(r["objective"] - ro["objective"]) <= bar * (1 + abs(ro["objective"]))
