"""CPU checks of bench.py's reporting contract: one GPU only, the roofline's
algorithmic byte count, and the measured HBM traffic it reports for the
kernel the autotuner picked."""
import os
import subprocess
import sys
from types import SimpleNamespace

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr


def test_primal_kernel_bytes_c2():
    """4,096 instances of m=3,000 n=8,000 nnz=48,000: per instance 6n + m
    values (x, x0, l, u read; x, x_r written; y once), shared 2n values and
    A' in CSR."""
    lp = SimpleNamespace(num_rows=3000, num_cols=8000, nnz=48000)
    assert bench.primal_kernel_bytes(lp, 4096, 4) == 836_064_004
    assert bench.primal_kernel_bytes(lp, 4096, 8) == 1_671_904_004
    # one instance, by hand: 6n + m values, then the shared part
    assert bench.primal_kernel_bytes(lp, 1, 8) == (6 * 8000 + 3000) * 8 + 2 * 8000 * 8 \
        + 48000 * 12 + 8001 * 4


def test_traffic_per_tuned_kernel():
    """Every primal variant the autotuner can pick has a measured HBM figure,
    above the algorithmic bytes (part of the gathers is refetched); unknown
    kernels and workloads give None."""
    for prec, sv in (("mixed", 4), ("fp64", 8)):
        t = "float" if prec == "mixed" else "double"
        for cw in (1, 2, 4):
            v = bench.read_traffic("c2", prec, f"kILPrimal<{t},{cw}>")
            assert isinstance(v, int) and v > 0, (prec, cw, v)
            lp = SimpleNamespace(num_rows=3000, num_cols=8000, nnz=48000)
            assert v >= bench.primal_kernel_bytes(lp, 4096, sv)
    assert bench.read_traffic("c2", "mixed", "kILPrimal<float,3>") is None
    assert bench.read_traffic("c9", "mixed", "kILPrimal<float,1>") is None


def test_iteration_bytes_match_baseline_table():
    """B_iter = 2 nnz (s_v + 4) + (m + n + 2) 4 + 9 n s_v + 7 m s_v: C0 is
    1.10 MB (FP64) / 0.66 MB (FP32), C1 500 MB / 304 MB per iteration."""
    assert round(bench.iter_bytes(2000, 5000, 25000, 8) / 1e6, 2) == 1.10
    assert round(bench.iter_bytes(2000, 5000, 25000, 4) / 1e6, 2) == 0.66
    assert round(bench.iter_bytes(1_000_000, 2_000_000, 12_000_000, 8) / 1e6) == 500
    assert round(bench.iter_bytes(1_000_000, 2_000_000, 12_000_000, 4) / 1e6) == 304


def test_whole_solve_bytes_shared():
    """Per instance (7 n + 4 m) s_v per iteration at the precision run, plus
    the shared matrix and vectors once per batched iteration."""
    lp = SimpleNamespace(num_rows=3000, num_cols=8000, nnz=48000)
    res = [dict(iterations=100, iterations_fp32=60), dict(iterations=50, iterations_fp32=0)]
    per4, per8 = (7 * 8000 + 4 * 3000) * 4, (7 * 8000 + 4 * 3000) * 8
    shared = 2 * 48000 * 8 + (3000 + 8000 + 2) * 4 + (2 * 8000 + 3 * 3000) * 4
    assert bench.whole_solve_bytes_shared(lp, res, 1, 200) == \
        60 * per4 + 40 * per8 + 50 * per8 + 200 * shared


def test_cpu_baseline_runs_the_restatement_only():
    """The CPU legs time the oracle (the restatement), never the reference's
    own build: that build stays in this container."""
    src = open(os.path.join(ROOT, "bench.py")).read()
    assert "ref_solve" not in src and "ref_available" not in src
    assert '"reference"' not in src


def test_autotuner_training_shapes_and_table():
    """The autotuner's training LPs fall in the bucket they tune, and the
    compiled table holds its measured choices (the seeded heterogeneous
    family's bucket (3, 4) included)."""
    import re
    from mipx_amd import autotune
    import numpy as np
    for bc, br in ((2, 4), (3, 4), (3, 5)):
        lps = autotune.training_batch(bc, br, 3)
        for lp in lps:
            assert round(np.log2(lp.nnz / lp.num_cols)) == bc
            assert round(np.log2(lp.nnz / lp.num_rows)) == br
    text = open(autotune.TABLE).read()
    rows = re.findall(r"\{([-\d, ]+)\},", text)
    table = [[int(v) for v in r.split(",")] for r in rows]
    assert len(table) == autotune.COL_BUCKETS and all(len(r) == autotune.ROW_BUCKETS for r in table)
    assert table[3][4] in range(8) and table[2][4] in range(8)
