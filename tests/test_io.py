"""MPS reader / writer and .solu reader of the C++ host mirror (no GPU):
the reference's tiny.mps / tiny.solu data files with the fields its MPS
reader test expects (tests/test_lp_problem.cpp:50-108), writer round trip,
gzip input, RANGES / BOUNDS / offset semantics."""
import os
import subprocess

from lp_cases import GOLDEN

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mipx_amd", "bin")


def test_mps_io(tmp_path):
    r = subprocess.run([os.path.join(BIN, "io_tests"), GOLDEN, str(tmp_path)],
                       capture_output=True, text=True, timeout=60)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert " 0 failed" in r.stdout
