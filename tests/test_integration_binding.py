"""The reference-side binding (mipx_amd/integration/pdlp_mi355x.cpp) compiles
against the reference's own headers with the macros its build would set, and
every library symbol it calls is exported by libmipx_pdlp.so.  Needs the
reference's include/ directory; skipped where it is absent."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INCLUDE = "/root/reference/include"
BINDING = os.path.join(ROOT, "mipx_amd", "integration", "pdlp_mi355x.cpp")
FLAGS = ["-std=c++20", "-Wall", "-Wextra", "-Werror", "-DMIPX_HAS_CUDA", "-DMIPX_HAS_MI355X",
         "-I" + REF_INCLUDE, "-I" + os.path.join(ROOT, "include")]

needs_ref = pytest.mark.skipif(
    not os.path.exists(os.path.join(REF_INCLUDE, "mipx", "pdlp.h")) or shutil.which("g++") is None,
    reason="reference headers or g++ not available")


@needs_ref
def test_binding_compiles_against_reference_headers():
    r = subprocess.run(["g++", "-fsyntax-only", *FLAGS, BINDING], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@needs_ref
def test_binding_calls_only_exported_symbols(tmp_path):
    obj = tmp_path / "pdlp_mi355x.o"
    r = subprocess.run(["g++", "-c", "-fPIC", *FLAGS, BINDING, "-o", str(obj)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    undef = subprocess.run(["nm", "-u", str(obj)], capture_output=True, text=True, check=True).stdout
    called = sorted(set(re.findall(r"\b(mpx_[a-z0-9_]+)\b", undef)))
    assert called == ["mpx_pdlp_default_options", "mpx_pdlp_solve"], called
    lib_path = os.path.join(ROOT, "mipx_amd", "libmipx_pdlp.so")
    if not os.path.exists(lib_path):
        pytest.skip("libmipx_pdlp.so not built")
    lib = C.CDLL(lib_path, mode=C.RTLD_LOCAL)
    for name in called:
        assert hasattr(lib, name), name
