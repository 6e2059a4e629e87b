"""CPU checks of the C-ABI library: it loads without a GPU and exports every
function include/mipx_pdlp.h declares; the ctypes structs match the header."""
import ctypes as C
import os
import re

from mipx_amd import _abi, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "mipx_pdlp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpx_[a-z0-9_]+)\s*\(", text)) - {"mpx_stop_fn"})


def test_exports_every_declared_function():
    lib = native.lib()
    names = header_functions()
    assert len(names) >= 16
    for nm in names:
        assert hasattr(lib, nm), nm
    assert set(names) <= set(native.EXPORTED)


def test_default_options_match_reference_defaults():
    o = _abi.MpxPdlpOptions()
    native.lib().mpx_pdlp_default_options(C.byref(o))
    for k, v in _abi.DEFAULT_OPTIONS.items():
        assert getattr(o, k) == v, k


def test_struct_sizes():
    # Field order mirrors the header; sizes follow the x86-64 C ABI.
    assert C.sizeof(_abi.MpxLp) == 4 * 4 + 8 * 8 + 8
    assert C.sizeof(_abi.MpxPdlpResult) == 2 * 4 + 9 * 8 + 2 * 4 + 2 * 8 + 2 * 4


def test_invalid_arguments_rejected_without_gpu():
    lib = native.lib()
    res = _abi.MpxPdlpResult()
    o = _abi.make_options()
    assert lib.mpx_pdlp_solve(None, C.byref(o), None, None, None, None, C.byref(res)) == 1
    assert "invalid" in native.last_error()
    # Option checks that return before any device work.
    from mipx_amd import lpgen
    lp = lpgen.c0_plain(seed=3, m=20, n=50).lp
    clp = lp.as_c()
    o.anderson_acceleration, o.anderson_window_size = 1, 65
    assert lib.mpx_pdlp_solve(C.byref(clp), C.byref(o), None, None, None, None, C.byref(res)) == 1
    assert "anderson_window_size" in native.last_error()
    o = _abi.make_options(matrix_zero_tol=1e-9)
    sp, _ = _abi.alloc_scaled(lp.num_rows, lp.num_cols, lp.nnz)
    assert lib.mpx_pdlp_scale_problem(C.byref(clp), C.byref(o), C.byref(sp)) == 4


def test_server_rejects_invalid_arguments_without_gpu():
    lib = native.lib()
    o = _abi.make_options()
    assert lib.mpx_server_create(C.byref(o), 0.0, 0, None) == 1
    assert lib.mpx_server_step(None, None) == 1
    assert lib.mpx_server_submit(None, None, None) == 1
    o.adaptive_step = 1
    h = C.c_void_p()
    assert lib.mpx_server_create(C.byref(o), 0.0, 0, C.byref(h)) == 4
    assert not h.value


def test_presolve_header_exported_by_host_library():
    """Every function include/mipx_presolve.h declares is exported by
    libmipx_host.so (the C++ host library)."""
    from mipx_amd import presolve
    text = open(os.path.join(ROOT, "include", "mipx_presolve.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(mpx_presolve_[a-z0-9_]+)\s*\(", text)))
    assert len(names) == 5
    lib = presolve.lib()
    for nm in names:
        assert hasattr(lib, nm), nm
