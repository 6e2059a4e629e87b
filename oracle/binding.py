"""ctypes access to oracle/liboracle.so and, if built, oracle/_ref/libref_pdlp.so.

Test infrastructure only (see oracle/pdlp_oracle.cpp for the reference
functions each entry point restates).
"""
import ctypes as C
import os

import numpy as np

from mipx_amd._abi import (MpxLp, MpxPdlpOptions, MpxPdlpResult, MpxScaledProblem,
                           alloc_scaled, make_options)

HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}


def _load(rel):
    if rel not in _libs:
        path = os.path.join(HERE, rel)
        if not os.path.exists(path):
            raise ImportError(f"oracle library {path} not built (make -C oracle)")
        lib = C.CDLL(path)
        _libs[rel] = lib
    return _libs[rel]


def oracle_lib():
    lib = _load("liboracle.so")
    lib.orc_pdlp_solve.argtypes = [C.POINTER(MpxLp), C.POINTER(MpxPdlpOptions), C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(MpxPdlpResult)]
    lib.orc_pdlp_solve_gpu.argtypes = lib.orc_pdlp_solve.argtypes
    lib.orc_scale_problem.argtypes = [C.POINTER(MpxLp), C.POINTER(MpxPdlpOptions),
                                      C.POINTER(MpxScaledProblem)]
    lib.orc_pdhg_iterate.argtypes = [C.POINTER(MpxLp), C.POINTER(MpxScaledProblem), C.c_int32,
                                     C.c_int32, C.c_double, C.c_double] + [C.c_void_p] * 6
    lib.orc_kkt_check.argtypes = [C.POINTER(MpxLp), C.c_void_p, C.c_void_p, C.c_void_p]
    lib.orc_spmv.argtypes = [C.c_int32, C.c_int32] + [C.c_void_p] * 5
    lib.orc_spmv_t.argtypes = [C.c_int32, C.c_int32] + [C.c_void_p] * 5
    lib.orc_reduced_costs.argtypes = [C.POINTER(MpxLp), C.c_void_p, C.c_void_p]
    return lib


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libref_pdlp.so"))


def ref_lib():
    lib = _load(os.path.join("_ref", "libref_pdlp.so"))
    lib.ref_pdlp_solve.argtypes = [C.POINTER(MpxLp), C.POINTER(MpxPdlpOptions), C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(MpxPdlpResult)]
    return lib


def _ptr(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data


def _solve_with(fn, lp, warm=None, **opts):
    o = opts.pop("_options", None) or make_options(**opts)
    clp = lp.as_c()
    x = np.zeros(lp.num_cols)
    y = np.zeros(lp.num_rows)
    res = MpxPdlpResult()
    wx = wy = None
    if warm is not None:
        wx = np.ascontiguousarray(warm[0], dtype=np.float64)
        wy = np.ascontiguousarray(warm[1], dtype=np.float64)
    rc = fn(C.byref(clp), C.byref(o), _ptr(wx), _ptr(wy), x.ctypes.data, y.ctypes.data,
            C.byref(res))
    if rc != 0:
        raise RuntimeError(f"oracle solve failed with code {rc}")
    return res.as_dict(), x, y


def solve(lp, warm=None, **opts):
    """CPU restatement of PdlpSolver::solve (CPU path)."""
    return _solve_with(oracle_lib().orc_pdlp_solve, lp, warm, **opts)


def solve_gpu(lp, warm=None, **opts):
    """The same solve with the GPU host loop's adaptive step (pdlp.cpp:2271-2307)."""
    return _solve_with(oracle_lib().orc_pdlp_solve_gpu, lp, warm, **opts)


def ref_solve(lp, warm=None, **opts):
    """The reference's own CPU PdlpSolver, built from its sources."""
    return _solve_with(ref_lib().ref_pdlp_solve, lp, warm, **opts)


def scale_problem(lp, **opts):
    o = make_options(**opts)
    sp, arrs = alloc_scaled(lp.num_rows, lp.num_cols, lp.nnz)
    clp = lp.as_c()
    rc = oracle_lib().orc_scale_problem(C.byref(clp), C.byref(o), C.byref(sp))
    assert rc == 0
    arrs["objective_scale"] = sp.objective_scale
    arrs["constraint_scale"] = sp.constraint_scale
    arrs["spectral_norm"] = sp.spectral_norm
    arrs["power_iterations"] = sp.power_iterations
    return arrs


def pdhg_iterate(lp, scaled, iters, inner, step, pw, state):
    """state: dict with cx, ix, px, cy, iy, py float64 arrays, updated in place."""
    from mipx_amd._abi import scaled_from_arrays
    sp = scaled_from_arrays(scaled)
    clp = lp.as_c()
    oracle_lib().orc_pdhg_iterate(C.byref(clp), C.byref(sp), iters, inner, step, pw,
                                  *[state[k].ctypes.data for k in ("cx", "ix", "px", "cy", "iy", "py")])
    return state


def kkt_check(lp, x, y):
    out = np.zeros(8)
    clp = lp.as_c()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    oracle_lib().orc_kkt_check(C.byref(clp), x.ctypes.data, y.ctypes.data, out.ctypes.data)
    keys = ["rel_primal", "rel_dual", "rel_gap", "primal_obj", "dual_obj", "bound_viol",
            "row_viol_max", "dual_sign_viol"]
    return dict(zip(keys, out.tolist()))


def spmv(lp, x):
    y = np.zeros(lp.num_rows)
    x = np.ascontiguousarray(x, dtype=np.float64)
    oracle_lib().orc_spmv(lp.num_rows, lp.num_cols, lp.row_starts.ctypes.data,
                          lp.col_indices.ctypes.data, lp.values.ctypes.data, x.ctypes.data,
                          y.ctypes.data)
    return y


def spmv_t(lp, x):
    y = np.zeros(lp.num_cols)
    x = np.ascontiguousarray(x, dtype=np.float64)
    oracle_lib().orc_spmv_t(lp.num_rows, lp.num_cols, lp.row_starts.ctypes.data,
                            lp.col_indices.ctypes.data, lp.values.ctypes.data, x.ctypes.data,
                            y.ctypes.data)
    return y


def reduced_costs(lp, y):
    rc = np.zeros(lp.num_cols)
    y = np.ascontiguousarray(y, dtype=np.float64)
    clp = lp.as_c()
    oracle_lib().orc_reduced_costs(C.byref(clp), y.ctypes.data, rc.ctypes.data)
    return rc
