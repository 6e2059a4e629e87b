"""ctypes access to the LP presolve of libmipx_host.so (include/mipx_presolve.h).

    pre = Presolved(lp)          # reduced LP in pre.reduced, counters in pre.stats
    x, y = pre.postsolve(x_red, y_red)
"""
import ctypes as C

import numpy as np

from ._abi import MpxLp
from ._loader import load_library
from .lp import LpProblem


class MpxPresolveStats(C.Structure):
    _fields_ = [("vars_removed", C.c_int32), ("rows_removed", C.c_int32),
                ("fixed_var_changes", C.c_int32), ("singleton_row_changes", C.c_int32),
                ("empty_row_changes", C.c_int32), ("empty_col_changes", C.c_int32),
                ("rounds", C.c_int32), ("infeasible", C.c_int32), ("seconds", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = load_library("libmipx_host.so", local=True)
        L.mpx_presolve_run.argtypes = [C.POINTER(MpxLp), C.POINTER(C.c_void_p)]
        L.mpx_presolve_reduced.argtypes = [C.c_void_p, C.POINTER(MpxLp)]
        L.mpx_presolve_get_stats.argtypes = [C.c_void_p, C.POINTER(MpxPresolveStats)]
        L.mpx_presolve_postsolve.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.mpx_presolve_destroy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


class Presolved:
    """A presolved LP: `reduced` (an LpProblem copy), `stats` (dict) and
    postsolve() back to the original space."""

    def __init__(self, lp):
        self.lp = lp
        self._h = C.c_void_p()
        clp = lp.as_c()
        rc = lib().mpx_presolve_run(C.byref(clp), C.byref(self._h))
        if rc != 0:
            raise RuntimeError(f"mpx_presolve_run failed (code {rc})")
        r = MpxLp()
        lib().mpx_presolve_reduced(self._h, C.byref(r))
        m, n, nnz = r.num_rows, r.num_cols, r.nnz
        arr = lambda p, k, dt: (np.ctypeslib.as_array(p, shape=(k,)).astype(dt).copy()
                                if k > 0 else np.zeros(0, dt))
        self.reduced = LpProblem(m, n, arr(r.row_starts, m + 1, np.int32) if m > 0 else
                                 np.zeros(1, np.int32), arr(r.col_indices, nnz, np.int32),
                                 arr(r.values, nnz, np.float64), arr(r.obj, n, np.float64),
                                 arr(r.col_lower, n, np.float64), arr(r.col_upper, n, np.float64),
                                 arr(r.row_lower, m, np.float64), arr(r.row_upper, m, np.float64),
                                 r.sense, r.obj_offset, name=lp.name + "-presolved")
        st = MpxPresolveStats()
        lib().mpx_presolve_get_stats(self._h, C.byref(st))
        self.stats = {k: getattr(st, k) for k, _ in MpxPresolveStats._fields_}

    def postsolve(self, x_red, y_red=None):
        x = np.zeros(self.lp.num_cols)
        y = np.zeros(self.lp.num_rows) if y_red is not None else None
        xr = np.ascontiguousarray(x_red, dtype=np.float64)
        yr = None if y_red is None else np.ascontiguousarray(y_red, dtype=np.float64)
        rc = lib().mpx_presolve_postsolve(self._h, xr.ctypes.data,
                                          None if yr is None else yr.ctypes.data, x.ctypes.data,
                                          None if y is None else y.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"mpx_presolve_postsolve failed (code {rc})")
        return x, y

    def close(self):
        if self._h:
            lib().mpx_presolve_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
