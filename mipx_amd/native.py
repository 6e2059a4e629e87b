"""Python access to libmipx_pdlp.so through its C-ABI (include/mipx_pdlp.h).

This is plumbing for tests, smoke and bench: every call goes straight to the
HIP library; there is no CPU fallback.  A missing library raises ImportError.
"""
import ctypes as C

import numpy as np

from ._abi import (MpxBatchStats, MpxLp, MpxPdlpOptions, MpxPdlpResult, MpxScaledProblem,
                   MpxServerStats, alloc_scaled, make_options, scaled_from_arrays, c_double_p)
from ._loader import load_library

_lib = None

EXPORTED = [
    "mpx_pdlp_default_options", "mpx_last_error", "mpx_version", "mpx_device_count",
    "mpx_device_synchronize", "mpx_copy_bench",
    "mpx_pdlp_solve", "mpx_pdlp_solve_batch_shared", "mpx_pdlp_solve_batch_shared_warm",
    "mpx_pdlp_solve_batch", "mpx_server_create", "mpx_server_add_base", "mpx_server_submit",
    "mpx_server_submit_bounds", "mpx_server_submit_bounds_warm",
    "mpx_server_step", "mpx_server_result", "mpx_server_poll", "mpx_server_get_stats",
    "mpx_server_destroy", "mpx_spmv_csr_f64", "mpx_spmv_csr_f32", "mpx_pdlp_scale_problem",
    "mpx_pdhg_iterate", "mpx_pdlp_scale_problem_batched", "mpx_batch_create_shared", "mpx_batch_create", "mpx_batch_run",
    "mpx_batch_results", "mpx_batch_get_stats", "mpx_batch_profile", "mpx_batch_destroy",
    "mpx_batch_profile_refine",
    "mpx_kkt_check_device", "mpx_kkt_check_device_noalloc",
]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    L = load_library("libmipx_pdlp.so")
    L.mpx_last_error.restype = C.c_char_p
    L.mpx_version.restype = C.c_char_p
    L.mpx_pdlp_default_options.argtypes = [C.POINTER(MpxPdlpOptions)]
    L.mpx_device_count.argtypes = [C.POINTER(C.c_int32)]
    L.mpx_device_synchronize.argtypes = [C.c_int32]
    L.mpx_copy_bench.argtypes = [C.c_int64, C.POINTER(C.c_double)]
    L.mpx_pdlp_solve.argtypes = [C.POINTER(MpxLp), C.POINTER(MpxPdlpOptions), C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(MpxPdlpResult)]
    L.mpx_pdlp_solve_batch_shared.argtypes = [C.POINTER(MpxLp), C.c_int32, C.c_void_p,
                                              C.c_void_p, C.POINTER(MpxPdlpOptions),
                                              C.c_void_p, C.c_void_p, C.POINTER(MpxPdlpResult)]
    L.mpx_pdlp_solve_batch_shared_warm.argtypes = [C.POINTER(MpxLp), C.c_int32, C.c_void_p,
                                                   C.c_void_p, C.c_void_p, C.c_void_p,
                                                   C.POINTER(MpxPdlpOptions), C.c_void_p,
                                                   C.c_void_p, C.POINTER(MpxPdlpResult)]
    L.mpx_pdlp_solve_batch.argtypes = [C.POINTER(MpxLp), C.c_int32, C.POINTER(MpxPdlpOptions),
                                       C.c_void_p, C.c_void_p, C.POINTER(MpxPdlpResult)]
    L.mpx_server_create.argtypes = [C.POINTER(MpxPdlpOptions), C.c_double, C.c_int32,
                                    C.POINTER(C.c_void_p)]
    L.mpx_server_add_base.argtypes = [C.c_void_p, C.POINTER(MpxLp), C.POINTER(C.c_int32)]
    L.mpx_server_submit.argtypes = [C.c_void_p, C.POINTER(MpxLp), C.POINTER(C.c_int64)]
    L.mpx_server_submit_bounds.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                           C.POINTER(C.c_int64)]
    L.mpx_server_submit_bounds_warm.argtypes = [C.c_void_p, C.c_int32] + [C.c_void_p] * 4 + [
        C.POINTER(C.c_int64)]
    L.mpx_server_get_stats.argtypes = [C.c_void_p, C.POINTER(MpxServerStats)]
    L.mpx_server_poll.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
    L.mpx_server_step.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    L.mpx_server_result.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int32), C.c_void_p,
                                    C.c_void_p, C.POINTER(MpxPdlpResult)]
    L.mpx_server_destroy.argtypes = [C.c_void_p]
    for nm in ("mpx_spmv_csr_f64", "mpx_spmv_csr_f32"):
        getattr(L, nm).argtypes = [C.c_int32, C.c_int32] + [C.c_void_p] * 5 + [C.c_int32,
                                                                            C.c_void_p]
    L.mpx_pdlp_scale_problem.argtypes = [C.POINTER(MpxLp), C.POINTER(MpxPdlpOptions),
                                         C.POINTER(MpxScaledProblem)]
    L.mpx_pdlp_scale_problem_batched.argtypes = [C.POINTER(MpxLp), C.POINTER(MpxPdlpOptions),
                                                 C.POINTER(MpxScaledProblem)]
    L.mpx_pdhg_iterate.argtypes = ([C.POINTER(MpxLp), C.POINTER(MpxScaledProblem), C.c_int32,
                                    C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32]
                                   + [C.c_void_p] * 6)
    L.mpx_batch_create_shared.argtypes = [C.POINTER(MpxLp), C.c_int32, C.c_void_p, C.c_void_p,
                                          C.POINTER(MpxPdlpOptions), C.POINTER(C.c_void_p)]
    L.mpx_batch_create.argtypes = [C.POINTER(MpxLp), C.c_int32, C.POINTER(MpxPdlpOptions),
                                   C.POINTER(C.c_void_p)]
    L.mpx_batch_run.argtypes = [C.c_void_p]
    L.mpx_batch_results.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(MpxPdlpResult)]
    L.mpx_batch_get_stats.argtypes = [C.c_void_p, C.POINTER(MpxBatchStats)]
    L.mpx_batch_profile.argtypes = [C.c_void_p, C.c_int32, C.POINTER(MpxBatchStats)]
    L.mpx_batch_profile_refine.argtypes = [C.c_void_p, C.c_int32, C.POINTER(MpxBatchStats)]
    L.mpx_batch_destroy.argtypes = [C.c_void_p]
    L.mpx_kkt_check_device.argtypes = [C.POINTER(MpxLp)] + [C.c_void_p] * 7
    L.mpx_kkt_check_device_noalloc.argtypes = [C.POINTER(MpxLp)] + [C.c_void_p] * 7
    _lib = L
    return L


def last_error():
    return lib().mpx_last_error().decode()


def device_count():
    c = C.c_int32(0)
    lib().mpx_device_count(C.byref(c))
    return c.value


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed (code {rc}): {last_error()}")


def solve(lp, warm=None, **opts):
    """One LP on the GPU.  Returns (result dict, x, y) in original space."""
    o = opts.pop("_options", None) or make_options(**opts)
    clp = lp.as_c()
    x = np.zeros(lp.num_cols)
    y = np.zeros(lp.num_rows)
    res = MpxPdlpResult()
    wx = wy = None
    if warm is not None:
        wx = np.ascontiguousarray(warm[0], dtype=np.float64)
        wy = np.ascontiguousarray(warm[1], dtype=np.float64)
    rc = lib().mpx_pdlp_solve(C.byref(clp), C.byref(o),
                              None if wx is None else wx.ctypes.data,
                              None if wy is None else wy.ctypes.data,
                              x.ctypes.data, y.ctypes.data, C.byref(res))
    _check(rc, "mpx_pdlp_solve")
    d = res.as_dict()
    if d["status"] == 6:
        d["error"] = last_error()
    return d, x, y


def solve_batch_shared(lp, col_lower=None, col_upper=None, warm=None, **opts):
    """count = rows of col_lower; returns (list of result dicts, X[count,n], Y[count,m]).
    warm: optional (WX[count,n], WY[count,m]) original-space warm starts."""
    o = opts.pop("_options", None) or make_options(**opts)
    count = opts.pop("count", None)
    if col_lower is not None:
        col_lower = np.ascontiguousarray(col_lower, dtype=np.float64)
        col_upper = np.ascontiguousarray(col_upper, dtype=np.float64)
        count = col_lower.shape[0]
    clp = lp.as_c()
    X = np.zeros((count, lp.num_cols))
    Y = np.zeros((count, lp.num_rows))
    res = (MpxPdlpResult * count)()
    if warm is None:
        rc = lib().mpx_pdlp_solve_batch_shared(
            C.byref(clp), count, None if col_lower is None else col_lower.ctypes.data,
            None if col_upper is None else col_upper.ctypes.data, C.byref(o), X.ctypes.data,
            Y.ctypes.data, res)
    else:
        wx = np.ascontiguousarray(warm[0], dtype=np.float64).reshape(count, lp.num_cols)
        wy = np.ascontiguousarray(warm[1], dtype=np.float64).reshape(count, lp.num_rows)
        rc = lib().mpx_pdlp_solve_batch_shared_warm(
            C.byref(clp), count, None if col_lower is None else col_lower.ctypes.data,
            None if col_upper is None else col_upper.ctypes.data, wx.ctypes.data,
            wy.ctypes.data, C.byref(o), X.ctypes.data, Y.ctypes.data, res)
    _check(rc, "mpx_pdlp_solve_batch_shared")
    return [r.as_dict() for r in res], X, Y


def solve_batch(lps, **opts):
    o = opts.pop("_options", None) or make_options(**opts)
    count = len(lps)
    arr = (MpxLp * count)(*[lp.as_c() for lp in lps])
    xs = [np.zeros(lp.num_cols) for lp in lps]
    ys = [np.zeros(lp.num_rows) for lp in lps]
    xp = (C.c_void_p * count)(*[x.ctypes.data for x in xs])
    yp = (C.c_void_p * count)(*[y.ctypes.data for y in ys])
    res = (MpxPdlpResult * count)()
    rc = lib().mpx_pdlp_solve_batch(arr, count, C.byref(o), C.cast(xp, C.c_void_p),
                                    C.cast(yp, C.c_void_p), res)
    _check(rc, "mpx_pdlp_solve_batch")
    return [r.as_dict() for r in res], xs, ys


def scale_problem(lp, batched=False, **opts):
    """Scaled problem from the single-LP setup, or (batched=True) from the
    batched engines' cohort setup."""
    o = make_options(**opts)
    sp, arrs = alloc_scaled(lp.num_rows, lp.num_cols, lp.nnz)
    clp = lp.as_c()
    f = lib().mpx_pdlp_scale_problem_batched if batched else lib().mpx_pdlp_scale_problem
    _check(f(C.byref(clp), C.byref(o), C.byref(sp)), "mpx_pdlp_scale_problem")
    arrs["objective_scale"] = sp.objective_scale
    arrs["constraint_scale"] = sp.constraint_scale
    arrs["spectral_norm"] = sp.spectral_norm
    arrs["power_iterations"] = sp.power_iterations
    return arrs


def pdhg_iterate(lp, scaled, iters, inner, step, pw, state, precision=0, group=1):
    sp = scaled_from_arrays(scaled)
    clp = lp.as_c()
    _check(lib().mpx_pdhg_iterate(C.byref(clp), C.byref(sp), iters, inner, step, pw, precision,
                                  group, *[state[k].ctypes.data for k in
                                           ("cx", "ix", "px", "cy", "iy", "py")]),
           "mpx_pdhg_iterate")
    return state


def spmv_device(rows, cols, rs, ci, val, x, y, group=0, stream=None):
    """y = A x on device tensors (torch tensors or raw pointers)."""
    ptr = lambda t: t if isinstance(t, int) else t.data_ptr()
    f = lib().mpx_spmv_csr_f64 if str(getattr(val, "dtype", "")) == "torch.float64" else \
        lib().mpx_spmv_csr_f32
    _check(f(rows, cols, ptr(rs), ptr(ci), ptr(val), ptr(x), ptr(y), group, stream),
           "mpx_spmv_csr")


class ResidentBatch:
    """A batch whose problem data stays in HBM between solves
    (mpx_batch_create_shared / mpx_batch_create, then mpx_batch_run)."""

    def __init__(self, lp=None, col_lower=None, col_upper=None, lps=None, **opts):
        L = lib()
        self._o = opts.pop("_options", None) or make_options(**opts)
        self._h = C.c_void_p()
        if lps is not None:
            self.lps = list(lps)
            self._clps = (MpxLp * len(self.lps))(*[x.as_c() for x in self.lps])
            rc = L.mpx_batch_create(self._clps, len(self.lps), C.byref(self._o), C.byref(self._h))
            self.count = len(self.lps)
        else:
            self.lps = [lp]
            self._clp = lp.as_c()
            self.col_lower = np.ascontiguousarray(col_lower, dtype=np.float64)
            self.col_upper = np.ascontiguousarray(col_upper, dtype=np.float64)
            self.count = self.col_lower.shape[0]
            rc = L.mpx_batch_create_shared(C.byref(self._clp), self.count,
                                           self.col_lower.ctypes.data, self.col_upper.ctypes.data,
                                           C.byref(self._o), C.byref(self._h))
        _check(rc, "mpx_batch_create")

    def run(self):
        _check(lib().mpx_batch_run(self._h), "mpx_batch_run")

    def results(self, want_vectors=True):
        res = (MpxPdlpResult * self.count)()
        xs = ys = None
        if want_vectors:
            if len(self.lps) == 1:
                xs = [np.zeros(self.lps[0].num_cols) for _ in range(self.count)]
                ys = [np.zeros(self.lps[0].num_rows) for _ in range(self.count)]
            else:
                xs = [np.zeros(l.num_cols) for l in self.lps]
                ys = [np.zeros(l.num_rows) for l in self.lps]
            xp = (C.c_void_p * self.count)(*[x.ctypes.data for x in xs])
            yp = (C.c_void_p * self.count)(*[y.ctypes.data for y in ys])
            rc = lib().mpx_batch_results(self._h, C.cast(xp, C.c_void_p), C.cast(yp, C.c_void_p), res)
        else:
            rc = lib().mpx_batch_results(self._h, None, None, res)
        _check(rc, "mpx_batch_results")
        return [r.as_dict() for r in res], xs, ys

    def stats(self):
        st = MpxBatchStats()
        _check(lib().mpx_batch_get_stats(self._h, C.byref(st)), "mpx_batch_get_stats")
        return st.as_dict()

    def profile(self, launches=50, refine=False):
        """Times `launches` launches of the iteration kernels (refine: the
        refinement's correction kernels, mixed precision)."""
        st = MpxBatchStats()
        f = lib().mpx_batch_profile_refine if refine else lib().mpx_batch_profile
        _check(f(self._h, launches, C.byref(st)), "mpx_batch_profile")
        return st.as_dict()

    def close(self):
        if self._h:
            lib().mpx_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching serving loop (mpx_server_*): submit LPs or bounds
    variants of a registered base at any time, call step() to advance every
    admitted request by one round, collect finished results by ticket."""

    def __init__(self, hbm_budget_bytes=0.0, max_slots=0, **opts):
        self._o = opts.pop("_options", None) or make_options(**opts)
        self._h = C.c_void_p()
        _check(lib().mpx_server_create(C.byref(self._o), float(hbm_budget_bytes), int(max_slots),
                                       C.byref(self._h)), "mpx_server_create")
        self._dims = {}   # ticket -> (n, m)
        self._base_dims = []

    def add_base(self, lp):
        bid = C.c_int32(-1)
        clp = lp.as_c()
        _check(lib().mpx_server_add_base(self._h, C.byref(clp), C.byref(bid)),
               "mpx_server_add_base")
        self._base_dims.append((lp.num_cols, lp.num_rows))
        return bid.value

    def submit(self, lp):
        t = C.c_int64(-1)
        clp = lp.as_c()
        _check(lib().mpx_server_submit(self._h, C.byref(clp), C.byref(t)), "mpx_server_submit")
        self._dims[t.value] = (lp.num_cols, lp.num_rows)
        return t.value

    def submit_bounds(self, base_id, col_lower, col_upper, warm=None):
        """warm: optional (x, y) in original space (a parent node's solution)."""
        lo = np.ascontiguousarray(col_lower, dtype=np.float64)
        hi = np.ascontiguousarray(col_upper, dtype=np.float64)
        t = C.c_int64(-1)
        if warm is None:
            _check(lib().mpx_server_submit_bounds(self._h, base_id, lo.ctypes.data,
                                                  hi.ctypes.data, C.byref(t)),
                   "mpx_server_submit_bounds")
        else:
            wx = np.ascontiguousarray(warm[0], dtype=np.float64)
            wy = np.ascontiguousarray(warm[1], dtype=np.float64)
            _check(lib().mpx_server_submit_bounds_warm(self._h, base_id, lo.ctypes.data,
                                                       hi.ctypes.data, wx.ctypes.data,
                                                       wy.ctypes.data, C.byref(t)),
                   "mpx_server_submit_bounds_warm")
        self._dims[t.value] = self._base_dims[base_id]
        return t.value

    def step(self):
        a = C.c_int32(0)
        _check(lib().mpx_server_step(self._h, C.byref(a)), "mpx_server_step")
        return a.value

    def result(self, ticket, want_vectors=True):
        """(result dict, x, y) once finished (then forgotten), else None."""
        n, m = self._dims[ticket]
        x = np.zeros(n) if want_vectors else None
        y = np.zeros(m) if want_vectors else None
        done = C.c_int32(0)
        res = MpxPdlpResult()
        _check(lib().mpx_server_result(self._h, ticket, C.byref(done),
                                       None if x is None else x.ctypes.data,
                                       None if y is None else y.ctypes.data, C.byref(res)),
               "mpx_server_result")
        if not done.value:
            return None
        del self._dims[ticket]
        return res.as_dict(), x, y

    def poll(self, capacity=4096):
        """Tickets finished since the last poll (completion order)."""
        buf = np.empty(capacity, np.int64)
        cnt = C.c_int32(0)
        _check(lib().mpx_server_poll(self._h, buf.ctypes.data, capacity, C.byref(cnt)),
               "mpx_server_poll")
        return buf[:cnt.value].tolist()

    def stats(self):
        st = MpxServerStats()
        _check(lib().mpx_server_get_stats(self._h, C.byref(st)), "mpx_server_get_stats")
        return st.as_dict()

    def close(self):
        if self._h:
            lib().mpx_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


KKT_FIELDS = ["rel_primal", "rel_dual", "rel_gap", "primal_obj", "dual_obj", "bound_viol",
              "row_viol_max", "dual_sign_viol"]


def kkt_check_device(lp, x, y, entry="mpx_kkt_check_device"):
    """Relative KKT residuals of (x, y) computed on the device
    (mpx_kkt_check_device, or another entry point with its signature): the
    LP, its CSC and the vectors are copied to device memory with torch, the
    eight outputs come back as a dict."""
    import scipy.sparse as sp
    import torch
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    a = sp.csr_matrix((lp.values, lp.col_indices, lp.row_starts),
                      shape=(lp.num_rows, lp.num_cols)).tocsc()
    a.sort_indices()
    keep = [dev(v) for v in (lp.row_starts, lp.col_indices, lp.values, lp.obj, lp.col_lower,
                             lp.col_upper, lp.row_lower, lp.row_upper,
                             a.indptr.astype(np.int32), a.indices.astype(np.int32), a.data,
                             np.asarray(x, np.float64), np.asarray(y, np.float64))]
    out = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = [t.data_ptr() for t in keep]
    from ._abi import c_double_p, c_int32_p
    ptrs = [C.cast(v, c_int32_p if k < 2 else c_double_p) for k, v in enumerate(p[:8])]
    clp = MpxLp(lp.num_rows, lp.num_cols, lp.nnz, lp.sense, *ptrs, lp.obj_offset)
    stream = torch.cuda.current_stream().cuda_stream
    _check(getattr(lib(), entry)(C.byref(clp), p[8], p[9], p[10], p[11], p[12],
                                 out.data_ptr(), stream), entry)
    torch.cuda.synchronize()
    return dict(zip(KKT_FIELDS, out.cpu().numpy().tolist()))


def synchronize(device=0):
    _check(lib().mpx_device_synchronize(device), "mpx_device_synchronize")


def copy_bench(nbytes):
    """Device-to-device copy of nbytes with a plain streaming kernel; returns
    nbytes (the kernel time is available as copy_bench_ms)."""
    ms = C.c_double(0.0)
    _check(lib().mpx_copy_bench(nbytes, C.byref(ms)), "mpx_copy_bench")
    global copy_bench_ms
    copy_bench_ms = ms.value
    return nbytes
