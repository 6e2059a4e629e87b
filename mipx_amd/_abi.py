"""ctypes mirror of include/mipx_pdlp.h (the C-ABI of libmipx_pdlp.so).

Only plain C types cross the boundary.  The structs below must keep the exact
field order of the header.
"""
import ctypes as C

MPX_OK = 0
MPX_ERR_INVALID = 1
MPX_ERR_DEVICE = 2
MPX_ERR_OUT_OF_MEMORY = 3
MPX_ERR_UNSUPPORTED = 4
MPX_ERR_BUSY = 5

STATUS_NAMES = ["Optimal", "Infeasible", "Unbounded", "IterLimit", "TimeLimit",
                "NodeLimit", "Error"]
OPTIMAL, INFEASIBLE, UNBOUNDED, ITER_LIMIT, TIME_LIMIT, NODE_LIMIT, ERROR = range(7)

PREC_FP64 = 0
PREC_MIXED = 1

c_int32_p = C.POINTER(C.c_int32)
c_double_p = C.POINTER(C.c_double)

STOP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)


class MpxLp(C.Structure):
    _fields_ = [
        ("num_rows", C.c_int32),
        ("num_cols", C.c_int32),
        ("nnz", C.c_int32),
        ("sense", C.c_int32),
        ("row_starts", c_int32_p),
        ("col_indices", c_int32_p),
        ("values", c_double_p),
        ("obj", c_double_p),
        ("col_lower", c_double_p),
        ("col_upper", c_double_p),
        ("row_lower", c_double_p),
        ("row_upper", c_double_p),
        ("obj_offset", C.c_double),
    ]


class MpxPdlpOptions(C.Structure):
    _fields_ = [
        ("max_iter", C.c_int32),
        ("optimality_tol", C.c_double),
        ("primal_tol", C.c_double),
        ("dual_tol", C.c_double),
        ("max_solve_seconds", C.c_double),
        ("optimality_norm", C.c_int32),
        ("sv_max_iter", C.c_int32),
        ("sv_tol", C.c_double),
        ("adaptive_step", C.c_int32),
        ("adaptive_step_down", C.c_double),
        ("adaptive_step_up", C.c_double),
        ("adaptive_step_min", C.c_double),
        ("adaptive_step_max", C.c_double),
        ("primal_weight", C.c_double),
        ("update_primal_weight", C.c_int32),
        ("pid_kp", C.c_double),
        ("pid_ki", C.c_double),
        ("pid_kd", C.c_double),
        ("pid_i_smooth", C.c_double),
        ("restart_sufficient_decay", C.c_double),
        ("restart_necessary_decay", C.c_double),
        ("restart_artificial_fraction", C.c_double),
        ("termination_eval_frequency", C.c_int32),
        ("matrix_zero_tol", C.c_double),
        ("enable_polishing", C.c_int32),
        ("polish_primal_tol", C.c_double),
        ("polish_dual_tol", C.c_double),
        ("polish_max_iter", C.c_int32),
        ("do_ruiz_scaling", C.c_int32),
        ("ruiz_iterations", C.c_int32),
        ("do_pock_chambolle_scaling", C.c_int32),
        ("do_bound_obj_rescaling", C.c_int32),
        ("preconditioner_refresh", C.c_int32),
        ("preconditioner_refresh_ratio", C.c_double),
        ("anderson_acceleration", C.c_int32),
        ("anderson_window_size", C.c_int32),
        ("anderson_regularization", C.c_double),
        ("use_gpu", C.c_int32),
        ("gpu_min_rows", C.c_int32),
        ("gpu_min_nnz", C.c_int32),
        ("verbose", C.c_int32),
        ("stop_fn", STOP_FN),
        ("stop_user", C.c_void_p),
        ("precision", C.c_int32),
        ("refine_tol", C.c_double),
        ("spmv_group", C.c_int32),
        ("device", C.c_int32),
        ("mixed_handover", C.c_double),
        ("spmv_chunk_cols", C.c_int32),
        ("packed_preset", C.c_int32),
        ("mixed_refine", C.c_int32),
    ]


class MpxPdlpResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("iterations", C.c_int32),
        ("objective", C.c_double),
        ("work_units", C.c_double),
        ("rel_primal", C.c_double),
        ("rel_dual", C.c_double),
        ("rel_gap", C.c_double),
        ("primal_obj", C.c_double),
        ("dual_obj", C.c_double),
        ("step", C.c_double),
        ("primal_weight", C.c_double),
        ("restarts", C.c_int32),
        ("iterations_fp32", C.c_int32),
        ("seconds", C.c_double),
        ("setup_seconds", C.c_double),
        ("iterations_refine", C.c_int32),
        ("pad", C.c_int32),
    ]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["status_name"] = STATUS_NAMES[self.status] if 0 <= self.status < 7 else "?"
        return d


class MpxScaledProblem(C.Structure):
    _fields_ = [
        ("values", c_double_p),
        ("at_values", c_double_p),
        ("at_row_starts", c_int32_p),
        ("at_col_indices", c_int32_p),
        ("cscaled", c_double_p),
        ("col_lower", c_double_p),
        ("col_upper", c_double_p),
        ("row_lower", c_double_p),
        ("row_upper", c_double_p),
        ("row_scale", c_double_p),
        ("col_scale", c_double_p),
        ("sigma_base", c_double_p),
        ("tau_base", c_double_p),
        ("objective_scale", C.c_double),
        ("constraint_scale", C.c_double),
        ("spectral_norm", C.c_double),
        ("power_iterations", C.c_int32),
    ]


# Defaults of mipx::PdlpOptions (include/mipx/pdlp.h:19-87).
DEFAULT_OPTIONS = dict(
    max_iter=4000000, optimality_tol=5e-5, primal_tol=5e-5, dual_tol=5e-5,
    max_solve_seconds=-1.0, optimality_norm=0, sv_max_iter=5000, sv_tol=1e-4,
    adaptive_step=0, adaptive_step_down=0.7, adaptive_step_up=1.0,
    adaptive_step_min=1e-6, adaptive_step_max=1e3, primal_weight=1.0,
    update_primal_weight=1, pid_kp=0.99, pid_ki=0.01, pid_kd=0.0, pid_i_smooth=0.3,
    restart_sufficient_decay=0.2, restart_necessary_decay=0.5,
    restart_artificial_fraction=0.36, termination_eval_frequency=200,
    matrix_zero_tol=0.0, enable_polishing=1, polish_primal_tol=1e-6,
    polish_dual_tol=1e-6, polish_max_iter=500000, do_ruiz_scaling=1,
    ruiz_iterations=10, do_pock_chambolle_scaling=1, do_bound_obj_rescaling=1,
    preconditioner_refresh=0, preconditioner_refresh_ratio=10.0,
    anderson_acceleration=0, anderson_window_size=5, anderson_regularization=1e-10,
    use_gpu=1, gpu_min_rows=512, gpu_min_nnz=10000, verbose=1,
    precision=PREC_FP64, refine_tol=-1.0, spmv_group=0, device=0, mixed_handover=3.0,
    spmv_chunk_cols=0, packed_preset=-1, mixed_refine=1,
)


def make_options(**overrides):
    o = MpxPdlpOptions()
    for k, v in DEFAULT_OPTIONS.items():
        setattr(o, k, v)
    for k, v in overrides.items():
        if k not in DEFAULT_OPTIONS and k not in ("stop_fn", "stop_user"):
            raise KeyError(f"unknown PDLP option {k!r}")
        setattr(o, k, v)
    return o


def alloc_scaled(m, n, nnz):
    """An MpxScaledProblem whose pointers refer to fresh numpy arrays."""
    import numpy as np
    arrs = dict(
        values=np.zeros(nnz), at_values=np.zeros(nnz),
        at_row_starts=np.zeros(n + 1, np.int32), at_col_indices=np.zeros(nnz, np.int32),
        cscaled=np.zeros(n), col_lower=np.zeros(n), col_upper=np.zeros(n),
        row_lower=np.zeros(m), row_upper=np.zeros(m), row_scale=np.zeros(m),
        col_scale=np.zeros(n), sigma_base=np.zeros(m), tau_base=np.zeros(n))
    sp = MpxScaledProblem()
    for k, a in arrs.items():
        ptype = c_int32_p if a.dtype == np.int32 else c_double_p
        setattr(sp, k, a.ctypes.data_as(ptype))
    return sp, arrs


def scaled_from_arrays(arrs):
    """An MpxScaledProblem viewing an existing dict of arrays."""
    import numpy as np
    sp = MpxScaledProblem()
    for k, a in arrs.items():
        if isinstance(a, np.ndarray):
            ptype = c_int32_p if a.dtype == np.int32 else c_double_p
            setattr(sp, k, a.ctypes.data_as(ptype))
    return sp


class MpxBatchStats(C.Structure):
    _fields_ = [
        ("upload_ms", C.c_double),
        ("setup_ms", C.c_double),
        ("run_ms", C.c_double),
        ("rounds", C.c_int32),
        ("interleaved", C.c_int32),
        ("instance_iterations", C.c_int64),
        ("latency_p50_ms", C.c_double),
        ("latency_p99_ms", C.c_double),
        ("latency_max_ms", C.c_double),
        ("primal_kernel_us", C.c_double),
        ("dual_kernel_us", C.c_double),
        ("profiled_slots", C.c_int32),
        ("profiled_launches", C.c_int32),
        ("columns_per_wave", C.c_int32),
        ("rows_per_wave", C.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class MpxServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64),
        ("finished", C.c_int64),
        ("queued", C.c_int32),
        ("preparing", C.c_int32),
        ("running", C.c_int32),
        ("slots", C.c_int32),
        ("rounds", C.c_int64),
        ("pool_bytes", C.c_double),
        ("pool_used_bytes", C.c_double),
        ("pool_peak_bytes", C.c_double),
        ("iterate_ms", C.c_double),
        ("instance_iterations", C.c_int64),
        ("admit_ms", C.c_double),
        ("activate_ms", C.c_double),
        ("retire_ms", C.c_double),
        ("max_step_ms", C.c_double),
        ("setup_query_ms", C.c_double),
        ("setup_wait_ms", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}
