// pdlp_mi355x.cpp -- PdlpSolver::solveGpu() of mipx on libmipx_pdlp.so.
//
// The file a mipx maintainer adds as src/lp/pdlp_mi355x.cpp to run the
// reference's GPU PDLP path on an MI355X.  It replaces the CUDA definition
// of solveGpu() (src/lp/pdlp.cpp:1726-2794), the kernels of
// src/lp/pdlp_kernels.cu and the cuSPARSE calls; the gate that calls it
// (pdlp.cpp:862-869), the CPU loop and every accessor and mutator stay as
// they are.  It is compiled against the reference's own headers:
//
//   MIPX_HAS_CUDA    declares solveGpu() (include/mipx/pdlp.h:131-133) and
//                    turns the gate on;
//   MIPX_HAS_MI355X  set by the build that links libmipx_pdlp.so; pdlp.cpp's
//                    two CUDA blocks (lines 10-17 and 1726-2794) are then
//                    skipped by `#if defined(MIPX_HAS_CUDA) && !defined(MIPX_HAS_MI355X)`.
//
// The library scales the original LP itself (Ruiz, Pock-Chambolle, bound /
// objective rescaling, all on the device), so the call passes original_,
// not the host-scaled copy, and receives the solution in original space.
#include <atomic>
#include <cstdint>

#include "mipx/pdlp.h"
#include "mipx_pdlp.h"

namespace mipx {

namespace {

int pollStopFlag(void* user) {
    return static_cast<const std::atomic<bool>*>(user)->load(std::memory_order_relaxed) ? 1 : 0;
}

// PdlpOptions (include/mipx/pdlp.h:19-87) -> mpx_pdlp_options, field by field.
mpx_pdlp_options toMi355x(const PdlpOptions& p) {
    mpx_pdlp_options o;
    mpx_pdlp_default_options(&o);
    o.max_iter = static_cast<int32_t>(p.max_iter);
    o.optimality_tol = p.optimality_tol;
    o.primal_tol = p.primal_tol;
    o.dual_tol = p.dual_tol;
    o.max_solve_seconds = p.max_solve_seconds;
    o.optimality_norm = p.optimality_norm == PdlpOptimalityNorm::LInf ? 1 : 0;
    o.sv_max_iter = static_cast<int32_t>(p.sv_max_iter);
    o.sv_tol = p.sv_tol;
    o.adaptive_step = p.adaptive_step ? 1 : 0;
    o.adaptive_step_down = p.adaptive_step_down;
    o.adaptive_step_up = p.adaptive_step_up;
    o.adaptive_step_min = p.adaptive_step_min;
    o.adaptive_step_max = p.adaptive_step_max;
    o.primal_weight = p.primal_weight;
    o.update_primal_weight = p.update_primal_weight ? 1 : 0;
    o.pid_kp = p.pid_kp;
    o.pid_ki = p.pid_ki;
    o.pid_kd = p.pid_kd;
    o.pid_i_smooth = p.pid_i_smooth;
    o.restart_sufficient_decay = p.restart_sufficient_decay;
    o.restart_necessary_decay = p.restart_necessary_decay;
    o.restart_artificial_fraction = p.restart_artificial_fraction;
    o.termination_eval_frequency = static_cast<int32_t>(p.termination_eval_frequency);
    o.matrix_zero_tol = p.matrix_zero_tol;
    o.enable_polishing = p.enable_polishing ? 1 : 0;
    o.polish_primal_tol = p.polish_primal_tol;
    o.polish_dual_tol = p.polish_dual_tol;
    o.polish_max_iter = static_cast<int32_t>(p.polish_max_iter);
    o.do_ruiz_scaling = p.do_ruiz_scaling ? 1 : 0;
    o.ruiz_iterations = static_cast<int32_t>(p.ruiz_iterations);
    o.do_pock_chambolle_scaling = p.do_pock_chambolle_scaling ? 1 : 0;
    o.do_bound_obj_rescaling = p.do_bound_obj_rescaling ? 1 : 0;
    o.preconditioner_refresh = p.preconditioner_refresh ? 1 : 0;
    o.preconditioner_refresh_ratio = p.preconditioner_refresh_ratio;
    o.anderson_acceleration = p.anderson_acceleration ? 1 : 0;
    o.anderson_window_size = static_cast<int32_t>(p.anderson_window_size);
    o.anderson_regularization = p.anderson_regularization;
    o.use_gpu = p.use_gpu ? 1 : 0;
    o.gpu_min_rows = static_cast<int32_t>(p.gpu_min_rows);
    o.gpu_min_nnz = static_cast<int32_t>(p.gpu_min_nnz);
    o.verbose = p.verbose ? 1 : 0;
    if (p.stop_flag != nullptr) {
        o.stop_fn = &pollStopFlag;
        o.stop_user = const_cast<std::atomic<bool>*>(p.stop_flag);
    }
    // The reference's GPU path computes in FP64 throughout.
    o.precision = MPX_PREC_FP64;
    return o;
}

}  // namespace

LpResult PdlpSolver::solveGpu() {
    used_gpu_ = false;
    const SparseMatrix& a = original_.matrix;
    mpx_lp lp{};
    lp.num_rows = original_.num_rows;
    lp.num_cols = original_.num_cols;
    lp.nnz = a.numNonzeros();
    lp.sense = original_.sense == Sense::Maximize ? 1 : 0;
    lp.row_starts = a.csr_row_starts().data();
    lp.col_indices = a.csr_col_indices().data();
    lp.values = a.csr_values().data();
    lp.obj = original_.obj.data();
    lp.col_lower = original_.col_lower.data();
    lp.col_upper = original_.col_upper.data();
    lp.row_lower = original_.row_lower.data();
    lp.row_upper = original_.row_upper.data();
    lp.obj_offset = original_.obj_offset;
    const mpx_pdlp_options o = toMi355x(options_);

    // setWarmStart vectors of the wrong size are ignored (pdlp.h:115-119).
    const bool warm = has_warm_start_ &&
                      static_cast<Index>(warm_start_x_.size()) == original_.num_cols &&
                      static_cast<Index>(warm_start_y_.size()) == original_.num_rows;
    primal_orig_.assign(static_cast<size_t>(original_.num_cols), 0.0);
    dual_orig_.assign(static_cast<size_t>(original_.num_rows), 0.0);
    mpx_pdlp_result r{};
    const int rc = mpx_pdlp_solve(&lp, &o, warm ? warm_start_x_.data() : nullptr,
                                  warm ? warm_start_y_.data() : nullptr, primal_orig_.data(),
                                  dual_orig_.data(), &r);
    if (rc == MPX_ERR_OUT_OF_MEMORY) {
        // The reference falls back to its CPU loop when device memory
        // cannot be had (pdlp.cpp:1767-1771): the same solve with the gate
        // closed for this one call.
        options_.use_gpu = false;
        const LpResult cpu = solve();
        options_.use_gpu = true;
        return cpu;
    }
    if (rc != MPX_OK) {
        // A device failure is an error status, as at pdlp.cpp:1979-1983;
        // mpx_last_error() holds the text.
        status_ = Status::Error;
        objective_ = 0.0;
        iterations_ = 0;
        return {status_, objective_, iterations_, 0.0};
    }
    used_gpu_ = r.status != MPX_STATUS_ERROR;
    // MPX_STATUS_* follow mipx::Status (include/mipx/core.h:11-19).
    status_ = static_cast<Status>(r.status);
    objective_ = r.objective;
    iterations_ = static_cast<Int>(r.iterations);
    // Work units: PDHG iterations x (4 nnz + 2 (m + n)) (pdlp.cpp:2786-2791).
    return {status_, objective_, iterations_, r.work_units};
}

}  // namespace mipx
