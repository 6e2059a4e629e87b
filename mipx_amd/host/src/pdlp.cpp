// pdlp.cpp -- mipx::PdlpSolver over the C-ABI of libmipx_pdlp.so.
//
// load() copies the problem (reference pdlp.cpp:445-455); solve() hands it to
// mpx_pdlp_solve(), which does the scaling, power iteration, PDHG loop,
// polish and recovery on the device; results are returned by value.  The
// mutators edit the stored problem and, like the reference's (pdlp.cpp:2833-2907),
// silently ignore invalid arguments; the next solve starts from the edited LP.
#include "mipx/pdlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "mipx_pdlp.h"

namespace mipx {

namespace {

int stopThunk(void* user) {
    const auto* flag = static_cast<const std::atomic<bool>*>(user);
    return flag->load(std::memory_order_relaxed) ? 1 : 0;
}

mpx_pdlp_options toC(const PdlpOptions& p) {
    mpx_pdlp_options o;
    mpx_pdlp_default_options(&o);
    o.max_iter = p.max_iter;
    o.optimality_tol = p.optimality_tol;
    o.primal_tol = p.primal_tol;
    o.dual_tol = p.dual_tol;
    o.max_solve_seconds = p.max_solve_seconds;
    o.optimality_norm = p.optimality_norm == PdlpOptimalityNorm::LInf ? 1 : 0;
    o.sv_max_iter = p.sv_max_iter;
    o.sv_tol = p.sv_tol;
    o.adaptive_step = p.adaptive_step;
    o.adaptive_step_down = p.adaptive_step_down;
    o.adaptive_step_up = p.adaptive_step_up;
    o.adaptive_step_min = p.adaptive_step_min;
    o.adaptive_step_max = p.adaptive_step_max;
    o.primal_weight = p.primal_weight;
    o.update_primal_weight = p.update_primal_weight;
    o.pid_kp = p.pid_kp;
    o.pid_ki = p.pid_ki;
    o.pid_kd = p.pid_kd;
    o.pid_i_smooth = p.pid_i_smooth;
    o.restart_sufficient_decay = p.restart_sufficient_decay;
    o.restart_necessary_decay = p.restart_necessary_decay;
    o.restart_artificial_fraction = p.restart_artificial_fraction;
    o.termination_eval_frequency = p.termination_eval_frequency;
    o.matrix_zero_tol = p.matrix_zero_tol;
    o.enable_polishing = p.enable_polishing;
    o.polish_primal_tol = p.polish_primal_tol;
    o.polish_dual_tol = p.polish_dual_tol;
    o.polish_max_iter = p.polish_max_iter;
    o.do_ruiz_scaling = p.do_ruiz_scaling;
    o.ruiz_iterations = p.ruiz_iterations;
    o.do_pock_chambolle_scaling = p.do_pock_chambolle_scaling;
    o.do_bound_obj_rescaling = p.do_bound_obj_rescaling;
    o.preconditioner_refresh = p.preconditioner_refresh;
    o.preconditioner_refresh_ratio = p.preconditioner_refresh_ratio;
    o.anderson_acceleration = p.anderson_acceleration;
    o.anderson_window_size = p.anderson_window_size;
    o.anderson_regularization = p.anderson_regularization;
    o.use_gpu = p.use_gpu;
    o.gpu_min_rows = p.gpu_min_rows;
    o.gpu_min_nnz = p.gpu_min_nnz;
    o.verbose = p.verbose;
    if (p.stop_flag) {
        o.stop_fn = &stopThunk;
        o.stop_user = const_cast<std::atomic<bool>*>(p.stop_flag);
    }
    o.precision = p.precision == PdlpPrecision::Mixed ? MPX_PREC_MIXED : MPX_PREC_FP64;
    o.spmv_group = p.spmv_group;
    o.device = p.device;
    o.mixed_handover = p.mixed_handover;
    o.spmv_chunk_cols = p.spmv_chunk_cols;
    o.refine_tol = p.refine_tol;
    o.mixed_refine = p.mixed_refine ? 1 : 0;
    return o;
}

}  // namespace

void PdlpSolver::load(const LpProblem& problem) {
    problem_ = problem;
    loaded_ = true;
    status_ = Status::Error;
    objective_ = 0.0;
    iterations_ = 0;
    has_warm_ = false;
    primal_.assign(static_cast<size_t>(problem_.num_cols), 0.0);
    dual_.assign(static_cast<size_t>(problem_.num_rows), 0.0);
}

void PdlpSolver::setWarmStart(std::span<const Real> x, std::span<const Real> y) {
    warm_x_.assign(x.begin(), x.end());
    warm_y_.assign(y.begin(), y.end());
    has_warm_ = true;
}

void PdlpSolver::clearWarmStart() {
    warm_x_.clear();
    warm_y_.clear();
    has_warm_ = false;
}

LpResult PdlpSolver::solve() {
    if (!loaded_) {
        status_ = Status::Error;
        return {status_, 0.0, 0, 0.0};
    }
    const auto& a = problem_.matrix;
    mpx_lp lp{};
    lp.num_rows = problem_.num_rows;
    lp.num_cols = problem_.num_cols;
    lp.nnz = a.numNonzeros();
    lp.sense = problem_.sense == Sense::Maximize ? 1 : 0;
    lp.row_starts = a.csr_row_starts().data();
    lp.col_indices = a.csr_col_indices().data();
    lp.values = a.csr_values().data();
    lp.obj = problem_.obj.data();
    lp.col_lower = problem_.col_lower.data();
    lp.col_upper = problem_.col_upper.data();
    lp.row_lower = problem_.row_lower.data();
    lp.row_upper = problem_.row_upper.data();
    lp.obj_offset = problem_.obj_offset;
    const mpx_pdlp_options o = toC(options_);
    const bool warm = has_warm_ && static_cast<Index>(warm_x_.size()) == problem_.num_cols &&
                      static_cast<Index>(warm_y_.size()) == problem_.num_rows;
    primal_.assign(static_cast<size_t>(problem_.num_cols), 0.0);
    dual_.assign(static_cast<size_t>(problem_.num_rows), 0.0);
    if (!passesGpuGate()) {
        // The reference runs its CPU loop for such LPs (pdlp.cpp:862-869);
        // that loop is not part of this library.
        used_gpu_ = false;
        last_error_ =
            "LP is below the GPU gate (use_gpu, gpu_min_rows, gpu_min_nnz): the CPU PDLP loop "
            "is not part of this library";
        status_ = Status::Error;
        objective_ = 0.0;
        iterations_ = 0;
        return {status_, 0.0, 0, 0.0};
    }
    mpx_pdlp_result r{};
    const int rc = mpx_pdlp_solve(&lp, &o, warm ? warm_x_.data() : nullptr,
                                  warm ? warm_y_.data() : nullptr, primal_.data(), dual_.data(), &r);
    used_gpu_ = rc == MPX_OK && problem_.num_cols > 0 && problem_.num_rows > 0 &&
                r.status != MPX_STATUS_ERROR;
    if (rc != MPX_OK) {
        last_error_ = mpx_last_error();
        status_ = Status::Error;
        objective_ = 0.0;
        iterations_ = 0;
        return {status_, 0.0, 0, 0.0};
    }
    if (r.status == MPX_STATUS_ERROR) last_error_ = mpx_last_error();
    status_ = static_cast<Status>(r.status);
    objective_ = r.objective;
    iterations_ = r.iterations;
    return {status_, objective_, iterations_, r.work_units};
}

bool PdlpSolver::passesGpuGate() const {
    // Trivial LPs are closed forms evaluated before the gate (pdlp.cpp:812-860).
    if (problem_.num_cols == 0 || problem_.num_rows == 0) return true;
    // The gate counts the entries that survive matrix_zero_tol (pdlp.cpp:499-518, 863-864).
    Index kept = problem_.matrix.numNonzeros();
    if (options_.matrix_zero_tol > 0.0) {
        kept = 0;
        for (const Real v : problem_.matrix.csr_values())
            if (std::abs(v) >= options_.matrix_zero_tol) ++kept;
    }
    return options_.use_gpu && problem_.num_rows >= options_.gpu_min_rows &&
           kept >= options_.gpu_min_nnz;
}

std::vector<Real> PdlpSolver::getReducedCosts() const {
    std::vector<Real> rc(static_cast<size_t>(problem_.num_cols), 0.0);
    if (!loaded_ || dual_.empty()) return rc;
    std::vector<Real> aty(rc.size(), 0.0);
    problem_.matrix.multiplyTranspose(dual_, aty);
    for (Index j = 0; j < problem_.num_cols; ++j) rc[j] = problem_.obj[j] + aty[j];
    return rc;
}

void PdlpSolver::addRows(std::span<const Index> starts, std::span<const Index> indices,
                         std::span<const Real> values, std::span<const Real> lower,
                         std::span<const Real> upper) {
    if (!loaded_ || starts.empty()) return;
    const Index count = static_cast<Index>(lower.size());
    if (count != static_cast<Index>(upper.size()) || static_cast<Index>(starts.size()) != count + 1)
        return;
    for (Index r = 0; r < count; ++r) {
        const Index s = starts[r], e = starts[r + 1];
        if (s < 0 || e < s || e > static_cast<Index>(indices.size()) ||
            e > static_cast<Index>(values.size()))
            return;
        problem_.matrix.addRow(indices.subspan(static_cast<size_t>(s), static_cast<size_t>(e - s)),
                               values.subspan(static_cast<size_t>(s), static_cast<size_t>(e - s)));
        problem_.row_lower.push_back(lower[r]);
        problem_.row_upper.push_back(upper[r]);
        problem_.row_names.emplace_back();
    }
    problem_.num_rows += count;
}

void PdlpSolver::removeRows(std::span<const Index> rows) {
    if (!loaded_ || rows.empty()) return;
    std::vector<Index> s(rows.begin(), rows.end());
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    for (auto it = s.rbegin(); it != s.rend(); ++it) {
        const Index r = *it;
        if (r < 0 || r >= problem_.num_rows) continue;
        problem_.matrix.removeRowStable(r);
        problem_.row_lower.erase(problem_.row_lower.begin() + r);
        problem_.row_upper.erase(problem_.row_upper.begin() + r);
        if (r < static_cast<Index>(problem_.row_names.size()))
            problem_.row_names.erase(problem_.row_names.begin() + r);
        --problem_.num_rows;
    }
}

void PdlpSolver::setColBounds(Index col, Real lower, Real upper) {
    if (!loaded_ || col < 0 || col >= problem_.num_cols) return;
    problem_.col_lower[col] = lower;
    problem_.col_upper[col] = upper;
}

void PdlpSolver::setObjective(std::span<const Real> obj) {
    if (!loaded_ || static_cast<Index>(obj.size()) != problem_.num_cols) return;
    problem_.obj.assign(obj.begin(), obj.end());
}

}  // namespace mipx
