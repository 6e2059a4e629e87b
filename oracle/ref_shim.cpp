// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C entry points around the reference's own CPU PdlpSolver, compiled together
// with the reference sources where they lie (src/lp/pdlp.cpp,
// src/lp/sparse_matrix.cpp, src/lp/lp_problem.cpp; see oracle/Makefile) into
// oracle/_ref/libref_pdlp.so.  The reference is built without MIPX_HAS_CUDA,
// so PdlpSolver::solve() runs its CPU path (pdlp.cpp:862-869).  This file
// holds no reference code: it only converts the mpx_lp / mpx_pdlp_options
// structs of include/mipx_pdlp.h into mipx::LpProblem / mipx::PdlpOptions and
// copies results back.

#include <atomic>
#include <cstring>
#include <vector>

#include "mipx/lp_problem.h"
#include "mipx/pdlp.h"
#include "mipx/sparse_matrix.h"

#include "../include/mipx_pdlp.h"

namespace {

mipx::LpProblem toProblem(const mpx_lp& lp) {
    mipx::LpProblem p;
    p.name = "ref_shim";
    p.sense = lp.sense ? mipx::Sense::Maximize : mipx::Sense::Minimize;
    p.num_cols = lp.num_cols;
    p.num_rows = lp.num_rows;
    p.obj.assign(lp.obj, lp.obj + lp.num_cols);
    p.col_lower.assign(lp.col_lower, lp.col_lower + lp.num_cols);
    p.col_upper.assign(lp.col_upper, lp.col_upper + lp.num_cols);
    p.col_type.assign(static_cast<size_t>(lp.num_cols), mipx::VarType::Continuous);
    p.row_lower.assign(lp.row_lower, lp.row_lower + lp.num_rows);
    p.row_upper.assign(lp.row_upper, lp.row_upper + lp.num_rows);
    const int nnz = lp.num_rows > 0 ? lp.row_starts[lp.num_rows] : 0;
    std::vector<double> v(lp.values, lp.values + nnz);
    std::vector<int> ci(lp.col_indices, lp.col_indices + nnz);
    std::vector<int> rs(lp.row_starts, lp.row_starts + lp.num_rows + 1);
    p.matrix = mipx::SparseMatrix(lp.num_rows, lp.num_cols, std::move(v), std::move(ci),
                                  std::move(rs));
    p.obj_offset = lp.obj_offset;
    return p;
}

mipx::PdlpOptions toOptions(const mpx_pdlp_options& o) {
    mipx::PdlpOptions r;
    r.max_iter = o.max_iter;
    r.optimality_tol = o.optimality_tol;
    r.primal_tol = o.primal_tol;
    r.dual_tol = o.dual_tol;
    r.max_solve_seconds = o.max_solve_seconds;
    r.optimality_norm =
        o.optimality_norm ? mipx::PdlpOptimalityNorm::LInf : mipx::PdlpOptimalityNorm::L2;
    r.sv_max_iter = o.sv_max_iter;
    r.sv_tol = o.sv_tol;
    r.adaptive_step = o.adaptive_step != 0;
    r.adaptive_step_down = o.adaptive_step_down;
    r.adaptive_step_up = o.adaptive_step_up;
    r.adaptive_step_min = o.adaptive_step_min;
    r.adaptive_step_max = o.adaptive_step_max;
    r.primal_weight = o.primal_weight;
    r.update_primal_weight = o.update_primal_weight != 0;
    r.pid_kp = o.pid_kp;
    r.pid_ki = o.pid_ki;
    r.pid_kd = o.pid_kd;
    r.pid_i_smooth = o.pid_i_smooth;
    r.restart_sufficient_decay = o.restart_sufficient_decay;
    r.restart_necessary_decay = o.restart_necessary_decay;
    r.restart_artificial_fraction = o.restart_artificial_fraction;
    r.termination_eval_frequency = o.termination_eval_frequency;
    r.matrix_zero_tol = o.matrix_zero_tol;
    r.enable_polishing = o.enable_polishing != 0;
    r.polish_primal_tol = o.polish_primal_tol;
    r.polish_dual_tol = o.polish_dual_tol;
    r.polish_max_iter = o.polish_max_iter;
    r.do_ruiz_scaling = o.do_ruiz_scaling != 0;
    r.ruiz_iterations = o.ruiz_iterations;
    r.do_pock_chambolle_scaling = o.do_pock_chambolle_scaling != 0;
    r.do_bound_obj_rescaling = o.do_bound_obj_rescaling != 0;
    r.preconditioner_refresh = o.preconditioner_refresh != 0;
    r.preconditioner_refresh_ratio = o.preconditioner_refresh_ratio;
    r.anderson_acceleration = o.anderson_acceleration != 0;
    r.anderson_window_size = o.anderson_window_size;
    r.anderson_regularization = o.anderson_regularization;
    r.use_gpu = false;
    r.verbose = o.verbose != 0;
    return r;
}

}  // namespace

extern "C" int ref_pdlp_solve(const mpx_lp* lp, const mpx_pdlp_options* opts,
                              const double* warm_x, const double* warm_y, double* x_out,
                              double* y_out, mpx_pdlp_result* res) {
    if (!lp || !opts) return MPX_ERR_INVALID;
    mipx::PdlpSolver solver;
    solver.setOptions(toOptions(*opts));
    solver.load(toProblem(*lp));
    if (warm_x && warm_y)
        solver.setWarmStart(std::span<const double>(warm_x, static_cast<size_t>(lp->num_cols)),
                            std::span<const double>(warm_y, static_cast<size_t>(lp->num_rows)));
    const mipx::LpResult r = solver.solve();
    const auto x = solver.getPrimalValues();
    const auto y = solver.getDualValues();
    if (x_out) std::memcpy(x_out, x.data(), x.size() * sizeof(double));
    if (y_out) std::memcpy(y_out, y.data(), y.size() * sizeof(double));
    if (res) {
        std::memset(res, 0, sizeof(*res));
        res->status = static_cast<int32_t>(r.status);
        res->objective = r.objective;
        res->iterations = r.iterations;
        res->work_units = r.work_units;
    }
    return MPX_OK;
}
