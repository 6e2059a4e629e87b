// pdlp.h -- PdlpSolver with the reference's interface (include/mipx/pdlp.h:14-174),
// running every solve on the MI355X through the C-ABI of include/mipx_pdlp.h.
//
// Options keep the reference's names and defaults.  The reference's GPU gate
// (pdlp.cpp:862-869: use_gpu && rows >= gpu_min_rows && kept nonzeros >=
// gpu_min_nnz) is honoured: an LP the gate would send to the reference's CPU
// loop gets Status::Error with lastError() saying so, and usedGpu() is false,
// because that CPU loop is not part of this library (a caller that wants it
// keeps the reference's own CPU branch, as the binding of the reference does).
// The trivial n == 0 / m == 0 cases are solved in closed form before the
// gate, as in the reference.  Extensions: `precision` (FP64 or mixed
// FP32/FP64), `spmv_group`, `device`, `mixed_handover`, `spmv_chunk_cols`.
#pragma once

#include <atomic>
#include <span>
#include <vector>

#include "mipx/core.h"
#include "mipx/lp_problem.h"
#include "mipx/lp_solver.h"

namespace mipx {

enum class PdlpOptimalityNorm { L2, LInf };
enum class PdlpPrecision { FP64, Mixed };

struct PdlpOptions {
    Int max_iter = 4000000;
    Real optimality_tol = 5e-5;
    Real primal_tol = 5e-5;
    Real dual_tol = 5e-5;
    Real max_solve_seconds = -1.0;
    PdlpOptimalityNorm optimality_norm = PdlpOptimalityNorm::L2;
    Int sv_max_iter = 5000;
    Real sv_tol = 1e-4;
    bool adaptive_step = false;
    Real adaptive_step_down = 0.7;
    Real adaptive_step_up = 1.0;
    Real adaptive_step_min = 1e-6;
    Real adaptive_step_max = 1e3;
    Real primal_weight = 1.0;
    bool update_primal_weight = true;
    Real pid_kp = 0.99;
    Real pid_ki = 0.01;
    Real pid_kd = 0.0;
    Real pid_i_smooth = 0.3;
    Real restart_sufficient_decay = 0.2;
    Real restart_necessary_decay = 0.5;
    Real restart_artificial_fraction = 0.36;
    Int termination_eval_frequency = 200;
    Real matrix_zero_tol = 0.0;
    bool enable_polishing = true;
    Real polish_primal_tol = 1e-6;
    Real polish_dual_tol = 1e-6;
    Int polish_max_iter = 500000;
    bool do_ruiz_scaling = true;
    Int ruiz_iterations = 10;
    bool do_pock_chambolle_scaling = true;
    bool do_bound_obj_rescaling = true;
    bool preconditioner_refresh = false;
    Real preconditioner_refresh_ratio = 10.0;
    bool anderson_acceleration = false;
    Int anderson_window_size = 5;
    Real anderson_regularization = 1e-10;
    bool use_gpu = true;
    Int gpu_min_rows = 512;
    Int gpu_min_nnz = 10000;
    bool verbose = true;
    const std::atomic<bool>* stop_flag = nullptr;
    // MI355X extensions.
    PdlpPrecision precision = PdlpPrecision::FP64;
    Int spmv_group = 0;
    Int device = 0;
    Real mixed_handover = 3.0;  // FP32 -> FP64 at this KKT ratio (see mipx_pdlp.h)
    Int spmv_chunk_cols = 0;    // column chunks of the dual step's SpMV (see mipx_pdlp.h)
    Real refine_tol = -1.0;     // mixed precision: tolerance of the refinement (see mipx_pdlp.h)
    bool mixed_refine = true;   // mixed precision: FP32 refinement around an FP64 base
};

class PdlpSolver : public LpSolver {
public:
    PdlpSolver() = default;

    void load(const LpProblem& problem) override;
    LpResult solve() override;

    Status getStatus() const override { return status_; }
    Real getObjective() const override { return objective_; }
    std::vector<Real> getPrimalValues() const override { return primal_; }
    std::vector<Real> getDualValues() const override { return dual_; }
    std::vector<Real> getReducedCosts() const override;
    std::vector<BasisStatus> getBasis() const override { return {}; }
    void setBasis(std::span<const BasisStatus>) override {}

    void addRows(std::span<const Index> starts, std::span<const Index> indices,
                 std::span<const Real> values, std::span<const Real> lower,
                 std::span<const Real> upper) override;
    void removeRows(std::span<const Index> rows) override;
    void setColBounds(Index col, Real lower, Real upper) override;
    void setObjective(std::span<const Real> obj) override;

    // Original-space warm start; ignored at solve() when sizes do not match.
    void setWarmStart(std::span<const Real> x, std::span<const Real> y);
    void clearWarmStart();

    void setOptions(const PdlpOptions& options) { options_ = options; }
    [[nodiscard]] const PdlpOptions& options() const { return options_; }
    [[nodiscard]] bool usedGpu() const { return used_gpu_; }
    // Text of the last device error (empty when none).
    [[nodiscard]] const std::string& lastError() const { return last_error_; }
    // Whether solve() would run this LP on the GPU (the reference's gate).
    [[nodiscard]] bool passesGpuGate() const;

private:
    PdlpOptions options_{};
    LpProblem problem_;
    bool loaded_ = false;
    std::vector<Real> primal_, dual_;
    Status status_ = Status::Error;
    Real objective_ = 0.0;
    Int iterations_ = 0;
    bool used_gpu_ = false;
    std::vector<Real> warm_x_, warm_y_;
    bool has_warm_ = false;
    std::string last_error_;
};

}  // namespace mipx
