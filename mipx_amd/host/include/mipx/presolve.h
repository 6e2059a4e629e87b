// presolve.h -- LP presolve / postsolve in front of the PDLP solve.
//
// Interface after the reference's mipx::Presolver (include/mipx/presolve.h:
// presolve(), postsolve(), isInfeasible(), stats(), colMapping(),
// setMaxRounds()), which its CLI runs before an LP solve by default
// (src/cli/main.cpp:455-467) and undoes on the primal values afterwards
// (560-568).  The reductions here are those that keep a PDLP answer whole:
// the reduced LP's primal AND dual solution map back to an original-space
// pair that satisfies the same KKT conditions, so statuses, solution and
// dual vectors of the solve survive (postsolveDual is an extension; the
// reference restores primal values only).
//
//   empty rows       0 in [rl, ru] or the LP is infeasible; dual 0
//   fixed columns    l == u: substituted into the rows and the objective offset
//   singleton rows   rl <= a x_j <= ru becomes a bound on x_j; the row's
//                    multiplier is recovered from x_j's reduced cost when the
//                    row's bound is the active one
//   empty columns    fixed at the bound their cost prefers (kept in the LP
//                    when that bound is infinite, for the solver to report
//                    unboundedness)
// repeated until nothing changes (or setMaxRounds rounds).
#pragma once

#include <vector>

#include "mipx/core.h"
#include "mipx/lp_problem.h"

namespace mipx {

struct PresolveStats {
    Index vars_removed = 0;
    Index rows_removed = 0;
    Index fixed_var_changes = 0;
    Index singleton_row_changes = 0;
    Index empty_row_changes = 0;
    Index empty_col_changes = 0;
    Index rounds = 0;
    Real time_seconds = 0.0;
};

class Presolver {
public:
    Presolver() = default;

    // Reduced copy of `problem` (the input is not modified).
    LpProblem presolve(const LpProblem& problem);
    // Original-space primal values from the reduced LP's.
    std::vector<Real> postsolve(const std::vector<Real>& presolved_solution) const;
    // Original-space duals (mipx's sign convention, reduced costs c + A'y)
    // from the reduced LP's duals and the postsolved primal values.
    std::vector<Real> postsolveDual(const std::vector<Real>& presolved_duals,
                                    const std::vector<Real>& full_primal) const;

    [[nodiscard]] bool isInfeasible() const { return infeasible_; }
    [[nodiscard]] const PresolveStats& stats() const { return stats_; }
    // Reduced column k is original column colMapping()[k]; same for rows.
    [[nodiscard]] const std::vector<Index>& colMapping() const { return col_map_; }
    [[nodiscard]] const std::vector<Index>& rowMapping() const { return row_map_; }
    void setMaxRounds(Index rounds) { max_rounds_ = rounds; }

private:
    enum class OpKind { FixCol, EmptyRow, SingletonRow };
    struct Op {
        OpKind kind;
        Index row = -1, col = -1;
        Real value = 0.0;        // FixCol: the fixed value; SingletonRow: the coefficient
        bool lo_from_row = false, hi_from_row = false;  // SingletonRow: bounds it set
        Real lo = 0.0, hi = 0.0;  // SingletonRow: the column's bounds after the row
    };

    LpProblem original_;
    std::vector<Index> cs_, ri_;  // A in CSC (rows increasing inside a column)
    std::vector<Real> cv_;
    std::vector<Op> ops_;
    std::vector<Index> col_map_, row_map_;
    PresolveStats stats_;
    bool infeasible_ = false;
    Index max_rounds_ = 100;
};

}  // namespace mipx
