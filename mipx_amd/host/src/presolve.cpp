// presolve.cpp -- mipx::Presolver: reductions that keep a PDLP answer whole.
//
// The LP is kept as given (CSR plus a CSC built once) with alive flags on
// rows and columns, per-row / per-column counts of live nonzeros, and the
// bounds as they tighten.  Rounds sweep fixed columns, then empty and
// singleton rows, then empty columns, until a round changes nothing.  Every
// reduction is recorded so that postsolve() and postsolveDual() can replay
// them backwards.  Reference counterpart: Presolver::presolve and
// PostsolveStack::postsolve (src/presolve/presolve.cpp:1628-1814, 231-320),
// whose reduction set is larger and restores primal values only.
#include "mipx/presolve.h"

#include <algorithm>
#include <chrono>
#include <cmath>

namespace mipx {

namespace {

// Bounds closer than this (relative) are taken as equal when a singleton
// row's implied bounds cross the column's by rounding.
constexpr Real kSnapTol = 1e-9;

Real tolOf(Real v) { return kSnapTol * (1.0 + std::abs(v)); }

}  // namespace

LpProblem Presolver::presolve(const LpProblem& problem) {
    const auto t0 = std::chrono::steady_clock::now();
    original_ = problem;
    ops_.clear();
    stats_ = PresolveStats{};
    infeasible_ = false;
    const Index m = problem.num_rows, n = problem.num_cols;
    const auto rs = problem.matrix.csr_row_starts();
    const auto ci = problem.matrix.csr_col_indices();
    const auto va = problem.matrix.csr_values();
    const Index nnz = problem.matrix.numNonzeros();
    // CSC of A (rows increasing inside a column), kept for postsolveDual.
    cs_.assign(static_cast<size_t>(n) + 1, 0);
    ri_.assign(static_cast<size_t>(nnz), 0);
    cv_.assign(static_cast<size_t>(nnz), 0.0);
    auto& cs = cs_;
    auto& ri = ri_;
    auto& cv = cv_;
    for (Index k = 0; k < nnz; ++k) ++cs[static_cast<size_t>(ci[k]) + 1];
    for (Index j = 0; j < n; ++j) cs[j + 1] += cs[j];
    {
        std::vector<Index> next(cs.begin(), cs.end() - 1);
        for (Index i = 0; i < m; ++i)
            for (Index k = rs[i]; k < rs[i + 1]; ++k) {
                const Index at = next[ci[k]]++;
                ri[at] = i;
                cv[at] = va[k];
            }
    }
    std::vector<char> row_alive(static_cast<size_t>(m), 1), col_alive(static_cast<size_t>(n), 1);
    std::vector<Index> row_cnt(static_cast<size_t>(m), 0), col_cnt(static_cast<size_t>(n), 0);
    for (Index i = 0; i < m; ++i)
        for (Index k = rs[i]; k < rs[i + 1]; ++k)
            if (va[k] != 0.0) {
                ++row_cnt[i];
                ++col_cnt[ci[k]];
            }
    std::vector<Real> lo = problem.col_lower, hi = problem.col_upper;
    std::vector<Real> rl = problem.row_lower, ru = problem.row_upper;
    Real offset = problem.obj_offset;
    const Real sgn = problem.sense == Sense::Maximize ? -1.0 : 1.0;

    for (Index round = 0; round < max_rounds_ && !infeasible_; ++round) {
        Index changed = 0;
        // Fixed columns: substitute the value into their rows.
        for (Index j = 0; j < n; ++j) {
            if (!col_alive[j] || lo[j] != hi[j] || !std::isfinite(lo[j])) continue;
            const Real v = lo[j];
            for (Index k = cs[j]; k < cs[j + 1]; ++k) {
                const Index i = ri[k];
                if (!row_alive[i] || cv[k] == 0.0) continue;
                rl[i] -= cv[k] * v;
                ru[i] -= cv[k] * v;
                --row_cnt[i];
            }
            offset += problem.obj[j] * v;
            col_alive[j] = 0;
            ops_.push_back({OpKind::FixCol, -1, j, v});
            ++stats_.fixed_var_changes;
            ++changed;
        }
        // Empty and singleton rows.
        for (Index i = 0; i < m && !infeasible_; ++i) {
            if (!row_alive[i] || row_cnt[i] > 1) continue;
            if (row_cnt[i] == 0) {
                if (rl[i] > tolOf(rl[i]) || ru[i] < -tolOf(ru[i])) {
                    infeasible_ = true;
                    break;
                }
                row_alive[i] = 0;
                ops_.push_back({OpKind::EmptyRow, i});
                ++stats_.empty_row_changes;
                ++changed;
                continue;
            }
            Index j = -1;
            Real a = 0.0;
            for (Index k = rs[i]; k < rs[i + 1]; ++k)
                if (col_alive[ci[k]] && va[k] != 0.0) {
                    j = ci[k];
                    a = va[k];
                    break;
                }
            const Real ilo = a > 0.0 ? rl[i] / a : ru[i] / a;
            const Real ihi = a > 0.0 ? ru[i] / a : rl[i] / a;
            const bool lo_from = ilo > lo[j], hi_from = ihi < hi[j];
            Real nlo = lo_from ? ilo : lo[j], nhi = hi_from ? ihi : hi[j];
            if (nlo > nhi) {
                if (nlo - nhi > tolOf(nlo)) {
                    infeasible_ = true;
                    break;
                }
                nlo = nhi = 0.5 * (nlo + nhi);
            }
            lo[j] = nlo;
            hi[j] = nhi;
            row_alive[i] = 0;
            --col_cnt[j];
            Op op{OpKind::SingletonRow, i, j, a};
            op.lo_from_row = lo_from;
            op.hi_from_row = hi_from;
            op.lo = nlo;
            op.hi = nhi;
            ops_.push_back(op);
            ++stats_.singleton_row_changes;
            ++changed;
        }
        if (infeasible_) break;
        // Empty columns: fixed where their cost points, when that bound is finite.
        for (Index j = 0; j < n; ++j) {
            if (!col_alive[j] || col_cnt[j] != 0) continue;
            if (lo[j] > hi[j] + tolOf(lo[j])) {
                infeasible_ = true;
                break;
            }
            const Real c = sgn * problem.obj[j];
            Real v;
            if (c > 0.0) {
                if (!std::isfinite(lo[j])) continue;
                v = lo[j];
            } else if (c < 0.0) {
                if (!std::isfinite(hi[j])) continue;
                v = hi[j];
            } else {
                v = std::clamp(0.0, lo[j], hi[j]);
            }
            offset += problem.obj[j] * v;
            col_alive[j] = 0;
            ops_.push_back({OpKind::FixCol, -1, j, v});
            ++stats_.empty_col_changes;
            ++changed;
        }
        ++stats_.rounds;
        if (changed == 0) break;
    }
    for (Index j = 0; j < n && !infeasible_; ++j)
        if (col_alive[j] && lo[j] > hi[j] + tolOf(lo[j])) infeasible_ = true;

    // The reduced LP: live rows and columns in their original order.
    col_map_.clear();
    row_map_.clear();
    std::vector<Index> new_col(static_cast<size_t>(n), -1);
    for (Index j = 0; j < n; ++j)
        if (col_alive[j]) {
            new_col[j] = static_cast<Index>(col_map_.size());
            col_map_.push_back(j);
        }
    for (Index i = 0; i < m; ++i)
        if (row_alive[i]) row_map_.push_back(i);
    LpProblem out;
    out.name = problem.name;
    out.sense = problem.sense;
    out.obj_offset = offset;
    out.num_cols = static_cast<Index>(col_map_.size());
    out.num_rows = static_cast<Index>(row_map_.size());
    for (const Index j : col_map_) {
        out.obj.push_back(problem.obj[j]);
        out.col_lower.push_back(lo[j]);
        out.col_upper.push_back(hi[j]);
        if (static_cast<size_t>(j) < problem.col_type.size()) out.col_type.push_back(problem.col_type[j]);
        if (static_cast<size_t>(j) < problem.col_names.size()) out.col_names.push_back(problem.col_names[j]);
    }
    std::vector<Real> values;
    std::vector<Index> cols, starts{0};
    for (const Index i : row_map_) {
        for (Index k = rs[i]; k < rs[i + 1]; ++k)
            if (col_alive[ci[k]]) {
                values.push_back(va[k]);
                cols.push_back(new_col[ci[k]]);
            }
        starts.push_back(static_cast<Index>(values.size()));
        out.row_lower.push_back(rl[i]);
        out.row_upper.push_back(ru[i]);
        if (static_cast<size_t>(i) < problem.row_names.size()) out.row_names.push_back(problem.row_names[i]);
    }
    out.matrix = SparseMatrix(out.num_rows, out.num_cols, std::move(values), std::move(cols),
                              std::move(starts));
    stats_.vars_removed = n - out.num_cols;
    stats_.rows_removed = m - out.num_rows;
    stats_.time_seconds =
        std::chrono::duration<Real>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

std::vector<Real> Presolver::postsolve(const std::vector<Real>& presolved_solution) const {
    std::vector<Real> x(static_cast<size_t>(original_.num_cols), 0.0);
    const size_t k_end = std::min(presolved_solution.size(), col_map_.size());
    for (size_t k = 0; k < k_end; ++k) x[static_cast<size_t>(col_map_[k])] = presolved_solution[k];
    for (auto it = ops_.rbegin(); it != ops_.rend(); ++it)
        if (it->kind == OpKind::FixCol) x[static_cast<size_t>(it->col)] = it->value;
    return x;
}

std::vector<Real> Presolver::postsolveDual(const std::vector<Real>& presolved_duals,
                                           const std::vector<Real>& full_primal) const {
    const Index m = original_.num_rows, n = original_.num_cols;
    std::vector<Real> y(static_cast<size_t>(m), 0.0);
    const size_t k_end = std::min(presolved_duals.size(), row_map_.size());
    for (size_t k = 0; k < k_end; ++k) y[static_cast<size_t>(row_map_[k])] = presolved_duals[k];
    if (static_cast<Index>(full_primal.size()) != n) return y;
    const Real sgn = original_.sense == Sense::Maximize ? -1.0 : 1.0;
    // Singleton rows backwards: when the bound a row imposed on its column
    // is the active one, the column's reduced cost (original objective and
    // every multiplier assigned so far) moves onto the row, leaving the
    // column with a zero reduced cost; the row's multiplier then has the
    // sign its active side allows (equality rows: either sign).
    for (auto it = ops_.rbegin(); it != ops_.rend(); ++it) {
        if (it->kind != OpKind::SingletonRow) continue;
        const Index j = it->col;
        Real g = original_.obj[j];
        for (Index k = cs_[j]; k < cs_[j + 1]; ++k) g += cv_[k] * y[static_cast<size_t>(ri_[k])];
        g *= sgn;
        const Real xj = full_primal[static_cast<size_t>(j)];
        const bool at_lo = it->lo_from_row && xj <= it->lo + tolOf(it->lo);
        const bool at_hi = it->hi_from_row && xj >= it->hi - tolOf(it->hi);
        if ((at_lo && at_hi) || (at_lo && g > 0.0) || (at_hi && g < 0.0))
            y[static_cast<size_t>(it->row)] = -g / (sgn * it->value);
    }
    return y;
}

}  // namespace mipx
