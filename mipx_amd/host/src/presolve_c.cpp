// presolve_c.cpp -- the C-ABI of include/mipx_presolve.h over mipx::Presolver.
#include <exception>
#include <vector>

#include "mipx/presolve.h"
#include "mipx_presolve.h"

struct mpx_presolve {
    mipx::Presolver pre;
    mipx::LpProblem reduced;
    mipx::Index orig_rows = 0, orig_cols = 0;
};

extern "C" {

int mpx_presolve_run(const mpx_lp* lp, mpx_presolve** out) {
    if (!lp || !out || lp->num_rows < 0 || lp->num_cols < 0 || lp->nnz < 0 ||
        (lp->num_rows > 0 && (!lp->row_starts || !lp->row_lower || !lp->row_upper)) ||
        (lp->num_cols > 0 && (!lp->obj || !lp->col_lower || !lp->col_upper)) ||
        (lp->nnz > 0 && (!lp->col_indices || !lp->values)))
        return MPX_ERR_INVALID;
    try {
        mipx::LpProblem p;
        p.num_rows = lp->num_rows;
        p.num_cols = lp->num_cols;
        p.sense = lp->sense ? mipx::Sense::Maximize : mipx::Sense::Minimize;
        p.obj_offset = lp->obj_offset;
        p.obj.assign(lp->obj, lp->obj + lp->num_cols);
        p.col_lower.assign(lp->col_lower, lp->col_lower + lp->num_cols);
        p.col_upper.assign(lp->col_upper, lp->col_upper + lp->num_cols);
        p.row_lower.assign(lp->row_lower, lp->row_lower + lp->num_rows);
        p.row_upper.assign(lp->row_upper, lp->row_upper + lp->num_rows);
        std::vector<mipx::Index> starts(1, 0);
        if (lp->num_rows > 0) starts.assign(lp->row_starts, lp->row_starts + lp->num_rows + 1);
        p.matrix = mipx::SparseMatrix(lp->num_rows, lp->num_cols,
                                      lp->nnz ? std::vector<double>(lp->values, lp->values + lp->nnz)
                                              : std::vector<double>(),
                                      lp->nnz ? std::vector<mipx::Index>(lp->col_indices,
                                                                         lp->col_indices + lp->nnz)
                                              : std::vector<mipx::Index>(),
                                      std::move(starts));
        auto* h = new mpx_presolve();
        h->orig_rows = lp->num_rows;
        h->orig_cols = lp->num_cols;
        h->reduced = h->pre.presolve(p);
        *out = h;
    } catch (const std::exception&) {
        return MPX_ERR_OUT_OF_MEMORY;
    }
    return MPX_OK;
}

int mpx_presolve_reduced(const mpx_presolve* p, mpx_lp* r) {
    if (!p || !r) return MPX_ERR_INVALID;
    const auto& q = p->reduced;
    r->num_rows = q.num_rows;
    r->num_cols = q.num_cols;
    r->nnz = q.matrix.numNonzeros();
    r->sense = q.sense == mipx::Sense::Maximize ? 1 : 0;
    r->row_starts = q.matrix.csr_row_starts().data();
    r->col_indices = q.matrix.csr_col_indices().data();
    r->values = q.matrix.csr_values().data();
    r->obj = q.obj.data();
    r->col_lower = q.col_lower.data();
    r->col_upper = q.col_upper.data();
    r->row_lower = q.row_lower.data();
    r->row_upper = q.row_upper.data();
    r->obj_offset = q.obj_offset;
    return MPX_OK;
}

int mpx_presolve_get_stats(const mpx_presolve* p, mpx_presolve_stats* s) {
    if (!p || !s) return MPX_ERR_INVALID;
    const auto& st = p->pre.stats();
    s->vars_removed = st.vars_removed;
    s->rows_removed = st.rows_removed;
    s->fixed_var_changes = st.fixed_var_changes;
    s->singleton_row_changes = st.singleton_row_changes;
    s->empty_row_changes = st.empty_row_changes;
    s->empty_col_changes = st.empty_col_changes;
    s->rounds = st.rounds;
    s->infeasible = p->pre.isInfeasible() ? 1 : 0;
    s->seconds = st.time_seconds;
    return MPX_OK;
}

int mpx_presolve_postsolve(const mpx_presolve* p, const double* x_red, const double* y_red,
                           double* x, double* y) {
    if (!p || !x_red || !x || (y && !y_red)) return MPX_ERR_INVALID;
    try {
        const auto& q = p->reduced;
        const std::vector<double> xr(x_red, x_red + q.num_cols);
        const std::vector<double> xf = p->pre.postsolve(xr);
        for (mipx::Index j = 0; j < p->orig_cols; ++j) x[j] = xf[static_cast<size_t>(j)];
        if (y) {
            const std::vector<double> yr(y_red, y_red + q.num_rows);
            const std::vector<double> yf = p->pre.postsolveDual(yr, xf);
            for (mipx::Index i = 0; i < p->orig_rows; ++i) y[i] = yf[static_cast<size_t>(i)];
        }
    } catch (const std::exception&) {
        return MPX_ERR_OUT_OF_MEMORY;
    }
    return MPX_OK;
}

int mpx_presolve_destroy(mpx_presolve* p) {
    delete p;
    return MPX_OK;
}

}  // extern "C"
