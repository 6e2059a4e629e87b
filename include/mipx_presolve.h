/*
 * mipx_presolve.h -- C-ABI of the LP presolve in front of the PDLP solve
 * (exported by libmipx_host.so, the C++ host library; no device code).
 *
 * Replaces the reference's mipx::Presolver as the CLI uses it before an LP
 * solve (src/cli/main.cpp:455-467: presolve unless --no-presolve;
 * 560-568: postsolve of the primal values and the objective recomputed on
 * them); reference implementation src/presolve/presolve.cpp:1628-1814
 * (Presolver::presolve) and 231-320 (PostsolveStack::postsolve).  The
 * reductions (empty rows, fixed columns, singleton rows to bounds, empty
 * columns) keep a PDLP answer whole: mpx_presolve_postsolve maps the
 * reduced LP's primal and dual solution to an original-space pair with the
 * same KKT quality.  Plain C types, int error codes (MPX_OK == 0, the codes
 * of mipx_pdlp.h).
 */
#ifndef MIPX_PRESOLVE_H
#define MIPX_PRESOLVE_H

#include "mipx_pdlp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpx_presolve mpx_presolve;

/* Counters of PresolveStats (reference include/mipx/presolve.h). */
typedef struct mpx_presolve_stats {
    int32_t vars_removed;
    int32_t rows_removed;
    int32_t fixed_var_changes;
    int32_t singleton_row_changes;
    int32_t empty_row_changes;
    int32_t empty_col_changes;
    int32_t rounds;
    int32_t infeasible;        /* 1: presolve proved the LP infeasible */
    double seconds;
} mpx_presolve_stats;

/* Presolves a copy of `lp` (arrays only read during the call). */
int mpx_presolve_run(const mpx_lp* lp, mpx_presolve** out);
/* The reduced LP (host arrays owned by the handle, valid until destroy). */
int mpx_presolve_reduced(const mpx_presolve* p, mpx_lp* reduced);
int mpx_presolve_get_stats(const mpx_presolve* p, mpx_presolve_stats* stats);
/* x_red[reduced cols], y_red[reduced rows] -> x[orig cols], y[orig rows];
 * y_red / y may be NULL (primal only, as the reference's postsolve). */
int mpx_presolve_postsolve(const mpx_presolve* p, const double* x_red, const double* y_red,
                           double* x, double* y);
int mpx_presolve_destroy(mpx_presolve* p);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* MIPX_PRESOLVE_H */
