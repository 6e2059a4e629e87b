"""Writes tests/golden/highs_optima.json: optimal objectives of seeded
synthetic LPs (mipx_amd.lpgen.c0_plain) from scipy's HiGHS (dual simplex /
IPM with crossover), as an independent optimum for the PDLP oracle and the
GPU solver.  Run from the repository root after building the generator."""
import json
import os
import sys

import numpy as np
import scipy.sparse as sp
from scipy.optimize import linprog

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from mipx_amd import lpgen  # noqa: E402

CASES = [(0xC0, 2000, 5000, 5), (101, 300, 750, 5), (102, 500, 1250, 4), (103, 200, 500, 6)]


def highs(lp):
    a = sp.csr_matrix((lp.values, lp.col_indices, lp.row_starts),
                      shape=(lp.num_rows, lp.num_cols))
    eq = lp.row_lower == lp.row_upper
    res = linprog(lp.obj, A_ub=a[~eq], b_ub=lp.row_upper[~eq], A_eq=a[eq], b_eq=lp.row_lower[eq],
                  bounds=list(zip(lp.col_lower, lp.col_upper)), method="highs")
    assert res.status == 0, res.message
    return float(res.fun)


def main():
    out = {"_source": "scipy.optimize.linprog(method='highs') on mipx_amd.lpgen.c0_plain LPs",
           "cases": []}
    for seed, m, n, k in CASES:
        g = lpgen.c0_plain(seed=seed, m=m, n=n, per_col=k)
        out["cases"].append(dict(seed=seed, m=m, n=n, per_col=k, objective=highs(g.lp)))
        print(out["cases"][-1])
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "highs_optima.json"),
              "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
