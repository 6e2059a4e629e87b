"""Loaders for the golden fixtures under tests/golden/."""
import json
import os

import numpy as np

from mipx_amd.lp import from_triplets

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STATUS = {"Optimal": 0, "Infeasible": 1, "Unbounded": 2, "IterLimit": 3, "TimeLimit": 4,
          "NodeLimit": 5, "Error": 6}


def _num(v):
    if v == "inf":
        return np.inf
    if v == "-inf":
        return -np.inf
    return float(v)


def closed_form_cases():
    with open(os.path.join(GOLDEN, "closed_form_lps.json")) as f:
        data = json.load(f)
    out = []
    for c in data["cases"]:
        lp = from_triplets(c["m"], c["n"], c["triplets"], [_num(v) for v in c["obj"]],
                           [_num(v) for v in c["col_lower"]], [_num(v) for v in c["col_upper"]],
                           [_num(v) for v in c["row_lower"]], [_num(v) for v in c["row_upper"]],
                           0, c.get("offset", 0.0), c["name"])
        out.append((c, lp))
    return out


def spmv_cases():
    with open(os.path.join(GOLDEN, "spmv_kat.json")) as f:
        data = json.load(f)
    out = []
    for c in data["cases"]:
        lp = from_triplets(c["m"], c["n"], c["triplets"], np.zeros(c["n"]), np.zeros(c["n"]),
                           np.zeros(c["n"]), np.zeros(c["m"]), np.zeros(c["m"]))
        out.append((c, lp))
    return out


def medium_lp():
    """The reference's buildMediumLp shape (tests/test_pdlp.cpp:260-299):
    50 x 80, ~60% dense, values 0.1 + 2U, <= rows with 50 + 2i, 0 <= x <= 100,
    alternating objective.  std::mt19937 is not reproducible from Python, so
    the density pattern and values come from numpy's PCG64 with seed 123."""
    rows, cols = 50, 80
    rng = np.random.Generator(np.random.PCG64(123))
    obj = np.array([1.0 + 0.1 * j if j % 2 == 0 else -(0.5 + 0.05 * j) for j in range(cols)])
    trips = []
    for i in range(rows):
        for j in range(cols):
            if rng.uniform() < 0.6:
                trips.append((i, j, 0.1 + rng.uniform() * 2.0))
    return from_triplets(rows, cols, trips, obj, np.zeros(cols), np.full(cols, 100.0),
                         np.full(rows, -np.inf), np.array([50.0 + 2.0 * i for i in range(rows)]))


def highs_objective(lp):
    """Exact optimum of a (feasible, bounded) LP from scipy's HiGHS, as an
    independent optimum for accuracy checks."""
    import scipy.sparse as sp
    from scipy.optimize import linprog
    a = sp.csr_matrix((lp.values, lp.col_indices, lp.row_starts), shape=(lp.num_rows, lp.num_cols))
    eq = lp.row_lower == lp.row_upper
    ub_hi = ~eq & np.isfinite(lp.row_upper)
    ub_lo = ~eq & np.isfinite(lp.row_lower)
    import scipy.sparse as sps
    a_ub = sps.vstack([a[ub_hi], -a[ub_lo]])
    b_ub = np.concatenate([lp.row_upper[ub_hi], -lp.row_lower[ub_lo]])
    sign = -1.0 if lp.sense else 1.0
    res = linprog(sign * lp.obj, A_ub=a_ub, b_ub=b_ub, A_eq=a[eq], b_eq=lp.row_lower[eq],
                  bounds=list(zip(lp.col_lower, lp.col_upper)), method="highs")
    assert res.status == 0, res.message
    return float(sign * res.fun + lp.obj_offset)
