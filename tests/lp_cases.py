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


def certify(lp, r, x, y, tol=1e-6, tag=""):
    """Long-double KKT certificate of a returned Optimal (x, y) in original
    units (oracle.orc_kkt_check), checked three ways:
      * primal residual and gap within the tolerance, dual signs exact;
      * the solver's own last-evaluated rel_primal / rel_gap agree with the
        certificate at the returned point (1e-3 of the tolerance);
      * the solver's rel_dual bounds the certificate's from above (the
        solver counts a column's reduced cost unless the iterate sits at the
        bound, the certificate whenever a finite bound can absorb it), so a
        dual within the tolerance is certified too.
    Returns the certificate."""
    from oracle import binding as ob
    k = ob.kkt_check(lp, x, y)
    assert k["rel_primal"] <= tol * 1.001, (tag, k)
    assert k["rel_gap"] <= tol * 1.001, (tag, k)
    assert k["rel_dual"] <= tol * 1.001, (tag, k)
    assert k["dual_sign_viol"] == 0.0 and k["bound_viol"] == 0.0, (tag, k)
    assert abs(k["rel_primal"] - r["rel_primal"]) <= 1e-3 * tol, (tag, k, r["rel_primal"])
    assert abs(k["rel_gap"] - r["rel_gap"]) <= 1e-3 * tol, (tag, k, r["rel_gap"])
    assert k["rel_dual"] <= r["rel_dual"] * (1 + 1e-9) + 1e-3 * tol, (tag, k, r["rel_dual"])
    assert abs(k["primal_obj"] - r["objective"]) <= 1e-12 * (1 + abs(r["objective"])), (tag, k, r)
    return k


def one_sided_lp(seed=0x5D1, m=300, n=750):
    """C0-family LP with 10% free, 15% lower-only and 15% upper-only columns:
    the dual residual is not identically zero at approximate optima (it is on
    boxed LPs), so dual feasibility is certified on its own."""
    from mipx_amd import lpgen
    return lpgen.c0_plain(seed=seed, m=m, n=n, free=0.1, lower_only=0.15, upper_only=0.15).lp


def farkas_bounds(lp, lo, hi, row):
    """Column bounds that make equality row `row` infeasible in a way the
    reference's Farkas rule (src/lp/pdlp.cpp:119-211) certifies: the row's
    columns become one-sided (a > 0: [10, inf), a < 0: (-inf, 0]) so its
    minimum activity exceeds the right-hand side and A' e_row meets the sign
    rules.  (An infeasibility made by boxed columns needs A'ray = 0 on every
    boxed column, which the rule never finds.)"""
    import numpy as np
    lo = lo.copy()
    hi = hi.copy()
    s, e = lp.row_starts[row], lp.row_starts[row + 1]
    assert lp.row_lower[row] == lp.row_upper[row]
    amin = 0.0
    for j, v in zip(lp.col_indices[s:e], lp.values[s:e]):
        if v > 0:
            lo[j], hi[j] = 10.0, np.inf
            amin += v * 10.0
        else:
            lo[j], hi[j] = -np.inf, 0.0
    assert amin > lp.row_lower[row]
    return lo, hi
