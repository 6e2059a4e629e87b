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


def planted_presolve_lp(seed=0x9E5, m=300, n=750):
    """A C0-family LP with structure presolve removes: 20 fixed columns
    (l = u at the generator's feasible point), singleton rows that bind at
    the optimum (x_j <= 9 or 2 x_j >= 1 on columns the LP pushes to a bound,
    an equality 3 x_j = 3 x0_j), two empty rows (one ranged around 0, one an
    equality 0 = 0) and three empty columns (cost +, -, 0).  Returns the LP,
    the generator's point x0, and the expected counts."""
    import numpy as np
    import scipy.sparse as sp
    from mipx_amd import lpgen
    from mipx_amd.lp import LpProblem
    from oracle import binding as ob
    g = lpgen.c0_plain(seed=seed, m=m, n=n)
    lp, x0 = g.lp, g.x0
    r, x, _ = ob.solve(lp, primal_tol=1e-6, dual_tol=1e-6, optimality_tol=1e-6, verbose=0)
    rng = np.random.default_rng(seed)
    lo, hi = lp.col_lower.copy(), lp.col_upper.copy()
    fixed = rng.choice(n, 20, replace=False)
    lo[fixed] = hi[fixed] = x0[fixed]
    rest = np.setdiff1d(np.arange(n), fixed)
    at_hi = [j for j in rest if x[j] > 9.9][:8]
    at_lo = [j for j in rest if x[j] < 0.1][:8]
    mid = [j for j in rest if 1.0 < x[j] < 9.0][:1]
    rows = []   # (col, coeff, rl, ru)
    rows += [(j, 1.0, -np.inf, 9.0) for j in at_hi]
    rows += [(j, 2.0, 1.0, np.inf) for j in at_lo]
    rows += [(j, 3.0, 3.0 * x0[j], 3.0 * x0[j]) for j in mid]
    a = sp.csr_matrix((lp.values, lp.col_indices, lp.row_starts), shape=(m, n))
    extra = sp.csr_matrix(([c for _, c, _, _ in rows], ([k for k in range(len(rows))],
                                                        [j for j, _, _, _ in rows])),
                          shape=(len(rows), n))
    empty_rows = sp.csr_matrix((2, n))
    a2 = sp.vstack([a, extra, empty_rows]).tocsr()
    a2 = sp.hstack([a2, sp.csr_matrix((a2.shape[0], 3))]).tocsr()
    a2.sort_indices()
    rl = np.concatenate([lp.row_lower, [r_[2] for r_ in rows], [-1.0, 0.0]])
    ru = np.concatenate([lp.row_upper, [r_[3] for r_ in rows], [2.0, 0.0]])
    obj = np.concatenate([lp.obj, [1.5, -2.0, 0.0]])
    lo = np.concatenate([lo, [-1.0, 0.0, -3.0]])
    hi = np.concatenate([hi, [4.0, 5.0, 3.0]])
    out = LpProblem(a2.shape[0], a2.shape[1], a2.indptr, a2.indices, a2.data, obj, lo, hi, rl, ru,
                    0, 0.0, name="planted")
    counts = dict(fixed=20, singleton_rows=len(rows), empty_rows=2, empty_cols=3)
    return out, np.concatenate([x0, [0.0, 0.0, 0.0]]), counts


def write_free_mps(lp, path):
    """Free-format MPS of an LpProblem (N row COST, then E / L / G / ranged
    rows, RHS, RANGES, BOUNDS), enough for the CLI's reader."""
    import numpy as np
    lines = [f"NAME {lp.name or 'LP'}", "ROWS", " N COST"]
    kinds = []
    for i in range(lp.num_rows):
        l, u = lp.row_lower[i], lp.row_upper[i]
        k = "E" if l == u else ("L" if np.isinf(l) else ("G" if np.isinf(u) else "R"))
        kinds.append(k)
        lines.append(f" {'L' if k == 'R' else k} R{i}")
    lines.append("COLUMNS")
    cols = [[] for _ in range(lp.num_cols)]
    for i in range(lp.num_rows):
        for q in range(lp.row_starts[i], lp.row_starts[i + 1]):
            cols[lp.col_indices[q]].append((i, lp.values[q]))
    for j in range(lp.num_cols):
        lines.append(f" C{j} COST {float(lp.obj[j])!r}")
        for i, v in cols[j]:
            lines.append(f" C{j} R{i} {float(v)!r}")
    lines.append("RHS")
    for i, k in enumerate(kinds):
        rhs = lp.row_lower[i] if k == "G" else lp.row_upper[i]
        if rhs != 0.0:
            lines.append(f" RHS R{i} {float(rhs)!r}")
    if "R" in kinds:
        lines.append("RANGES")
        for i, k in enumerate(kinds):
            if k == "R":
                lines.append(f" RNG R{i} {float(lp.row_upper[i] - lp.row_lower[i])!r}")
    lines.append("BOUNDS")
    for j in range(lp.num_cols):
        l, u = lp.col_lower[j], lp.col_upper[j]
        if l == u:
            lines.append(f" FX BND C{j} {float(l)!r}")
            continue
        if np.isinf(l) and np.isinf(u):
            lines.append(f" FR BND C{j}")
            continue
        if np.isinf(l):
            lines.append(f" MI BND C{j}")
        elif l != 0.0:
            lines.append(f" LO BND C{j} {float(l)!r}")
        if not np.isinf(u):
            lines.append(f" UP BND C{j} {float(u)!r}")
    lines.append("ENDATA")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def work_unit_workload():
    """The seeded solves of the device work-units gate
    (tests/test_gpu_work_units_gate.py, baseline written by
    tests/golden/make_device_work_units.py): name -> (kind, payload).
    kind 'single' (one LP), 'batch' (independent LPs), 'shared' (an LP and
    bounds variants), 'server' (fresh LPs and variants of a base)."""
    from mipx_amd import lpgen
    singles = {f"c0_{seed}_{m}x{n}": lpgen.c0_plain(seed=seed, m=m, n=n).lp
               for seed, m, n in ((101, 300, 750), (102, 500, 1250), (103, 200, 500),
                                  (0xC0, 2000, 5000))}
    batch = [lpgen.c3_instance(i, m).lp for i, m in enumerate((200, 600, 1500, 3000))]
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    count = 64
    lo = np.empty((count, 800))
    hi = np.empty((count, 800))
    for i in range(count):
        lo[i], hi[i] = g.branch_bounds(lpgen.c2_instance_seed(i))
    fresh = [lpgen.c3_instance(40 + i, m).lp for i, m in enumerate((400, 2500))]
    return {"singles": singles, "batch": batch, "shared": (g.lp, lo, hi),
            "server": (fresh, g.lp, lo[:8], hi[:8])}


def run_work_unit_workload(native, tight):
    """Solves the gate's workload on the device; returns name -> dict(status,
    iterations, work_units)."""
    w = work_unit_workload()
    out = {}

    def rec(name, r):
        out[name] = dict(status=int(r["status"]), iterations=int(r["iterations"]),
                         work_units=float(r["work_units"]))

    for prec, tag in ((0, "fp64"), (1, "mixed")):
        for name, lp in w["singles"].items():
            rec(f"single/{tag}/{name}", native.solve(lp, precision=prec, **tight)[0])
        res, _, _ = native.solve_batch(w["batch"], precision=prec, **tight)
        for i, r in enumerate(res):
            rec(f"batch/{tag}/{i}", r)
        lp, lo, hi = w["shared"]
        res, _, _ = native.solve_batch_shared(lp, lo, hi, precision=prec, **tight)
        for i, r in enumerate(res):
            rec(f"shared/{tag}/{i}", r)
    fresh, base_lp, lo, hi = w["server"]
    srv = native.Server(hbm_budget_bytes=4e9, max_slots=64, precision=1, **tight)
    base = srv.add_base(base_lp)
    tickets = [(f"server/mixed/fresh{i}", srv.submit(lp)) for i, lp in enumerate(fresh)]
    tickets += [(f"server/mixed/variant{i}", srv.submit_bounds(base, lo[i], hi[i]))
                for i in range(lo.shape[0])]
    got = {}
    while len(got) < len(tickets):
        srv.step()
        for t in srv.poll():
            got[t] = srv.result(t)
    srv.close()
    for name, t in tickets:
        rec(name, got[t][0])
    return out
