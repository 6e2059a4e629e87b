"""CPU checks of the LP presolve (include/mipx_presolve.h, libmipx_host.so):
the reductions it must find on a planted LP, postsolve of primal and dual
solutions certified on the original LP, infeasibility detection, and LPs
left alone.  The reduced and original LPs are solved by the oracle."""
import numpy as np
import pytest

from conftest import TIGHT
from lp_cases import closed_form_cases, highs_objective, planted_presolve_lp
from mipx_amd import lpgen
from mipx_amd.lp import from_dense
from mipx_amd.presolve import Presolved
from oracle import binding as ob


def test_planted_reductions_and_postsolve_certified():
    lp, x0, want = planted_presolve_lp()
    pre = Presolved(lp)
    st = pre.stats
    print(st)
    assert st["fixed_var_changes"] >= want["fixed"]
    assert st["singleton_row_changes"] == want["singleton_rows"]
    assert st["empty_row_changes"] == want["empty_rows"]
    assert st["empty_col_changes"] == want["empty_cols"]
    assert st["infeasible"] == 0
    red = pre.reduced
    assert red.num_rows == lp.num_rows - want["singleton_rows"] - want["empty_rows"]
    assert red.num_cols <= lp.num_cols - want["fixed"] - want["empty_cols"]
    rr, xr, yr = ob.solve(red, **TIGHT)
    assert rr["status"] == 0
    x, y = pre.postsolve(xr, yr)
    # The reduced LP's objective (with the offset of fixed columns) is the
    # original objective at the postsolved point.
    assert abs((lp.obj @ x + lp.obj_offset) - rr["objective"]) <= 1e-9 * (1 + abs(rr["objective"]))
    # Against the exact optimum of the original LP: the reduced LP solved to
    # 1e-8 lands within 1e-6 of it.
    opt = highs_objective(lp)
    r8, _, _ = ob.solve(red, primal_tol=1e-8, dual_tol=1e-8, optimality_tol=1e-8, verbose=0,
                        max_iter=2000000)
    assert r8["status"] == 0
    assert abs(r8["objective"] - opt) <= 1e-6 * (1 + abs(opt)), (r8["objective"], opt)
    # Postsolved primal and dual certified on the original LP; the binding
    # singleton rows carry multipliers (a primal-only postsolve would leave
    # those columns with uncovered reduced costs).
    k = ob.kkt_check(lp, x, y)
    print(k)
    assert k["rel_primal"] <= 1.01e-6 and k["rel_gap"] <= 1.01e-6 and k["rel_dual"] <= 1.01e-6
    assert k["dual_sign_viol"] == 0.0 and k["bound_viol"] == 0.0
    singles = slice(lp.num_rows - want["singleton_rows"] - 2, lp.num_rows - 2)
    assert np.count_nonzero(y[singles]) >= 8
    # Dropping the singleton rows' multipliers breaks the certificate.
    yb = y.copy()
    yb[singles] = 0.0
    assert ob.kkt_check(lp, x, yb)["rel_gap"] > 1e-5


def test_fixed_columns_stall_the_reference_loop_without_presolve():
    """Why presolve matters for PDLP: the reference's dual-residual test
    checks a column's lower bound first (src/lp/pdlp.cpp:1127-1142), so a
    fixed column with a negative reduced cost never counts as dual feasible
    and the planted LP does not converge without presolve (the oracle
    follows the reference bit for bit); its reduced LP does."""
    lp, _, _ = planted_presolve_lp()
    r0, _, _ = ob.solve(lp, max_iter=40000, **TIGHT)
    rr, _, _ = ob.solve(Presolved(lp).reduced, max_iter=40000, **TIGHT)
    assert r0["status"] == 3 and rr["status"] == 0


def test_generated_lps_left_alone():
    """The seeded workload families have no fixed columns, singleton or
    empty rows: presolve keeps every row and column."""
    for lp in (lpgen.c0_plain(seed=101, m=300, n=750).lp,
               lpgen.c2_shared(seed=0xC2, m=300, n=800).lp):
        pre = Presolved(lp)
        assert pre.stats["vars_removed"] == pre.stats["rows_removed"] == 0
        red = pre.reduced
        assert np.array_equal(red.values, lp.values) and np.array_equal(red.row_starts, lp.row_starts)
        x = np.arange(lp.num_cols, dtype=float)
        assert np.array_equal(pre.postsolve(x)[0], x)


def test_infeasibility_found_in_presolve():
    inf = np.inf
    # empty row 0 in [1, 2]
    lp1 = from_dense(np.array([[0.0, 0.0], [1.0, 1.0]]), [1.0, 1.0], [0.0, 0.0], [5.0, 5.0],
                     [1.0, -inf], [2.0, 4.0])
    assert Presolved(lp1).stats["infeasible"] == 1
    # singleton row x0 >= 6 against x0 <= 5
    lp2 = from_dense(np.array([[1.0, 0.0], [1.0, 1.0]]), [1.0, 1.0], [0.0, 0.0], [5.0, 5.0],
                     [6.0, -inf], [inf, 4.0])
    assert Presolved(lp2).stats["infeasible"] == 1
    # fixed column makes a row impossible: x0 = 3, then x0 + x1 = 1 with x1 >= 0
    # leaves the singleton row x1 = -2 against x1 >= 0
    lp3 = from_dense(np.array([[1.0, 1.0]]), [1.0, 1.0], [3.0, 0.0], [3.0, 5.0], [1.0], [1.0])
    assert Presolved(lp3).stats["infeasible"] == 1


def test_closed_forms_through_presolve():
    """The reference's closed-form LPs solve to the same objective with and
    without presolve (oracle on both); an empty column with an infinite
    preferred bound stays in the LP for the solver to report unboundedness."""
    for c, lp in closed_form_cases():
        if lp.num_cols == 0 or c.get("options"):
            continue
        pre = Presolved(lp)
        if pre.stats["infeasible"]:
            assert c["status"] == "Infeasible", c["name"]
            continue
        ro, _, _ = ob.solve(lp, verbose=0, max_iter=100000)
        rr, xr, yr = ob.solve(pre.reduced, verbose=0, max_iter=100000)
        print(c["name"], pre.stats, ro["status"], rr["status"])
        assert rr["status"] == ro["status"], c["name"]
        if ro["status"] == 0:
            x, _ = pre.postsolve(xr, yr)
            assert abs(lp.obj @ x + lp.obj_offset - ro["objective"]) <= 1e-4, c["name"]


def _random_feasible_lp(rng):
    """A small LP that x0 satisfies, with the structures presolve removes:
    fixed columns, empty and singleton rows, empty columns; every other
    column boxed, so the LP is bounded."""
    m, n = int(rng.integers(2, 9)), int(rng.integers(2, 11))
    a = np.zeros((m, n))
    for i in range(m):
        kind = rng.integers(0, 4)
        if kind == 1:
            a[i, rng.integers(0, n)] = rng.uniform(0.5, 3.0) * rng.choice([-1.0, 1.0])
        elif kind >= 2:
            mask = rng.random(n) < 0.5
            a[i, mask] = rng.uniform(-3.0, 3.0, mask.sum())
    lo = -rng.uniform(0.0, 3.0, n)
    hi = rng.uniform(0.0, 3.0, n)
    fixed = rng.random(n) < 0.25
    v = rng.uniform(-2.0, 2.0, n)
    lo[fixed] = hi[fixed] = v[fixed]
    x0 = np.clip(rng.uniform(-3.0, 3.0, n), lo, hi)
    ax = a @ x0
    rl, ru = ax - rng.uniform(0.0, 1.0, m), ax + rng.uniform(0.0, 1.0, m)
    kinds = rng.integers(0, 4, m)
    rl[kinds == 0] = ru[kinds == 0] = ax[kinds == 0]
    rl[kinds == 1] = -np.inf
    ru[kinds == 2] = np.inf
    return from_dense(a, rng.uniform(-2.0, 2.0, n), lo, hi, rl, ru)


def test_random_lps_presolve_sound():
    """Random small LPs that are feasible by construction: presolve never
    declares one infeasible, and the oracle's solution of the reduced LP,
    postsolved, has the original LP's exact optimal objective (scipy HiGHS)
    and passes the long-double KKT check on the original LP (primal and
    dual).  The original LPs themselves often stall in the oracle, as in the
    reference, when a fixed column has a negative reduced cost."""
    rng = np.random.default_rng(0x9E5)
    removed = 0
    for case in range(60):
        lp = _random_feasible_lp(rng)
        pre = Presolved(lp)
        assert pre.stats["infeasible"] == 0, case
        removed += pre.stats["vars_removed"] + pre.stats["rows_removed"]
        rr, xr, yr = ob.solve(pre.reduced, max_iter=500000, **TIGHT)
        assert rr["status"] == 0, (case, rr)
        x, y = pre.postsolve(xr, yr)
        obj = lp.obj @ x + lp.obj_offset
        assert abs(obj - rr["objective"]) <= 1e-9 * (1 + abs(rr["objective"])), case
        opt = highs_objective(lp)
        assert abs(obj - opt) <= 1e-5 * (1 + abs(opt)), (case, obj, opt)
        k = ob.kkt_check(lp, x, y)
        assert k["rel_primal"] <= 1e-5 and k["rel_dual"] <= 1e-5 and k["rel_gap"] <= 1e-5, (case, k)
        assert k["bound_viol"] == 0.0, (case, k)
    assert removed > 100
