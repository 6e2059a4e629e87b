"""The option variants of the reference's own PDLP tests (tests/test_pdlp.cpp)
on the device, each against the oracle's solve with the same options:
polishing off, on with tight polish tolerances, and after a loose main
tolerance (1366-1500); bound / objective rescaling and the two scalings
switched off (572-737); the L-inf optimality norm (156-171).  One lane per
column and row (spmv_group = 1 on the single-LP path, one lane per column in
packed tiles, one lane per instance for bounds variants) sums every product
in the reference's order, so the device follows the oracle's trajectory:
iterations within one evaluation block of 200, restarts within one.

Also the reference's "polishing respects time limit" (1470-1493): a solve
that has converged and is polishing when the clock runs out ends Optimal at
its polished-so-far point, on the single-LP path, a batch and the server."""
import time

import numpy as np
import pytest

from conftest import TIGHT
from lp_cases import certify, closed_form_cases, medium_lp
from mipx_amd import lpgen, native
from oracle import binding as ob

pytestmark = pytest.mark.gpu

OPTIONS = {
    "no_polish": dict(enable_polishing=0),
    "tight_polish": dict(enable_polishing=1, polish_primal_tol=1e-7, polish_dual_tol=1e-7),
    "loose_main_then_polish": dict(primal_tol=1e-4, dual_tol=1e-4, optimality_tol=1e-4,
                                   enable_polishing=1, polish_primal_tol=1e-7,
                                   polish_dual_tol=1e-7),
    "no_rescaling": dict(do_bound_obj_rescaling=0),
    "no_ruiz": dict(do_ruiz_scaling=0),
    "no_pock_chambolle": dict(do_pock_chambolle_scaling=0),
    "no_scaling": dict(do_ruiz_scaling=0, do_pock_chambolle_scaling=0, do_bound_obj_rescaling=0),
    "linf": dict(optimality_norm=1),
    "fixed_primal_weight": dict(update_primal_weight=0, primal_weight=2.0),
    "eval_every_50": dict(termination_eval_frequency=50),
    "restart_rules": dict(restart_sufficient_decay=0.1, restart_necessary_decay=0.5,
                          restart_artificial_fraction=0.5),
    "pid_gains": dict(pid_kp=0.5, pid_ki=0.05, pid_kd=0.1, pid_i_smooth=0.5),
    "short_polish_budget": dict(enable_polishing=1, polish_primal_tol=1e-12,
                                polish_dual_tol=1e-12, polish_max_iter=400),
    "iteration_limit": dict(max_iter=1000),
}


def _closed(name):
    return {c["name"]: (c, lp) for c, lp in closed_form_cases()}[name]


def _lps():
    return {"simple": _closed("simple")[1], "equality": _closed("equality")[1],
            "medium": medium_lp(), "c0_300x750": lpgen.c0_plain(seed=101, m=300, n=750).lp}


def _opts(name):
    o = dict(TIGHT, max_iter=500000)
    o.update(OPTIONS[name])
    return o


def _check(lp, r, x, y, ro, o, tag):
    assert r["status"] == ro["status"], (tag, r, ro)
    block = o.get("termination_eval_frequency", 200)
    assert abs(r["iterations"] - ro["iterations"]) <= block, (tag, r["iterations"], ro["iterations"])
    assert abs(r["restarts"] - ro["restarts"]) <= 1, (tag, r["restarts"], ro["restarts"])
    assert r["work_units"] == r["iterations"] * (4.0 * lp.nnz + 2.0 * (lp.num_rows + lp.num_cols))
    if ro["status"] != 0:
        # Stopped by the iteration limit: no answer (pdlp.cpp:1705-1708).
        assert r["objective"] == 0.0 and not np.any(x) and not np.any(y), (tag, r)
        return
    assert abs(r["objective"] - ro["objective"]) <= 1e-6 * (1 + abs(ro["objective"])), (tag, r, ro)
    if o.get("optimality_norm", 0) == 1:
        return  # the certificate measures in L2
    tol = max(o["primal_tol"], o["optimality_tol"])
    if o.get("polish_max_iter", 500000) < 500000:
        # A polish phase that exhausts its budget keeps its last iterate,
        # which may sit just outside the main tolerance (pdlp.cpp:1665-1680:
        # "exhausted budget", then recovery from pdhg_x); the oracle's
        # trajectory, matched above, ends at the same point.
        tol = 10 * tol
    certify(lp, r, x, y, tol=tol, tag=tag)
    if o.get("enable_polishing", 1) and o.get("polish_max_iter", 500000) >= 500000:
        # Polish drives both residuals to its own tolerance (pdlp.cpp:1449).
        assert r["rel_primal"] <= o.get("polish_primal_tol", 1e-6) * 1.001, (tag, r)
        assert r["rel_dual"] <= o.get("polish_dual_tol", 1e-6) * 1.001, (tag, r)


@pytest.mark.parametrize("opt", sorted(OPTIONS))
@pytest.mark.parametrize("which", ["simple", "equality", "medium", "c0_300x750"])
def test_single_lp_options_vs_oracle(which, opt):
    lp = _lps()[which]
    o = _opts(opt)
    r, x, y = native.solve(lp, spmv_group=1, **o)
    ro, _, _ = ob.solve_gpu(lp, **o)
    _check(lp, r, x, y, ro, o, f"{which}/{opt}")


@pytest.mark.parametrize("opt", ["no_polish", "tight_polish", "loose_main_then_polish",
                                 "no_rescaling", "no_scaling", "linf", "fixed_primal_weight",
                                 "eval_every_50", "restart_rules", "pid_gains",
                                 "short_polish_budget", "iteration_limit"])
def test_batched_options_vs_oracle(opt):
    """The same option sets through a packed batch of independent LPs (one
    lane per column and row) and a batch of bounds variants of one matrix
    (one lane per instance), every instance against the oracle."""
    o = _opts(opt)
    lps = [lpgen.c3_instance(i, m).lp for i, m in enumerate([60, 150, 240])] + [medium_lp()]
    res, xs, ys = native.solve_batch(lps, precision=0, spmv_group=1, **o)
    for i, (lp, r, x, y) in enumerate(zip(lps, res, xs, ys)):
        ro, _, _ = ob.solve(lp, **o)
        _check(lp, r, x, y, ro, o, f"packed{i}/{opt}")
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    count = 6
    lo = np.empty((count, 800))
    hi = np.empty((count, 800))
    for i in range(count):
        lo[i], hi[i] = g.branch_bounds(lpgen.c2_instance_seed(i))
    res, X, Y = native.solve_batch_shared(g.lp, lo, hi, precision=0, **o)
    for i in range(count):
        one = g.lp.with_col_bounds(lo[i], hi[i])
        ro, _, _ = ob.solve(one, **o)
        _check(one, res[i], X[i], Y[i], ro, o, f"shared{i}/{opt}")


POLISH_FOREVER = dict(verbose=0, max_iter=500000, enable_polishing=1, polish_primal_tol=1e-15,
                      polish_dual_tol=1e-15, polish_max_iter=100000000, max_solve_seconds=3.0)


def test_polishing_respects_time_limit_single_batch_and_server():
    """tests/test_pdlp.cpp:1470-1493 on a 300 x 750 LP of the seeded family
    (its equality rows keep a nonzero residual, so polishing towards 1e-15
    never finishes; the reference's medium LP shape polishes to exactly zero
    residuals in half a second): the main loop converges, the 3 s limit cuts
    the polish, and the result is Optimal at the polished-so-far point (the
    oracle's objective, certified): one solve, a batch and the server."""
    lp = lpgen.c0_plain(seed=101, m=300, n=750).lp
    ro, _, _ = ob.solve(lp, max_iter=500000, **TIGHT)
    assert ro["status"] == 0

    def ok(r, x, y, elapsed, tag):
        assert elapsed >= 2.9, (tag, elapsed, r)  # stopped by the clock, inside polish
        assert r["status"] == 0, (tag, r)
        assert abs(r["objective"] - ro["objective"]) <= 1e-5 * (1 + abs(ro["objective"])), (tag, r)
        certify(lp, r, x, y, tag=tag)

    t0 = time.perf_counter()
    r, x, y = native.solve(lp, **POLISH_FOREVER)
    ok(r, x, y, time.perf_counter() - t0, "single")
    t0 = time.perf_counter()
    res, xs, ys = native.solve_batch([lp, lp], **POLISH_FOREVER)
    el = time.perf_counter() - t0
    for i in range(2):
        ok(res[i], xs[i], ys[i], el, f"batch{i}")
    srv = native.Server(hbm_budget_bytes=2e9, max_slots=8, **POLISH_FOREVER)
    tk = srv.submit(lp)
    t0 = time.perf_counter()
    got = None
    while got is None:
        srv.step()
        for t in srv.poll():
            if t == tk:
                got = srv.result(tk)
        assert time.perf_counter() - t0 < 60.0
    el = time.perf_counter() - t0
    srv.close()
    ok(got[0], got[1], got[2], el, "server")


MIXED_OPTIONS = ["no_polish", "tight_polish", "loose_main_then_polish", "no_rescaling",
                 "no_scaling", "linf", "fixed_primal_weight", "eval_every_50", "restart_rules",
                 "pid_gains"]


@pytest.mark.parametrize("opt", MIXED_OPTIONS)
def test_mixed_precision_options_certified(opt):
    """The option sets in mixed precision (FP32 iterations, FP64 evaluation
    and refinement) on the single-LP path, a packed batch and a batch of
    bounds variants: the oracle's status, the objective within the 5e-6
    envelope of two points of the 1e-6 KKT region, every point certified
    (L2 norms) at the tolerance."""
    o = _opts(opt)
    lps = [lpgen.c0_plain(seed=101, m=300, n=750).lp] + [
        lpgen.c3_instance(i, m).lp for i, m in enumerate([60, 150, 240])]
    runs = [(lps[0],) + native.solve(lps[0], precision=1, **o)]
    res, xs, ys = native.solve_batch(lps[1:], precision=1, **o)
    runs += list(zip(lps[1:], res, xs, ys))
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    lo = np.empty((4, 800))
    hi = np.empty((4, 800))
    for i in range(4):
        lo[i], hi[i] = g.branch_bounds(lpgen.c2_instance_seed(i))
    res, X, Y = native.solve_batch_shared(g.lp, lo, hi, precision=1, **o)
    runs += [(g.lp.with_col_bounds(lo[i], hi[i]), res[i], X[i], Y[i]) for i in range(4)]
    for k, (lp, r, x, y) in enumerate(runs):
        ro, _, _ = ob.solve(lp, **o)
        tag = f"run{k}/{opt}"
        assert r["status"] == ro["status"] == 0, (tag, r, ro)
        assert r["iterations_fp32"] > 0, (tag, r)
        assert abs(r["objective"] - ro["objective"]) <= 5e-6 * (1 + abs(ro["objective"])), (tag, r, ro)
        if o.get("optimality_norm", 0) != 1:
            certify(lp, r, x, y, tol=max(o["primal_tol"], o["optimality_tol"]), tag=tag)


def test_mixed_precision_handover_to_fp64():
    """mixed_refine = 0: the FP32 phase hands over to plain FP64 PDHG (the
    FP32 point becomes the FP64 iterate, restarted there) instead of the
    refinement, on the single-LP path, a packed batch and a batch of bounds
    variants: both phases run, the result has the oracle's status, is
    certified and lies within the 5e-6 objective envelope."""
    o = dict(TIGHT, max_iter=500000, precision=1, mixed_refine=0)
    lps = [lpgen.c0_plain(seed=101, m=300, n=750).lp] + [
        lpgen.c3_instance(i, m).lp for i, m in enumerate([60, 150, 240])]
    runs = [(lps[0],) + native.solve(lps[0], **o)]
    res, xs, ys = native.solve_batch(lps[1:], **o)
    runs += list(zip(lps[1:], res, xs, ys))
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    lo = np.empty((4, 800))
    hi = np.empty((4, 800))
    for i in range(4):
        lo[i], hi[i] = g.branch_bounds(lpgen.c2_instance_seed(i))
    res, X, Y = native.solve_batch_shared(g.lp, lo, hi, **o)
    runs += [(g.lp.with_col_bounds(lo[i], hi[i]), res[i], X[i], Y[i]) for i in range(4)]
    for k, (lp, r, x, y) in enumerate(runs):
        ro, _, _ = ob.solve(lp, **TIGHT)
        tag = f"run{k}"
        assert r["status"] == ro["status"] == 0, (tag, r, ro)
        assert 0 < r["iterations_fp32"] < r["iterations"], (tag, r)
        assert r["iterations_refine"] == 0, (tag, r)
        assert abs(r["objective"] - ro["objective"]) <= 5e-6 * (1 + abs(ro["objective"])), (tag, r, ro)
        certify(lp, r, x, y, tag=tag)


def _same(a, b, tag):
    """Bit-identical results, the wall-clock fields aside."""
    (ra, xa, ya), (rb, xb, yb) = a, b
    for k in ra:
        if "seconds" not in k:
            assert ra[k] == rb[k], (tag, k, ra[k], rb[k])
    assert np.array_equal(xa, xb) and np.array_equal(ya, yb), tag


@pytest.mark.parametrize("precision", [0, 1])
def test_identical_options_identical_runs(precision):
    """tests/test_pdlp.cpp:1310-1340: Anderson acceleration switched off
    explicitly and left at its default give the same iterations and
    objective.  The device holds this bit for bit, and for any repeat of a
    solve: every reduction runs in a fixed order, without atomics, and the
    lane-per-instance autotuner's choice of columns per wave leaves each
    lane's sums unchanged.  Single LP, packed batch, bounds variants (enough
    rounds for the autotuner to run) and the server, each solved twice."""
    base = dict(TIGHT, max_iter=500000, precision=precision)
    lp = lpgen.c0_plain(seed=101, m=300, n=750).lp
    one = native.solve(lp, **base)
    _same(one, native.solve(lp, anderson_acceleration=0, **base), "single/explicit-off")
    _same(one, native.solve(lp, **base), "single/repeat")
    lps = [lpgen.c3_instance(i, m).lp for i, m in enumerate([60, 150, 240])] + [medium_lp()]
    a = native.solve_batch(lps, **base)
    b = native.solve_batch(lps, anderson_acceleration=0, **base)
    for i in range(len(lps)):
        _same((a[0][i], a[1][i], a[2][i]), (b[0][i], b[1][i], b[2][i]), f"packed{i}")
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    count = 70
    lo = np.empty((count, 800))
    hi = np.empty((count, 800))
    for i in range(count):
        lo[i], hi[i] = g.branch_bounds(lpgen.c2_instance_seed(i))
    a = native.solve_batch_shared(g.lp, lo, hi, **base)
    b = native.solve_batch_shared(g.lp, lo, hi, **base)
    for i in range(count):
        _same((a[0][i], a[1][i], a[2][i]), (b[0][i], b[1][i], b[2][i]), f"shared{i}")
    srv_res = []
    for _ in range(2):
        srv = native.Server(hbm_budget_bytes=2e9, max_slots=8, **base)
        tk = srv.submit(lp)
        got = None
        while got is None:
            srv.step()
            for t in srv.poll():
                if t == tk:
                    got = srv.result(tk)
        srv.close()
        srv_res.append(got)
    _same(tuple(srv_res[0]), tuple(srv_res[1]), "server/repeat")
