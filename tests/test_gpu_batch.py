"""GPU parity tests of the batched path (one workgroup per LP) against the
oracle, through the C-ABI."""
import numpy as np
import pytest

from conftest import TIGHT
from lp_cases import certify, farkas_bounds, highs_objective
from mipx_amd import lpgen, native
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def _check_vs_oracle(lp, r, x, y, tag="", mixed=False, same_order=False):
    """FP64: objective within 1e-6 of the oracle's (same trajectory); with
    `same_order` (every column and row summed by one lane in storage order:
    the lane-per-instance kernels, or packed tiles with one lane per column)
    the iteration count is the oracle's to within one evaluation block and
    the restart count to within one (the power iteration's last bits are
    device reductions; measured equal on the C2 family).
    Mixed precision (FP32 iterations, FP64 evaluation and refinement) lands
    on a different point of the 1e-6 KKT region.  A 1e-6 relative KKT
    tolerance does not bound the objective error by 1e-6: the oracle itself
    (bit-identical to the reference) is up to 3.9e-6 from the exact optimum
    on these LP families.  Mixed results at tolerance 1e-6 are therefore held
    to 3e-6 of the exact optimum (scipy HiGHS) and KKT-certified by the
    independent long-double checker; the 1e-6 bar against the oracle holds
    when both solve to 1e-8 (test_mixed_refined_to_1e8_matches_oracle)."""
    ro, xo, yo = ob.solve(lp, **TIGHT)
    assert r["status"] == ro["status"], (tag, r, ro["status"])
    if ro["status"] == 0:
        if mixed:
            opt = highs_objective(lp)
            assert abs(r["objective"] - opt) <= 3e-6 * (1 + abs(opt)), (tag, r["objective"], opt)
        else:
            rel = abs(r["objective"] - ro["objective"]) / (1 + abs(ro["objective"]))
            assert rel <= 1e-6, (tag, r["objective"], ro["objective"], rel)
            if same_order:
                assert abs(r["iterations"] - ro["iterations"]) <= 200, (tag, r, ro["iterations"])
                assert abs(r["restarts"] - ro["restarts"]) <= 1, (tag, r, ro["restarts"])
        certify(lp, r, x, y, tag=tag)
    return ro


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_shared_bounds_vs_oracle(precision):
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    count = 24
    lo = np.empty((count, 800))
    hi = np.empty((count, 800))
    for i in range(count):
        lo[i], hi[i] = g.branch_bounds(lpgen.c2_instance_seed(i))
    res, X, Y = native.solve_batch_shared(g.lp, lo, hi, precision=precision, max_iter=200000,
                                          **TIGHT)
    statuses = [r["status"] for r in res]
    print("statuses", statuses, "iters", [r["iterations"] for r in res])
    for i in range(count):
        _check_vs_oracle(g.lp.with_col_bounds(lo[i], hi[i]), res[i], X[i], Y[i], f"inst{i}",
                         mixed=precision == 1, same_order=True)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_shared_infeasible_and_one_sided_vs_oracle(precision):
    """The lane-per-instance path on a base with free and one-sided columns,
    with bounds variants of which three are made infeasible (farkas_bounds):
    those return Infeasible from the device certificate (bCertificates) at
    the same evaluation as the oracle, with the Farkas ray on the violated
    row; the rest match the oracle and are certified."""
    from mipx_amd.lpgen import Generated, params
    g = Generated(params(300, 800, 0, 6, 6, value_kind=2, vlo=0.0, vhi=5.0, seed=0xC25,
                         free=0.05, lower_only=0.1, upper_only=0.1))
    count = 16
    lo = np.empty((count, 800))
    hi = np.empty((count, 800))
    bad = {3: 17, 8: 101, 13: 3}
    for i in range(count):
        lo[i], hi[i] = g.branch_bounds(lpgen.c2_instance_seed(i))
        if i in bad:
            lo[i], hi[i] = farkas_bounds(g.lp, lo[i], hi[i], bad[i])
    res, X, Y = native.solve_batch_shared(g.lp, lo, hi, precision=precision, max_iter=200000,
                                          **TIGHT)
    print("statuses", [r["status"] for r in res], "iters", [r["iterations"] for r in res])
    for i in range(count):
        lp = g.lp.with_col_bounds(lo[i], hi[i])
        ro = _check_vs_oracle(lp, res[i], X[i], Y[i], f"inst{i}", mixed=precision == 1,
                              same_order=True)
        if i in bad:
            assert res[i]["status"] == ro["status"] == 1 and ro["iterations"] == 2000
            # FP32 iterates pass the same test at the same check (FP64) or
            # at a later one (mixed precision).
            assert (res[i]["iterations"] == 2000 if precision == 0
                    else res[i]["iterations"] % 2000 == 0)
            ray = Y[i] / np.linalg.norm(Y[i])
            assert abs(abs(ray[bad[i]]) - 1.0) <= 1e-3, i
        else:
            assert res[i]["status"] == 0, i


@pytest.mark.parametrize("group", [0, 1])
@pytest.mark.parametrize("precision", [0, 1])
def test_batch_heterogeneous_vs_oracle(precision, group):
    """Independent LPs in packed tiles: each LP's own (autotuned) lane
    groups, or one lane per column and row (spmv_group=1), whose FP64 runs
    follow the oracle's trajectory."""
    sizes = [40, 75, 120, 200, 333, 60, 90, 150]
    gens = [lpgen.c3_instance(i, m) for i, m in enumerate(sizes)]
    lps = [g.lp for g in gens]
    res, xs, ys = native.solve_batch(lps, precision=precision, max_iter=200000,
                                     spmv_group=group, **TIGHT)
    for i, lp in enumerate(lps):
        _check_vs_oracle(lp, res[i], xs[i], ys[i], f"lp{i}", mixed=precision == 1,
                         same_order=group == 1)


def test_batch_result_same_alone_and_in_mix():
    """An instance's result does not depend on what else is in the launch."""
    sizes = [80, 150, 60, 220]
    lps = [lpgen.c3_instance(i, m).lp for i, m in enumerate(sizes)]
    mix, xm, ym = native.solve_batch(lps, precision=1, max_iter=200000, **TIGHT)
    for i in (1, 3):
        alone, xa, ya = native.solve_batch([lps[i]], precision=1, max_iter=200000, **TIGHT)
        assert alone[0]["iterations"] == mix[i]["iterations"]
        assert alone[0]["objective"] == mix[i]["objective"]
        assert np.array_equal(xa[0], xm[i]) and np.array_equal(ya[0], ym[i])


def test_batch_detects_infeasible_and_unbounded():
    from lp_cases import closed_form_cases
    cases = {c["name"]: lp for c, lp in closed_form_cases()}
    lps = [cases["infeasible"], cases["unbounded_constrained"], cases["simple"]]
    res, xs, ys = native.solve_batch(lps, verbose=0, max_iter=100000)
    assert [r["status"] for r in res] == [1, 2, 0]
    assert abs(res[2]["objective"] + 4.0) <= 1e-4


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_certificates_match_oracle(precision):
    """The infeasibility and unboundedness certificates (tests/test_pdlp.cpp:
    347-496, src/lp/pdlp.cpp:119-284) on the batched paths, whose evaluation
    parts gather the certificates' terms and whose control kernel compares
    them: every closed-form LP with rows and columns in one packed batch of
    independent LPs, and bounds variants of one matrix on the lane-per-
    instance path (x1 - x2 <= 0, min -x1: as given unbounded, x2 <= 3 and
    x1 <= 2 bounded, x1 >= 5 with x2 <= 3 infeasible through the bounds
    alone, which the reference's Farkas test does not certify: it stops at
    the iteration limit, on the device as in the oracle), each with the
    oracle's status and, where optimal, its objective."""
    from lp_cases import closed_form_cases
    cases = [(c, lp) for c, lp in closed_form_cases()
             if c["status"] in ("Optimal", "Infeasible", "Unbounded") and c["m"] > 0 and c["n"] > 0]
    lps = [lp for _, lp in cases]
    res, xs, ys = native.solve_batch(lps, precision=precision, max_iter=100000)
    for (c, lp), r in zip(cases, res):
        ro, _, _ = ob.solve(lp, max_iter=100000, verbose=0)
        assert r["status"] == ro["status"], (c["name"], r, ro)
        if ro["status"] == 0:
            assert abs(r["objective"] - c["objective"]) <= 1e-4 * (1 + abs(c["objective"])), (c, r)
    base = {c["name"]: lp for c, lp in cases}["unbounded_constrained"]
    inf = np.inf
    lo = np.array([[0, 0], [0, 0], [5, 0], [0, 0]], dtype=float)
    hi = np.array([[inf, inf], [inf, 3], [inf, 3], [2, inf]], dtype=float)
    res, X, Y = native.solve_batch_shared(base, lo, hi, precision=precision, max_iter=20000)
    want = [(2, None), (0, -3.0), (3, None), (0, -2.0)]
    for i, (st, obj) in enumerate(want):
        one = base.with_col_bounds(lo[i], hi[i])
        ro, _, _ = ob.solve(one, max_iter=20000, verbose=0)
        assert res[i]["status"] == ro["status"] == st, (i, res[i], ro)
        if obj is not None:
            assert abs(res[i]["objective"] - obj) <= 1e-4, (i, res[i])


@pytest.mark.parametrize("precision", [0, 1])
def test_staged_evaluation_boundary(precision):
    """The evaluation copies an instance's y^, y0 and x^ into LDS when 2m
    and n are at most 8,192 (and the whole evaluation is one part): LPs
    exactly at that size and one row or column past it, in one packed
    batch and as bounds variants of a shared matrix, each Optimal and
    certified by the long-double checker, with the single-LP path's
    objective."""
    lps = [lpgen.c0_plain(seed=61, m=4096, n=8192, per_col=4).lp,   # staged, at the limit
           lpgen.c0_plain(seed=62, m=4097, n=8192, per_col=4).lp,   # 2m one past
           lpgen.c0_plain(seed=63, m=2000, n=8193, per_col=4).lp]   # n one past
    opts = dict(precision=precision, max_iter=300000, **TIGHT)
    res, xs, ys = native.solve_batch(lps, **opts)
    for i, (lp, r, x, y) in enumerate(zip(lps, res, xs, ys)):
        assert r["status"] == 0, (i, r)
        certify(lp, r, x, y, tag=f"packed{i}")
        one, _, _ = native.solve(lp, **opts)
        assert abs(one["objective"] - r["objective"]) <= 5e-6 * (1 + abs(one["objective"])), (i, r, one)
    g = lpgen.c2_shared(seed=0xC6, m=4096, n=8192)
    lo = np.empty((6, 8192))
    hi = np.empty((6, 8192))
    for i in range(6):
        lo[i], hi[i] = g.branch_bounds(lpgen.c2_instance_seed(i, base=0xC6))
    res, X, Y = native.solve_batch_shared(g.lp, lo, hi, **opts)
    for i in range(6):
        lp = g.lp.with_col_bounds(lo[i], hi[i])
        if res[i]["status"] == 0:
            certify(lp, res[i], X[i], Y[i], tag=f"shared{i}")
        else:
            ro, _, _ = ob.solve(lp, max_iter=300000, **TIGHT)
            assert res[i]["status"] == ro["status"], (i, res[i], ro)


@pytest.mark.parametrize("which", ["c1_small", "c2_small"])
def test_batched_setup_vs_oracle(which):
    """The batched engines' cohort setup: element-wise scaling bit-identical
    to the reference's (oracle), global scales to 1e-14, the power
    iteration's spectral norm to its tolerance, and repeatable."""
    lp = (lpgen.c1_large(seed=3, m=3000, n=6000) if which == "c1_small"
          else lpgen.c2_shared(seed=0xC2, m=300, n=800)).lp
    got = native.scale_problem(lp, batched=True)
    again = native.scale_problem(lp, batched=True)
    want = ob.scale_problem(lp)
    for k in ("values", "at_values", "at_row_starts", "at_col_indices", "row_scale",
              "col_scale", "sigma_base", "tau_base"):
        assert np.array_equal(got[k], want[k]), k
    for k in ("objective_scale", "constraint_scale"):
        assert abs(got[k] - want[k]) <= 1e-14 * abs(want[k]), (k, got[k], want[k])
    for k in ("cscaled", "col_lower", "col_upper", "row_lower", "row_upper"):
        np.testing.assert_allclose(got[k], want[k], rtol=1e-14, atol=0)
    print("spectral norm", got["spectral_norm"], want["spectral_norm"], "iters",
          got["power_iterations"], want["power_iterations"])
    assert abs(got["spectral_norm"] - want["spectral_norm"]) <= 1e-6 * want["spectral_norm"]
    for k, v in got.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(v, again[k]), k
        else:
            assert v == again[k], k


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_shared_warm_start_vs_oracle(precision):
    """Branch-and-bound feed: node LPs (bounds variants of one base)
    warm-started from the root LP's solution.  Each instance matches the
    oracle's warm-started solve of its own LP (setColBounds + setWarmStart,
    pdlp.cpp:2890-2898, 915-930); in FP64 the iteration count follows the
    oracle's.  Restarting the instances from their own solutions takes fewer
    iterations in total than the cold solves (the reference's warm-start test,
    tests/test_pdlp.cpp:1076-1113, for one LP; per instance the reference
    itself sometimes needs more)."""
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    count = 12
    lo = np.empty((count, 800))
    hi = np.empty((count, 800))
    for i in range(count):
        lo[i], hi[i] = g.branch_bounds(lpgen.c2_instance_seed(i))
    r0, x0, y0 = native.solve(g.lp, **TIGHT)
    assert r0["status"] == 0
    opts = dict(precision=precision, max_iter=200000, **TIGHT)
    warm = (np.tile(x0, (count, 1)), np.tile(y0, (count, 1)))
    res, X, Y = native.solve_batch_shared(g.lp, lo, hi, warm=warm, **opts)
    cold, _, _ = native.solve_batch_shared(g.lp, lo, hi, **opts)
    print("warm iters", [r["iterations"] for r in res], "cold", [r["iterations"] for r in cold])
    both = []
    for i in range(count):
        lp = g.lp.with_col_bounds(lo[i], hi[i])
        ro, _, _ = ob.solve(lp, warm=(x0, y0), **TIGHT)
        assert res[i]["status"] == ro["status"] == cold[i]["status"], (i, res[i], ro["status"])
        if ro["status"] == 0:
            rel = abs(res[i]["objective"] - ro["objective"]) / (1 + abs(ro["objective"]))
            assert rel <= (5e-6 if precision else 1e-6), (i, res[i]["objective"], ro["objective"])
            if not precision:
                assert abs(res[i]["iterations"] - ro["iterations"]) <= 200, (i, res[i], ro)
            certify(lp, res[i], X[i], Y[i], tag=f"warm{i}")
            both.append(i)
    assert both
    _, XC, YC = native.solve_batch_shared(g.lp, lo, hi, **opts)
    again, _, _ = native.solve_batch_shared(g.lp, lo, hi, warm=(XC, YC), **opts)
    print("re-warm iters", [again[i]["iterations"] for i in both])
    assert all(again[i]["status"] == 0 for i in both)
    assert sum(again[i]["iterations"] for i in both) < sum(cold[i]["iterations"] for i in both)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_preconditioner_refresh(precision):
    """The batched path applies the same refresh rule: forced refreshes
    (ratio 1.01) leave every instance's run bit-identical."""
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    count = 8
    lo = np.empty((count, 800))
    hi = np.empty((count, 800))
    for i in range(count):
        lo[i], hi[i] = g.branch_bounds(lpgen.c2_instance_seed(i))
    base = dict(TIGHT, precision=precision, max_iter=200000, primal_weight=1.0)
    r0, X0, Y0 = native.solve_batch_shared(g.lp, lo, hi, **base)
    r1, X1, Y1 = native.solve_batch_shared(g.lp, lo, hi, preconditioner_refresh=1,
                                           preconditioner_refresh_ratio=1.01, **base)
    for a, b in zip(r0, r1):
        assert a["status"] == b["status"] == 0
        assert (a["iterations"], a["objective"]) == (b["iterations"], b["objective"])
    assert np.array_equal(X0, X1) and np.array_equal(Y0, Y1)
    assert any(abs(r["primal_weight"] - 1.0) > 0.01 for r in r1)


REFINED = dict(primal_tol=1e-8, dual_tol=1e-8, optimality_tol=1e-8, verbose=0)


def test_mixed_refined_to_1e8_matches_oracle():
    """Mixed precision refined to 1e-8 (FP32 iterations, FP64 phase to
    refine_tol, the final refinement to 1e-8 of the large-LP workload) against the
    oracle solved to 1e-8: objective within 1e-6 of the oracle's and of the
    exact optimum, the returned point certified at 1e-8, on the single-LP
    path, a shared-matrix batch (lane-per-instance kernels) and a packed
    batch."""
    opts = dict(TIGHT, precision=1, refine_tol=1e-8, max_iter=2000000)
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    count = 6
    lo = np.empty((count, 800))
    hi = np.empty((count, 800))
    for i in range(count):
        lo[i], hi[i] = g.branch_bounds(lpgen.c2_instance_seed(i))
    res, X, Y = native.solve_batch_shared(g.lp, lo, hi, **opts)
    cases = [(g.lp.with_col_bounds(lo[i], hi[i]), res[i], X[i], Y[i]) for i in range(count)]
    lps = [lpgen.c3_instance(i, m).lp for i, m in enumerate([60, 150, 240])]
    pr, xs, ys = native.solve_batch(lps, **opts)
    cases += list(zip(lps, pr, xs, ys))
    one = lpgen.c0_plain(seed=101, m=300, n=750).lp
    cases.append((one,) + native.solve(one, **opts))
    for k, (lp, r, x, y) in enumerate(cases):
        ro, _, _ = ob.solve(lp, max_iter=2000000, **REFINED)
        opt = highs_objective(lp)
        print(k, r["objective"], ro["objective"], opt, r["iterations"], r["iterations_fp32"])
        assert r["status"] == ro["status"] == 0, (k, r)
        assert r["iterations_fp32"] > 0
        assert abs(r["objective"] - ro["objective"]) <= 1e-6 * (1 + abs(ro["objective"])), k
        assert abs(r["objective"] - opt) <= 1e-6 * (1 + abs(opt)), k
        certify(lp, r, x, y, tol=1e-8, tag=f"case{k}")


def test_packed_preset_from_autotuned_table():
    """Packed LPs take their lanes per column / row from the autotuned table
    (mipx_amd/csrc/packed_presets.h) by shape alone: the default solve is
    bit-identical to forcing that table entry, alone and in a mix, and
    differs from another preset in the last bits only (same status, the
    objective to 1e-6)."""
    from mipx_amd.autotune import training_batch
    lp34 = lpgen.c3_instance(5, 300).lp         # bucket (3, 4): preset 4
    lp35 = training_batch(3, 5, 1)[0]            # bucket (3, 5): preset 6
    for lp, tuned in ((lp34, 4), (lp35, 6)):
        d, xd, yd = native.solve_batch([lp], precision=1, max_iter=200000, **TIGHT)
        f, xf, yf = native.solve_batch([lp], precision=1, max_iter=200000, packed_preset=tuned,
                                       **TIGHT)
        o, _, _ = native.solve_batch([lp], precision=1, max_iter=200000, packed_preset=0, **TIGHT)
        assert d[0]["status"] == f[0]["status"] == o[0]["status"] == 0
        assert (d[0]["iterations"], d[0]["objective"]) == (f[0]["iterations"], f[0]["objective"])
        assert np.array_equal(xd[0], xf[0]) and np.array_equal(yd[0], yf[0])
        assert abs(o[0]["objective"] - d[0]["objective"]) <= 1e-5 * (1 + abs(d[0]["objective"]))
    mix, xm, _ = native.solve_batch([lp35, lp34], precision=1, max_iter=200000, **TIGHT)
    alone, xa, _ = native.solve_batch([lp34], precision=1, max_iter=200000, **TIGHT)
    assert mix[1]["objective"] == alone[0]["objective"] and np.array_equal(xm[1], xa[0])
