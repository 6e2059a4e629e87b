"""CPU checks of the oracle (oracle/pdlp_oracle.cpp) and the generator.

Pins for the oracle: the reference's closed-form LPs, the SpMV known answers,
dense brute force, the adjoint identity, scaling invariants, an SVD check of
the power iteration, scipy-HiGHS optima of seeded LPs, and the reference's own
CPU PdlpSolver built from its sources (oracle/_ref, when present).
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import TIGHT
from lp_cases import STATUS, closed_form_cases, medium_lp, spmv_cases, GOLDEN
from mipx_amd import lpgen
from mipx_amd.lp import from_dense
from oracle import binding as ob


@pytest.mark.parametrize("case", spmv_cases(), ids=lambda c: c[0]["name"])
def test_oracle_spmv_known_answers(case):
    c, lp = case
    assert np.array_equal(ob.spmv(lp, c["x"]), np.array(c["Ax"], dtype=float))
    assert np.array_equal(ob.spmv_t(lp, c["xt"]), np.array(c["Atx"], dtype=float))


def test_oracle_spmv_brute_force_and_adjoint():
    g = lpgen.c0_plain(seed=7, m=40, n=90, per_col=4)
    lp = g.lp
    a = lp.dense()
    rng = np.random.default_rng(1)
    x = rng.standard_normal(lp.num_cols)
    y = rng.standard_normal(lp.num_rows)
    ax = ob.spmv(lp, x)
    aty = ob.spmv_t(lp, y)
    np.testing.assert_allclose(ax, a @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(aty, a.T @ y, rtol=1e-13, atol=1e-13)
    assert abs(ax @ y - x @ aty) <= 1e-11 * (1 + abs(ax @ y))
    # A transposed-index mistake would break this: row 0 of A·e_j is a[0, j].
    e = np.zeros(lp.num_cols)
    e[3] = 1.0
    assert np.array_equal(ob.spmv(lp, e), a[:, 3])


@pytest.mark.parametrize("case", closed_form_cases(), ids=lambda c: c[0]["name"])
def test_oracle_closed_form(case):
    c, lp = case
    opts = dict(verbose=0)
    opts.update(c.get("options", {}))
    r, x, y = ob.solve(lp, **opts)
    assert r["status"] == STATUS[c["status"]], r
    if "objective" in c:
        assert abs(r["objective"] - c["objective"]) <= c["tol"]
    if "x" in c:
        np.testing.assert_allclose(x, c["x"], atol=1e-5)
    if "iterations" in c:
        assert r["iterations"] == c["iterations"]
    if c.get("farkas_check"):
        ray = y[0] / abs(y[0])
        assert ray >= -1e-6
        assert ray * (-1.0) - ray * 1.0 < -1e-6


def test_oracle_scaling_invariants():
    g = lpgen.c0_plain(seed=11, m=60, n=150)
    lp = g.lp
    s = ob.scale_problem(lp)
    a = lp.dense()
    R, Cs = s["row_scale"], s["col_scale"]
    # Scaled matrix = R A C exactly up to rounding, in both orientations.
    sa = from_dense(np.diag(R) @ a @ np.diag(Cs), np.zeros(lp.num_cols), np.zeros(lp.num_cols),
                    np.zeros(lp.num_cols), np.zeros(lp.num_rows), np.zeros(lp.num_rows))
    np.testing.assert_allclose(s["values"], sa.values, rtol=1e-12)
    # Ruiz equilibrates the inf-norms towards 1.
    d = np.abs(np.diag(R) @ a @ np.diag(Cs))
    assert np.all(np.abs(d.max(axis=1) - 1) < 0.05)
    assert np.all(np.abs(d.max(axis=0) - 1) < 0.05)
    # Pock-Chambolle bases are 1/||row||_2 and 1/||col||_2 of the scaled matrix.
    np.testing.assert_allclose(s["sigma_base"], 1 / np.linalg.norm(np.diag(R) @ a @ np.diag(Cs), axis=1), rtol=1e-13)
    np.testing.assert_allclose(s["tau_base"], 1 / np.linalg.norm(np.diag(R) @ a @ np.diag(Cs), axis=0), rtol=1e-13)
    # Objective rescaling closed form: os = 1/(||c~||+1) with c~ = c*C before os.
    ct = lp.obj * Cs
    assert math.isclose(s["objective_scale"], 1 / (np.linalg.norm(ct) + 1), rel_tol=1e-14)
    np.testing.assert_allclose(s["cscaled"], ct * s["objective_scale"], rtol=1e-15)
    # Power iteration: sigma_max of the scaled matrix (SVD), to the 1e-4 stop.
    smax = np.linalg.svd(np.diag(R) @ a @ np.diag(Cs), compute_uv=False)[0]
    assert abs(s["spectral_norm"] - smax) <= 1e-3 * smax


def test_oracle_matches_reference_build():
    if not ob.ref_available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    for seed, (m, n) in [(0xC0, (2000, 5000)), (3, (120, 300))]:
        lp = lpgen.c0_plain(seed=seed, m=m, n=n).lp
        r1, x1, y1 = ob.solve(lp, **TIGHT)
        r2, x2, y2 = ob.ref_solve(lp, **TIGHT)
        assert r1["status"] == r2["status"]
        assert r1["iterations"] == r2["iterations"]
        assert r1["objective"] == r2["objective"]
        assert r1["work_units"] == r2["work_units"]
        assert np.array_equal(x1, x2) and np.array_equal(y1, y2)
    for c, lp in closed_form_cases():
        if c["name"] == "time_limit_zero":
            continue
        opts = dict(verbose=0)
        opts.update(c.get("options", {}))
        r1, x1, y1 = ob.solve(lp, **opts)
        r2, x2, y2 = ob.ref_solve(lp, **opts)
        assert (r1["status"], r1["iterations"]) == (r2["status"], r2["iterations"]), c["name"]
        assert r1["objective"] == r2["objective"]


def test_oracle_medium_lp_kkt():
    lp = medium_lp()
    r, x, y = ob.solve(lp, **TIGHT)
    assert r["status"] == 0
    k = ob.kkt_check(lp, x, y)
    assert k["rel_primal"] <= 2e-6 and k["rel_gap"] <= 2e-6


def test_oracle_highs_fixtures():
    path = os.path.join(GOLDEN, "highs_optima.json")
    with open(path) as f:
        data = json.load(f)
    for case in data["cases"]:
        lp = lpgen.c0_plain(seed=case["seed"], m=case["m"], n=case["n"],
                            per_col=case["per_col"]).lp
        r, x, y = ob.solve(lp, **TIGHT)
        assert r["status"] == 0
        rel = abs(r["objective"] - case["objective"]) / (1 + abs(case["objective"]))
        assert rel <= 5e-6, (case, r["objective"], rel)


def test_kkt_checker_detects_planted_errors():
    """The certificate must fail for a planted mistake in (x, y)."""
    lp = lpgen.c0_plain(seed=5, m=50, n=120).lp
    r, x, y = ob.solve(lp, **TIGHT)
    good = ob.kkt_check(lp, x, y)
    assert good["rel_primal"] <= 2e-6 and good["rel_gap"] <= 2e-6
    bad_y = ob.kkt_check(lp, x, -y)          # wrong dual sign
    assert bad_y["rel_gap"] > 1e-3 or bad_y["dual_sign_viol"] > 1e-3
    xb = x.copy()
    xb[0] += 1.0                               # perturbed primal
    assert ob.kkt_check(lp, xb, y)["rel_primal"] > 1e-5


def test_generator_is_deterministic_and_feasible():
    g1 = lpgen.c0_plain(seed=42, m=100, n=250)
    g2 = lpgen.c0_plain(seed=42, m=100, n=250)
    assert np.array_equal(g1.lp.values, g2.lp.values)
    lp = g1.lp
    ax = ob.spmv(lp, g1.x0)
    eq = lp.row_lower == lp.row_upper
    np.testing.assert_allclose(ax[eq], lp.row_lower[eq], rtol=1e-12, atol=1e-9)
    assert np.all(ax[~eq] <= lp.row_upper[~eq] + 1e-9)
    assert np.all(np.diff(lp.row_starts) >= 0)
    for i in range(lp.num_rows):
        seg = lp.col_indices[lp.row_starts[i]:lp.row_starts[i + 1]]
        assert np.all(np.diff(seg) > 0)
    assert lp.nnz == 250 * 5


def test_oracle_objective_envelope_vs_exact_optimum():
    """Pins how far the reference's PDLP (oracle, bit-identical to it) lands
    from the exact optimum at tol 1e-6: the mixed-precision bar is set from
    this envelope."""
    from lp_cases import highs_objective
    errs = []
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    for i in range(8):
        lo, hi = g.branch_bounds(lpgen.c2_instance_seed(i))
        lp = g.lp.with_col_bounds(lo, hi)
        r, _, _ = ob.solve(lp, **TIGHT)
        opt = highs_objective(lp)
        errs.append(abs(r["objective"] - opt) / (1 + abs(opt)))
    assert max(errs) <= 3e-6


# ---- non-default options: Anderson acceleration, adaptive step, matrix filter ----

OPTIONAL = [
    dict(anderson_acceleration=1),
    dict(anderson_acceleration=1, anderson_window_size=3, anderson_regularization=1e-8),
    dict(adaptive_step=1, adaptive_step_up=1.05, adaptive_step_down=0.5, max_iter=12000),
    dict(adaptive_step=1, anderson_acceleration=1),
]


def _with_small_entries(lp, every=7, factor=1e-7):
    """The same LP with every `every`-th matrix entry shrunk by `factor`."""
    import copy
    out = copy.deepcopy(lp)
    out.values = out.values.copy()
    out.values[::every] *= factor
    return out


def _dropped(lp, tol):
    """The LP with entries |a| < tol removed, as a new CSR (no filtering option)."""
    import copy
    keep = np.abs(lp.values) >= tol
    rows = np.repeat(np.arange(lp.num_rows), np.diff(lp.row_starts))
    out = copy.deepcopy(lp)
    out.values = lp.values[keep].copy()
    out.col_indices = lp.col_indices[keep].copy()
    out.row_starts = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=lp.num_rows))]).astype(np.int32)
    return out


@pytest.mark.parametrize("opts", OPTIONAL, ids=lambda o: "-".join(sorted(o)))
def test_oracle_optional_features_match_reference_build(opts):
    """Anderson acceleration (pdlp.cpp:293-437, 1339-1378) and the CPU loop's
    adaptive step (1004-1040): bit-identical to the reference built from its
    sources, including runs whose step grows and shrinks."""
    if not ob.ref_available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    for seed, (m, n) in [(3, (120, 300)), (0xC0, (600, 1500))]:
        lp = lpgen.c0_plain(seed=seed, m=m, n=n).lp
        o = dict(TIGHT)
        o.update(opts)
        r1, x1, y1 = ob.solve(lp, **o)
        r2, x2, y2 = ob.ref_solve(lp, **o)
        assert (r1["status"], r1["iterations"]) == (r2["status"], r2["iterations"])
        assert r1["objective"] == r2["objective"]
        assert np.array_equal(x1, x2) and np.array_equal(y1, y2)


def test_oracle_matrix_zero_tol():
    """pdlp.cpp:499-518: entries with |a| < matrix_zero_tol are dropped before
    scaling; the result equals solving the LP with those entries removed, and
    the reference build's."""
    lp = _with_small_entries(lpgen.c0_plain(seed=5, m=150, n=400).lp)
    tol = 1e-5
    r1, x1, y1 = ob.solve(lp, matrix_zero_tol=tol, **TIGHT)
    r2, x2, y2 = ob.solve(_dropped(lp, tol), **TIGHT)
    assert r1["status"] == 0
    assert r1["iterations"] == r2["iterations"] and r1["objective"] == r2["objective"]
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2)
    r0, _, _ = ob.solve(lp, **TIGHT)  # the filter changes the run
    assert r0["iterations"] != r1["iterations"] or r0["objective"] != r1["objective"]
    if ob.ref_available():
        r3, x3, y3 = ob.ref_solve(lp, matrix_zero_tol=tol, **TIGHT)
        assert r3["iterations"] == r1["iterations"] and r3["objective"] == r1["objective"]
        assert np.array_equal(x1, x3) and np.array_equal(y1, y3)


@pytest.mark.parametrize("N,up,down", [(200, 1.05, 0.5), (20, 3.0, 0.5), (50, 4.0, 0.25)])
def test_oracle_gpu_adaptive_step_rule(N, up, down):
    """The GPU loop's adaptive step (pdlp.cpp:2271-2307), rebuilt outside the
    oracle: the test drives the Halpern iteration itself (orc_pdhg_iterate),
    takes x^ and A x^ (scipy) at each evaluation and applies the rule; with
    restarts limited to the first evaluation the oracle's step after each
    evaluation must be that value.  The chosen factors make the step both grow
    and shrink."""
    import scipy.sparse as sps
    lp = lpgen.c0_plain(seed=3, m=120, n=300).lp
    o = dict(TIGHT, adaptive_step=1, adaptive_step_up=up, adaptive_step_down=down,
             termination_eval_frequency=N, restart_artificial_fraction=1.0,
             restart_sufficient_decay=0.0, restart_necessary_decay=0.0)
    sc = ob.scale_problem(lp, **o)
    a = sps.csr_matrix((sc["values"], lp.col_indices, lp.row_starts),
                       shape=(lp.num_rows, lp.num_cols))
    sig, tau = sc["sigma_base"], sc["tau_base"]
    step0 = 0.998 / sc["spectral_norm"]
    r1, _, _ = ob.solve_gpu(lp, **dict(o, max_iter=N))
    assert r1["restarts"] == 1
    m, n = lp.num_rows, lp.num_cols
    st = {k: np.zeros(n) for k in ("cx", "ix", "px")}
    st.update({k: np.zeros(m) for k in ("cy", "iy", "py")})
    ob.pdhg_iterate(lp, sc, N, 0, step0, 1.0, st)
    for dst, src in (("cx", "px"), ("ix", "px"), ("cy", "py"), ("iy", "py")):
        st[dst][:] = st[src]
    step, prev, inner = step0, None, 0
    seen = set()
    for e in range(2, 7):
        ob.pdhg_iterate(lp, sc, N, inner, step, r1["primal_weight"], st)
        inner += N
        z = st["px"].copy()
        az = a @ z
        if prev is not None:
            ad = np.sum(sig * (az - prev[1]) ** 2)
            xd = np.sum((z - prev[0]) ** 2 / tau)
            local = 0.998 / np.sqrt(max(ad / xd, 1e-30))
            assert abs(step - local) > 1e-6 * local  # no near-tie
            seen.add(step > local)
            step = min(max(step * (down if step > local else up), step0 * 1e-6), step0 * 1e3)
        prev = (z, az)
        r, _, _ = ob.solve_gpu(lp, **dict(o, max_iter=e * N))
        assert r["restarts"] == 1
        assert r["step"] == pytest.approx(step, rel=1e-14), e
    if up > 2:
        assert seen == {True, False}


@pytest.mark.parametrize("opts", [dict(adaptive_step=1), dict(anderson_acceleration=1)],
                         ids=["adaptive", "anderson"])
def test_oracle_closed_form_with_optional_features(opts):
    """The reference's feature-toggle tests (tests/test_pdlp.cpp:739-843,
    1207-1340): the closed-form LPs keep their optima with the feature on, in
    both loops."""
    for c, lp in closed_form_cases():
        if c["status"] != "Optimal" or "objective" not in c:
            continue
        o = dict(verbose=0)
        o.update(c.get("options", {}))
        o.update(opts)
        for solve in (ob.solve, ob.solve_gpu):
            r, x, y = solve(lp, **o)
            assert r["status"] == 0, c["name"]
            assert abs(r["objective"] - c["objective"]) <= max(c["tol"], 1e-4), c["name"]


def test_oracle_preconditioner_refresh():
    """pdlp.cpp:1325-1340 / 713-742: on a restart whose primal weight drifted
    past the ratio the reference recomputes Pock-Chambolle sigma/tau from the
    scaled matrix's row and column 2-norms.  Those are the values setup
    computed from the same matrix, so a refresh only moves the recorded
    weight: runs with refreshes forced (ratio 1.01) equal runs without them,
    and equal the reference build's."""
    for seed, (m, n) in [(3, (120, 300)), (0xC0, (600, 1500))]:
        lp = lpgen.c0_plain(seed=seed, m=m, n=n).lp
        base = dict(TIGHT, primal_weight=1.0)
        on = dict(base, preconditioner_refresh=1, preconditioner_refresh_ratio=1.01)
        r1, x1, y1 = ob.solve(lp, **on)
        r0, x0, y0 = ob.solve(lp, **base)
        assert r1["status"] == 0
        # The weight moves only at restarts and ends more than 1% from its
        # start, so at least one restart refreshed.
        assert r1["restarts"] > 0 and abs(r1["primal_weight"] - 1.0) > 0.01
        assert (r1["iterations"], r1["objective"]) == (r0["iterations"], r0["objective"])
        assert np.array_equal(x1, x0) and np.array_equal(y1, y0)
        if ob.ref_available():
            r2, x2, y2 = ob.ref_solve(lp, **on)
            assert (r1["status"], r1["iterations"]) == (r2["status"], r2["iterations"])
            assert r1["objective"] == r2["objective"]
            assert np.array_equal(x1, x2) and np.array_equal(y1, y2)


def test_oracle_reduced_costs_dense():
    """orc_reduced_costs = c + A'y in original space (src/lp/pdlp.cpp:2808-2819),
    against a dense product; a transposed or sign-flipped product fails it
    (the matrix is not square and y has zero entries that the scatter skips)."""
    lp = lpgen.c0_plain(seed=11, m=40, n=90, per_col=4).lp
    rng = np.random.default_rng(3)
    y = rng.standard_normal(lp.num_rows)
    y[::5] = 0.0
    rc = ob.reduced_costs(lp, y)
    want = lp.obj + lp.dense().T @ y
    np.testing.assert_allclose(rc, want, rtol=1e-13, atol=1e-13)
    assert np.abs(rc - (lp.obj - lp.dense().T @ y)).max() > 1e-3
    # At an optimum of a boxed LP, columns strictly inside their bounds have
    # zero reduced cost, those at a lower (upper) bound a nonnegative
    # (nonpositive) one, up to the solve's tolerance.
    lp2 = lpgen.c0_plain(seed=5, m=50, n=120).lp
    r, x, yy = ob.solve(lp2, **TIGHT)
    rc2 = ob.reduced_costs(lp2, yy)
    scale = 1e-4 * (1 + np.abs(lp2.obj).max())
    inside = (x > lp2.col_lower + 1e-6) & (x < lp2.col_upper - 1e-6)
    assert inside.any() and np.abs(rc2[inside]).max() <= scale
    assert rc2[x <= lp2.col_lower + 1e-9].min() >= -scale
    assert rc2[x >= lp2.col_upper - 1e-9].max() <= scale


@pytest.mark.parametrize("which", ["boxed", "one_sided"])
def test_certificate_agrees_with_oracle_report(which):
    """The certify() helper used by the GPU tests holds for the oracle's own
    results: reported residuals agree with the long-double certificate."""
    from lp_cases import certify, one_sided_lp
    lp = (one_sided_lp(seed=0x5D0, m=200, n=500) if which == "one_sided"
          else lpgen.c0_plain(seed=101, m=300, n=750).lp)
    r, x, y = ob.solve(lp, **TIGHT)
    assert r["status"] == 0
    k = certify(lp, r, x, y)
    if which == "one_sided":
        assert k["rel_dual"] > 0.0
        # A planted dual error shows up in the dual residual itself.
        bad = ob.kkt_check(lp, x, 1.01 * y)
        assert bad["rel_dual"] > 1e-4


def test_farkas_bounds_variant_detected_by_oracle():
    """A bounds variant made infeasible by farkas_bounds is certified by the
    reference's own rule (oracle/_ref when built) and by the oracle at the
    first certificate check, with the ray on the violated row."""
    from lp_cases import farkas_bounds
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    lo, hi = farkas_bounds(g.lp, *g.branch_bounds(lpgen.c2_instance_seed(7)), 17)
    lp = g.lp.with_col_bounds(lo, hi)
    r, x, y = ob.solve(lp, max_iter=100000, **TIGHT)
    assert r["status"] == 1 and r["iterations"] == 2000
    ray = y / np.linalg.norm(y)
    assert abs(abs(ray[17]) - 1.0) <= 1e-3
    if ob.ref_available():
        rr, _, yr = ob.ref_solve(lp, max_iter=100000, **TIGHT)
        assert rr["status"] == 1 and rr["iterations"] == 2000 and np.array_equal(y, yr)
