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
