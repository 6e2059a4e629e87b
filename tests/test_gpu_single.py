"""GPU parity tests of the single-LP path (mpx_pdlp_solve and its kernels)
against the oracle, through the C-ABI."""
import numpy as np
import pytest

from conftest import TIGHT
from lp_cases import STATUS, certify, closed_form_cases, medium_lp, one_sided_lp, spmv_cases
from mipx_amd import lpgen, native
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.mark.parametrize("group", [1, 2, 4, 8, 16, 64])
def test_spmv_device_parity(group):
    torch = _torch()
    lp = lpgen.c1_large(seed=9, m=20000, n=40000).lp
    x = np.random.default_rng(0).standard_normal(lp.num_cols)
    want = ob.spmv(lp, x)
    dev = lambda a: torch.from_numpy(a).cuda()
    y = torch.zeros(lp.num_rows, dtype=torch.float64, device="cuda")
    native.spmv_device(lp.num_rows, lp.num_cols, dev(lp.row_starts), dev(lp.col_indices),
                       dev(lp.values), dev(x), y, group)
    torch.cuda.synchronize()
    got = y.cpu().numpy()
    if group == 1:
        assert np.array_equal(got, want)      # reference summation order
    else:
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("case", spmv_cases(), ids=lambda c: c[0]["name"])
def test_spmv_known_answers_on_device(case):
    torch = _torch()
    c, lp = case
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    y = torch.zeros(lp.num_rows, dtype=torch.float64, device="cuda")
    native.spmv_device(lp.num_rows, lp.num_cols, dev(lp.row_starts), dev(lp.col_indices),
                       dev(lp.values), dev(np.array(c["x"], dtype=float)), y, 1)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), np.array(c["Ax"], dtype=float))


def test_scaling_bitwise_vs_oracle():
    lp = lpgen.c1_large(seed=3, m=3000, n=6000).lp
    got = native.scale_problem(lp)
    want = ob.scale_problem(lp)
    for k in ("values", "at_values", "at_row_starts", "at_col_indices", "row_scale",
              "col_scale", "sigma_base", "tau_base"):
        assert np.array_equal(got[k], want[k]), k
    # Global sums are reduced in a different (fixed) order on the device.
    for k in ("objective_scale", "constraint_scale"):
        assert abs(got[k] - want[k]) <= 1e-14 * abs(want[k]), k
    for k in ("cscaled", "col_lower", "col_upper", "row_lower", "row_upper"):
        np.testing.assert_allclose(got[k], want[k], rtol=1e-14, atol=0)
    assert abs(got["spectral_norm"] - want["spectral_norm"]) <= 1e-6 * want["spectral_norm"]


@pytest.mark.parametrize("group", [1, 4])
def test_halpern_iterations_vs_oracle(group):
    lp = lpgen.c0_plain(seed=21, m=400, n=1000).lp
    s = ob.scale_problem(lp)
    step = 0.998 / s["spectral_norm"]
    pw = 0.7
    rng = np.random.default_rng(4)
    def state():
        return dict(cx=rng.uniform(0, 1, lp.num_cols) * 0, ix=np.zeros(lp.num_cols),
                    px=np.zeros(lp.num_cols), cy=np.zeros(lp.num_rows),
                    iy=np.zeros(lp.num_rows), py=np.zeros(lp.num_rows))
    a, b = state(), state()
    ob.pdhg_iterate(lp, s, 60, 3, step, pw, a)
    native.pdhg_iterate(lp, s, 60, 3, step, pw, b, precision=0, group=group)
    for k in ("cx", "px", "cy", "py"):
        if group == 1:
            assert np.array_equal(a[k], b[k]), k
        else:
            np.testing.assert_allclose(b[k], a[k], rtol=1e-9, atol=1e-12)


def test_halpern_fp32_iterations_close_to_fp64():
    lp = lpgen.c0_plain(seed=22, m=400, n=1000).lp
    s = ob.scale_problem(lp)
    step = 0.998 / s["spectral_norm"]
    mk = lambda: dict(cx=np.zeros(lp.num_cols), ix=np.zeros(lp.num_cols),
                      px=np.zeros(lp.num_cols), cy=np.zeros(lp.num_rows),
                      iy=np.zeros(lp.num_rows), py=np.zeros(lp.num_rows))
    a, b = mk(), mk()
    ob.pdhg_iterate(lp, s, 40, 0, step, 1.0, a)
    native.pdhg_iterate(lp, s, 40, 0, step, 1.0, b, precision=1, group=4)
    for k in ("cx", "px", "cy", "py"):
        scale = max(1e-30, np.abs(a[k]).max())
        assert np.abs(a[k] - b[k]).max() <= 1e-4 * scale, k


@pytest.mark.parametrize("case", closed_form_cases(), ids=lambda c: c[0]["name"])
def test_closed_form_on_gpu(case):
    c, lp = case
    opts = dict(verbose=0)
    opts.update(c.get("options", {}))
    r, x, y = native.solve(lp, **opts)
    assert r["status"] == STATUS[c["status"]], r
    if "objective" in c:
        assert abs(r["objective"] - c["objective"]) <= c["tol"]
    if "x" in c:
        np.testing.assert_allclose(x, c["x"], atol=1e-5)
    if "iterations" in c:
        assert r["iterations"] == c["iterations"]
    if c.get("farkas_check"):
        ray = y[0] / abs(y[0])
        assert ray >= -1e-6 and ray * (-1.0) - ray * 1.0 < -1e-6
    ro, _, _ = ob.solve(lp, **opts)
    assert ro["status"] == r["status"]


FP64_CASES = [("c0", 0xC0, 2000, 5000), ("c0", 101, 300, 750), ("one_sided", 0x5D1, 300, 750)]


def _fp64_lp(kind, seed, m, n):
    return (one_sided_lp(seed, m, n) if kind == "one_sided"
            else lpgen.c0_plain(seed=seed, m=m, n=n).lp)


@pytest.mark.parametrize("group", [1, 0])
@pytest.mark.parametrize("case", FP64_CASES, ids=lambda c: f"{c[0]}-{c[1]:#x}-{c[2]}x{c[3]}")
def test_solve_fp64_vs_oracle(case, group):
    """Plain FP64 GPU PDLP against the oracle: same status, objective within
    1e-6, the returned point certified (with the device's own residuals
    agreeing with the certificate), same work-unit formula.  With one lane
    per row / column (group 1) the device follows the oracle's trajectory:
    the iteration count is the oracle's to within one evaluation block (the
    power iteration's last bits differ, since its norms are device
    reductions) and so is the restart count."""
    kind, seed, m, n = case
    lp = _fp64_lp(kind, seed, m, n)
    r, x, y = native.solve(lp, spmv_group=group, **TIGHT)
    ro, xo, yo = ob.solve(lp, **TIGHT)
    print(f"gpu iters={r['iterations']} restarts={r['restarts']} obj={r['objective']!r}  "
          f"oracle iters={ro['iterations']} restarts={ro['restarts']} obj={ro['objective']!r}")
    assert r["status"] == ro["status"] == 0
    assert abs(r["objective"] - ro["objective"]) <= 1e-6 * (1 + abs(ro["objective"]))
    certify(lp, r, x, y)
    assert r["work_units"] == r["iterations"] * (4.0 * lp.nnz + 2.0 * (m + n))
    if group == 1:
        assert abs(r["iterations"] - ro["iterations"]) <= 200
        assert abs(r["restarts"] - ro["restarts"]) <= 1
    else:
        assert abs(r["iterations"] - ro["iterations"]) <= 0.25 * ro["iterations"]


@pytest.mark.parametrize("seed,m,n", [(0xC0, 2000, 5000), (101, 300, 750), (103, 200, 500)])
def test_solve_mixed_precision(seed, m, n):
    """FP32 iterations, FP64 evaluation and refinement: objective within the
    reference's own error envelope at tol 1e-6 (3e-6 of the committed
    scipy-HiGHS optimum, 5e-6 of the oracle) and KKT-certified."""
    import json, os
    from lp_cases import GOLDEN
    fixture = json.load(open(os.path.join(GOLDEN, "highs_optima.json")))["cases"]
    case = [c for c in fixture if c["seed"] == seed][0]
    opt = case["objective"]
    lp = lpgen.c0_plain(seed=seed, m=m, n=n, per_col=case["per_col"]).lp
    r, x, y = native.solve(lp, precision=1, **TIGHT)
    ro, _, _ = ob.solve(lp, **TIGHT)
    print(r)
    assert r["status"] == 0 and r["iterations_fp32"] > 0
    assert abs(r["objective"] - opt) <= 3e-6 * (1 + abs(opt))
    assert abs(r["objective"] - ro["objective"]) <= 5e-6 * (1 + abs(ro["objective"]))
    certify(lp, r, x, y)


def test_medium_lp_gpu_matches_cpu():
    """Reference [gpu] tests (tests/test_pdlp.cpp:517-546): GPU and CPU
    objectives agree within 1e-4 on buildMediumLp (default options)."""
    lp = medium_lp()
    r, _, _ = native.solve(lp, verbose=0, max_iter=200000)
    ro, _, _ = ob.solve(lp, verbose=0, max_iter=200000)
    assert r["status"] == ro["status"] == 0
    assert abs(r["objective"] - ro["objective"]) <= 1e-4


def test_warm_start_reduces_iterations():
    """Reference warm-start test (tests/test_pdlp.cpp:1076-1113)."""
    lp = lpgen.c0_plain(seed=101, m=300, n=750).lp
    r1, x1, y1 = native.solve(lp, **TIGHT)
    r2, x2, y2 = native.solve(lp, warm=(x1, y1), **TIGHT)
    assert r2["status"] == 0 and r2["iterations"] < r1["iterations"]
    ro, _, _ = ob.solve(lp, warm=(x1, y1), **TIGHT)
    assert abs(r2["objective"] - ro["objective"]) <= 1e-6 * (1 + abs(ro["objective"]))


def test_stop_callback_reports_iter_limit():
    from mipx_amd._abi import STOP_FN
    lp = lpgen.c0_plain(seed=101, m=300, n=750).lp
    calls = []
    def stop(_user):
        calls.append(1)
        return 1 if len(calls) > 2 else 0
    cb = STOP_FN(stop)
    r, _, _ = native.solve(lp, stop_fn=cb, **TIGHT)
    assert r["status"] == 3 and r["iterations"] == 400


# ---- non-default options: Anderson acceleration, adaptive step, matrix filter ----

OPTIONAL = [
    dict(anderson_acceleration=1),
    dict(anderson_acceleration=1, anderson_window_size=3, anderson_regularization=1e-8),
    dict(adaptive_step=1, adaptive_step_up=1.05, adaptive_step_down=0.5, max_iter=30000),
    dict(adaptive_step=1, anderson_acceleration=1),
]


@pytest.mark.parametrize("group", [1, 0])
@pytest.mark.parametrize("opts", OPTIONAL, ids=lambda o: "-".join(sorted(o)))
def test_optional_features_vs_oracle(opts, group):
    """Anderson acceleration at restarts and the GPU loop's adaptive step:
    same status and objective (1e-6) as the oracle's GPU-loop solve, KKT
    certified; with one lane per row / column the iteration count is the
    oracle's to within one evaluation block (the Gram matrix and safeguard
    sums are device reductions) and the restart count to within one."""
    lp = lpgen.c0_plain(seed=101, m=300, n=750).lp
    o = dict(TIGHT)
    o.update(opts)
    r, x, y = native.solve(lp, spmv_group=group, **o)
    ro, _, _ = ob.solve_gpu(lp, **o)
    print(f"gpu iters={r['iterations']} step={r['step']!r} obj={r['objective']!r}  "
          f"oracle iters={ro['iterations']} step={ro['step']!r} obj={ro['objective']!r}")
    assert r["status"] == ro["status"] == 0
    assert abs(r["objective"] - ro["objective"]) <= 1e-6 * (1 + abs(ro["objective"]))
    certify(lp, r, x, y)
    if group == 1:
        assert abs(r["iterations"] - ro["iterations"]) <= 200
        assert abs(r["restarts"] - ro["restarts"]) <= 1
    else:
        assert abs(r["iterations"] - ro["iterations"]) <= 0.25 * ro["iterations"]


@pytest.mark.parametrize("N,up,down", [(200, 1.05, 0.5), (20, 3.0, 0.5), (50, 4.0, 0.25)])
def test_adaptive_step_trajectory_vs_oracle(N, up, down):
    """With one restart (after the first evaluation) the step after each of
    the next five evaluations equals the oracle's GPU-loop step: the same
    grow / shrink decisions (group 1: the reference's SpMV order)."""
    lp = lpgen.c0_plain(seed=3, m=120, n=300).lp
    o = dict(TIGHT, adaptive_step=1, adaptive_step_up=up, adaptive_step_down=down,
             termination_eval_frequency=N, restart_artificial_fraction=1.0,
             restart_sufficient_decay=0.0, restart_necessary_decay=0.0)
    for e in range(2, 7):
        r, _, _ = native.solve(lp, spmv_group=1, **dict(o, max_iter=e * N))
        ro, _, _ = ob.solve_gpu(lp, **dict(o, max_iter=e * N))
        assert r["restarts"] == ro["restarts"] == 1
        # The power iteration's last bits differ; the grow / shrink factors
        # applied to it must be the same.
        assert r["step"] == pytest.approx(ro["step"], rel=1e-12), (e, r["step"], ro["step"])


def test_adaptive_and_anderson_mixed_precision():
    """FP32 iterations with both accelerations, FP64 refinement: KKT
    certified and within 5e-6 of the oracle's objective."""
    lp = lpgen.c0_plain(seed=0xC0, m=2000, n=5000).lp
    o = dict(TIGHT, adaptive_step=1, anderson_acceleration=1)
    r, x, y = native.solve(lp, precision=1, **o)
    ro, _, _ = ob.solve_gpu(lp, **o)
    print(r)
    assert r["status"] == ro["status"] == 0 and r["iterations_fp32"] > 0
    assert abs(r["objective"] - ro["objective"]) <= 5e-6 * (1 + abs(ro["objective"]))
    certify(lp, r, x, y)


def test_matrix_zero_tol_on_gpu():
    """pdlp.cpp:499-518 on the device path: the filtered solve equals the solve
    of the LP with the small entries removed (bit for bit), matches the
    oracle's objective, and the work units count the kept entries."""
    import copy
    lp = lpgen.c0_plain(seed=5, m=150, n=400).lp
    lp = copy.deepcopy(lp)
    lp.values = lp.values.copy()
    lp.values[::7] *= 1e-7
    tol = 1e-5
    keep = np.abs(lp.values) >= tol
    rows = np.repeat(np.arange(lp.num_rows), np.diff(lp.row_starts))
    dropped = copy.deepcopy(lp)
    dropped.values = lp.values[keep].copy()
    dropped.col_indices = lp.col_indices[keep].copy()
    dropped.row_starts = np.concatenate(
        [[0], np.cumsum(np.bincount(rows[keep], minlength=lp.num_rows))]).astype(np.int32)
    r1, x1, y1 = native.solve(lp, matrix_zero_tol=tol, **TIGHT)
    r2, x2, y2 = native.solve(dropped, **TIGHT)
    ro, _, _ = ob.solve(lp, matrix_zero_tol=tol, **TIGHT)
    assert r1["status"] == r2["status"] == ro["status"] == 0
    assert r1["iterations"] == r2["iterations"] and r1["objective"] == r2["objective"]
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2)
    assert abs(r1["objective"] - ro["objective"]) <= 1e-6 * (1 + abs(ro["objective"]))
    assert r1["work_units"] == r1["iterations"] * (4.0 * int(keep.sum()) + 2.0 * (150 + 400))


def test_concurrent_solves_and_race_stop():
    """The concurrent-root race contract (mip_solver.cpp:4478-4500): solves on
    their own threads share no state (own stream and memory per solve) and
    give results identical to sequential solves; a shared stop flag set by the
    first Optimal finisher cancels the other, which reports IterLimit with
    work units = iterations * (4 nnz + 2 (m + n)), no wall-clock term."""
    import threading
    from mipx_amd._abi import STOP_FN
    small = lpgen.c0_plain(seed=101, m=300, n=750).lp
    big = lpgen.c0_plain(seed=0xC0, m=2000, n=5000).lp
    seq = {k: native.solve(lp, **TIGHT) for k, lp in (("small", small), ("big", big))}

    def both(stop_fn=None, on_done=None):
        out = {}

        def run(key, lp):
            kw = dict(TIGHT)
            if stop_fn is not None:
                kw["stop_fn"] = stop_fn
            out[key] = native.solve(lp, **kw)
            if on_done:
                on_done(out[key])
        ts = [threading.Thread(target=run, args=kv) for kv in (("small", small), ("big", big))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        return out

    par = both()
    for k in ("small", "big"):
        (ra, xa, ya), (rb, xb, yb) = par[k], seq[k]
        assert (ra["status"], ra["iterations"], ra["objective"]) == \
            (rb["status"], rb["iterations"], rb["objective"]), k
        assert np.array_equal(xa, xb) and np.array_equal(ya, yb), k

    flag = threading.Event()
    cb = STOP_FN(lambda _u: 1 if flag.is_set() else 0)
    race = both(cb, lambda r: flag.set() if r[0]["status"] == 0 else None)
    rs, rbg = race["small"][0], race["big"][0]
    assert rs["status"] == 0 and rs["iterations"] == seq["small"][0]["iterations"]
    assert rbg["status"] == 3 and 0 <= rbg["iterations"] < seq["big"][0]["iterations"]
    assert rbg["iterations"] % 200 == 0
    assert rbg["work_units"] == rbg["iterations"] * (4.0 * big.nnz + 2.0 * (2000 + 5000))


@pytest.mark.parametrize("precision", [0, 1])
def test_chunked_dual_spmv(precision):
    """The dual step's A x_r in column chunks (automatic for n > 2^20, forced
    here with spmv_chunk_cols): same status and objective as the one-launch
    SpMV and as the oracle, KKT certified."""
    lp = lpgen.c0_plain(seed=0xC0, m=2000, n=5000).lp
    base = dict(TIGHT, precision=precision)
    r1, x1, y1 = native.solve(lp, spmv_chunk_cols=1300, **base)   # 4 chunks
    r0, _, _ = native.solve(lp, spmv_chunk_cols=-1, **base)
    ro, _, _ = ob.solve(lp, **TIGHT)
    print(r1["iterations"], r0["iterations"], ro["iterations"])
    assert r1["status"] == r0["status"] == ro["status"] == 0
    tol = 1e-6 if precision == 0 else 5e-6
    assert abs(r1["objective"] - ro["objective"]) <= tol * (1 + abs(ro["objective"]))
    assert abs(r1["objective"] - r0["objective"]) <= tol * (1 + abs(r0["objective"]))
    assert abs(r1["iterations"] - r0["iterations"]) <= 0.25 * r0["iterations"]
    certify(lp, r1, x1, y1)


def test_chunked_dual_one_lane_follows_oracle():
    """With one lane per row the chunked A x_r continues each row's running
    sum chunk after chunk, the reference's order: the chunked solve is bit
    for bit the one-launch solve, and both follow the oracle's trajectory
    (iterations within one evaluation block, restarts within one)."""
    lp = lpgen.c0_plain(seed=0xC0, m=2000, n=5000).lp
    ro, _, _ = ob.solve(lp, **TIGHT)
    runs = {}
    for chunk in (1300, -1):
        r, x, y = native.solve(lp, spmv_group=1, spmv_chunk_cols=chunk, **TIGHT)
        runs[chunk] = (r, x, y)
        assert r["status"] == ro["status"] == 0, (chunk, r)
        assert abs(r["iterations"] - ro["iterations"]) <= 200, (chunk, r["iterations"], ro)
        assert abs(r["restarts"] - ro["restarts"]) <= 1, (chunk, r, ro)
        assert abs(r["objective"] - ro["objective"]) <= 1e-6 * (1 + abs(ro["objective"]))
        certify(lp, r, x, y)
    (ra, xa, ya), (rb, xb, yb) = runs[1300], runs[-1]
    assert ra["iterations"] == rb["iterations"] and ra["objective"] == rb["objective"]
    assert np.array_equal(xa, xb) and np.array_equal(ya, yb)


def test_preconditioner_refresh_single_path():
    """pdlp.cpp:1325-1340 / 713-742: a refresh recomputes sigma/tau from the
    unchanged scaled matrix, which leaves them as they were (pinned against
    the reference build in test_oracle_preconditioner_refresh), so runs with
    refreshes forced (ratio 1.01) are bit-identical to runs without them."""
    lp = lpgen.c0_plain(seed=3, m=120, n=300).lp
    base = dict(TIGHT, primal_weight=1.0)
    r0, x0, y0 = native.solve(lp, **base)
    r1, x1, y1 = native.solve(lp, preconditioner_refresh=1, preconditioner_refresh_ratio=1.01,
                              **base)
    assert r1["status"] == 0 and r1["restarts"] > 0 and abs(r1["primal_weight"] - 1.0) > 0.01
    assert (r1["iterations"], r1["objective"]) == (r0["iterations"], r0["objective"])
    assert np.array_equal(x1, x0) and np.array_equal(y1, y0)


@pytest.mark.parametrize("case", spmv_cases(), ids=lambda c: c[0]["name"])
def test_transpose_spmv_known_answers_on_device(case):
    """A'y on the device (sparse_matrix.cpp:195-208 known answers, the
    reference's multiplyTranspose cases): the library's own transpose of A
    (scaling switched off, so its values are A's) fed to the device SpMV."""
    torch = _torch()
    c, lp = case
    t = native.scale_problem(lp, do_ruiz_scaling=0, do_pock_chambolle_scaling=0,
                             do_bound_obj_rescaling=0)
    assert np.array_equal(np.sort(t["at_values"]), np.sort(lp.values))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    y = torch.zeros(lp.num_cols, dtype=torch.float64, device="cuda")
    native.spmv_device(lp.num_cols, lp.num_rows, dev(t["at_row_starts"]),
                       dev(t["at_col_indices"]), dev(t["at_values"]),
                       dev(np.array(c["xt"], dtype=float)), y, 1)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), np.array(c["Atx"], dtype=float))


@pytest.mark.parametrize("kind", ["c0", "one_sided"])
def test_preconditioner_refresh_vs_oracle(kind):
    """pdlp.cpp:713-742, 1325-1340 with refreshes forced (ratio 1.01) on the
    device and in the oracle: same status, objective within 1e-6, certified,
    iteration count within one evaluation block (one lane per row)."""
    lp = (lpgen.c0_plain(seed=3, m=120, n=300).lp if kind == "c0"
          else one_sided_lp(seed=0x5D0, m=200, n=500))
    o = dict(TIGHT, primal_weight=1.0, preconditioner_refresh=1,
             preconditioner_refresh_ratio=1.01)
    r, x, y = native.solve(lp, spmv_group=1, **o)
    ro, _, _ = ob.solve(lp, **o)
    print(r["iterations"], ro["iterations"], r["primal_weight"], ro["primal_weight"])
    assert r["status"] == ro["status"] == 0 and r["restarts"] > 0
    assert abs(r["objective"] - ro["objective"]) <= 1e-6 * (1 + abs(ro["objective"]))
    assert abs(r["iterations"] - ro["iterations"]) <= 200
    certify(lp, r, x, y)


def test_reduced_costs_at_device_optimum():
    """Reduced costs c + A'y (pdlp.cpp:2808-2819, oracle.orc_reduced_costs)
    of the device's dual: zero on columns strictly inside their bounds, of
    the right sign at bounds and on one-sided columns, and equal to the
    oracle's reduced costs at the oracle's optimum to the solve tolerance."""
    lp = one_sided_lp(seed=0x5D1, m=300, n=750)
    r, x, y = native.solve(lp, **TIGHT)
    assert r["status"] == 0
    rc = ob.reduced_costs(lp, y)
    np.testing.assert_allclose(rc, lp.obj + lp.dense().T @ y, rtol=1e-12, atol=1e-10)
    scale = 1e-3 * (1 + np.abs(lp.obj).max())
    lo, hi = lp.col_lower, lp.col_upper
    inside = (x > lo + 1e-5) & (x < hi - 1e-5)
    assert inside.any() and np.abs(rc[inside]).max() <= scale
    assert rc[np.isinf(hi)].min() >= -scale     # lower-only and free columns
    assert rc[np.isinf(lo)].max() <= scale      # upper-only and free columns

