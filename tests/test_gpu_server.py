"""GPU tests of the continuous-batching serving loop (mpx_server_*), through
the C-ABI: requests admitted at different times, sharing the device with
whatever else is in flight, give bit-identical results to the same LP solved
alone by the batched entry points (themselves checked against the oracle in
test_gpu_batch.py): a fresh LP as mpx_pdlp_solve_batch of that LP, a bounds
variant as mpx_pdlp_solve_batch_shared of its base with those bounds.
Admission respects the memory budget and the slot count."""
import numpy as np
import pytest

from conftest import TIGHT
from lp_cases import certify, closed_form_cases
from mipx_amd import lpgen, native
from oracle import binding as ob

pytestmark = pytest.mark.gpu

MAX_STEPS = 20000


def _drain(srv, tickets, submit_at=None):
    """Steps the server until every ticket is finished; `submit_at` maps a
    step number to a callable that submits more requests (returns tickets)."""
    out = {}
    submit_at = dict(submit_at or {})
    for k in range(MAX_STEPS):
        if k in submit_at:
            tickets += submit_at.pop(k)()
        active = srv.step()
        for t in srv.poll():
            r = srv.result(t)
            assert r is not None, t
            out[t] = r
        if active == 0 and not submit_at and len(out) == len(tickets):
            return out
    raise AssertionError(f"server did not drain in {MAX_STEPS} steps: {srv.stats()}")


def _same(a, b, tag):
    ra, xa, ya = a
    rb, xb, yb = b
    assert ra["status"] == rb["status"], (tag, ra, rb, native.last_error())
    assert ra["iterations"] == rb["iterations"], (tag, ra["iterations"], rb["iterations"])
    assert ra["objective"] == rb["objective"], (tag, ra["objective"], rb["objective"])
    assert np.array_equal(xa, xb) and np.array_equal(ya, yb), tag


@pytest.mark.parametrize("precision", [0, 1])
def test_server_results_identical_to_solo_batch(precision):
    opts = dict(precision=precision, max_iter=200000, **TIGHT)
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    fresh = [lpgen.c3_instance(100 + i, m).lp for i, m in enumerate([60, 140, 90, 200, 75])]
    bounds = [g.branch_bounds(lpgen.c2_instance_seed(i)) for i in range(8)]
    srv = native.Server(max_slots=16, **opts)
    base = srv.add_base(g.lp)
    kinds = {}

    def sub_variants(idx):
        ts = []
        for i in idx:
            t = srv.submit_bounds(base, *bounds[i])
            kinds[t] = ("variant", i)
            ts.append(t)
        return ts

    def sub_fresh(idx):
        ts = []
        for i in idx:
            t = srv.submit(fresh[i])
            kinds[t] = ("fresh", i)
            ts.append(t)
        return ts

    tickets = sub_variants([0, 1, 2]) + sub_fresh([0])
    got = _drain(srv, tickets, submit_at={
        3: lambda: sub_fresh([1, 2]) + sub_variants([3, 4, 5]),
        20: lambda: sub_variants([6, 7]) + sub_fresh([3, 4]),
    })
    st = srv.stats()
    print("server stats", st)
    assert st["finished"] == len(tickets) and st["running"] == 0 and st["pool_used_bytes"] > 0
    srv.close()
    for t, (kind, i) in kinds.items():
        if kind == "variant":
            lp = g.lp.with_col_bounds(*bounds[i])
            res, xs, ys = native.solve_batch_shared(g.lp, bounds[i][0][None], bounds[i][1][None],
                                                    **opts)
        else:
            lp = fresh[i]
            res, xs, ys = native.solve_batch([lp], **opts)
        _same(got[t], (res[0], xs[0], ys[0]), f"{kind}{i}")
        r, x, y = got[t]
        if r["status"] == 0:
            certify(lp, r, x, y, tag=f"{kind}{i}")


def test_server_two_bases_identical_to_solo_batch():
    """Bounds variants of two registered bases (different matrices, each
    iterated in its own lane-per-instance groups) and a fresh LP, submitted
    interleaved and in two waves: every result is bit-identical to the same
    request solved alone, and certified."""
    opts = dict(precision=1, max_iter=200000, **TIGHT)
    ga = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    gb = lpgen.c2_shared(seed=0xC5, m=200, n=500)
    ba = [ga.branch_bounds(lpgen.c2_instance_seed(i)) for i in range(5)]
    bb = [gb.branch_bounds(lpgen.c2_instance_seed(i, base=0xC5)) for i in range(5)]
    fresh = lpgen.c3_instance(7, 150).lp
    srv = native.Server(max_slots=16, **opts)
    base_a, base_b = srv.add_base(ga.lp), srv.add_base(gb.lp)
    kinds = {}

    def sub(which, idx):
        ts = []
        for i in idx:
            t = (srv.submit_bounds(base_a, *ba[i]) if which == "a"
                 else srv.submit_bounds(base_b, *bb[i]))
            kinds[t] = (which, i)
            ts.append(t)
        return ts

    tickets = sub("a", [0, 1]) + sub("b", [0, 1, 2])
    t = srv.submit(fresh)
    kinds[t] = ("fresh", 0)
    tickets.append(t)
    got = _drain(srv, tickets, submit_at={5: lambda: sub("b", [3, 4]) + sub("a", [2, 3, 4])})
    srv.close()
    assert len(got) == 11
    for t, (kind, i) in kinds.items():
        if kind == "fresh":
            lp = fresh
            res, xs, ys = native.solve_batch([lp], **opts)
            alone = (res[0], xs[0], ys[0])
        else:
            g, b = (ga, ba[i]) if kind == "a" else (gb, bb[i])
            lp = g.lp.with_col_bounds(*b)
            alone = _alone_shared(g, b, opts)
        _same(got[t], alone, f"{kind}{i}")
        r, x, y = got[t]
        if r["status"] == 0:
            certify(lp, r, x, y, tag=f"{kind}{i}")


def _alone_shared(g, b, opts):
    res, xs, ys = native.solve_batch_shared(g.lp, b[0][None], b[1][None], **opts)
    return res[0], xs[0], ys[0]


def test_server_slot_admission():
    """Two slots: requests queue and are admitted as slots free up."""
    opts = dict(max_iter=200000, **TIGHT)
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    bounds = [g.branch_bounds(lpgen.c2_instance_seed(40 + i)) for i in range(6)]
    srv = native.Server(max_slots=2, **opts)
    base = srv.add_base(g.lp)
    tickets = [srv.submit_bounds(base, *b) for b in bounds]
    srv.step()
    st = srv.stats()
    assert st["running"] + st["preparing"] <= 2 and st["queued"] >= 4, st
    got = _drain(srv, tickets)
    srv.close()
    for t, b in zip(tickets, bounds):
        _same(got[t], _alone_shared(g, b, opts), f"variant{t}")


def test_server_memory_admission():
    """A pool that holds the base and about two requests: the rest wait for
    memory (first in, first out); a request larger than the whole pool
    finishes with status Error."""
    opts = dict(max_iter=200000, **TIGHT)
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    bounds = [g.branch_bounds(lpgen.c2_instance_seed(60 + i)) for i in range(5)]
    probe = native.Server(hbm_budget_bytes=64e6, max_slots=64, **opts)
    base = probe.add_base(g.lp)
    b0 = probe.stats()["pool_used_bytes"]
    probe.submit_bounds(base, *bounds[0])
    probe.step()
    per = probe.stats()["pool_used_bytes"] - b0
    probe.close()
    assert per > 0
    budget = b0 + 2.5 * per
    srv = native.Server(hbm_budget_bytes=budget, max_slots=64, **opts)
    base = srv.add_base(g.lp)
    big = lpgen.c0_plain(seed=7, m=4000, n=10000).lp  # far more than the pool
    tickets = [srv.submit_bounds(base, *b) for b in bounds]
    t_big = srv.submit(big)
    srv.step()
    st = srv.stats()
    assert st["running"] + st["preparing"] == 2 and st["queued"] >= 3, st
    got = _drain(srv, tickets + [t_big])
    st = srv.stats()
    assert st["pool_peak_bytes"] <= st["pool_bytes"] <= budget + 1, st
    srv.close()
    assert got[t_big][0]["status"] == 6
    for t, b in zip(tickets, bounds):
        _same(got[t], _alone_shared(g, b, opts), f"variant{t}")


def test_server_statuses_and_time_limit():
    cases = {c["name"]: lp for c, lp in closed_form_cases()}
    srv = native.Server(max_slots=8, max_iter=100000, verbose=0)
    ts = [srv.submit(cases[k]) for k in ("infeasible", "unbounded_constrained", "simple")]
    got = _drain(srv, ts)
    assert [got[t][0]["status"] for t in ts] == [1, 2, 0]
    assert abs(got[ts[2]][0]["objective"] + 4.0) <= 1e-4
    srv.close()
    # A zero time limit stops a request after its first round.
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    srv = native.Server(max_slots=4, max_solve_seconds=0.0, **TIGHT)
    t = srv.submit(g.lp)
    r = _drain(srv, [t])[t][0]
    srv.close()
    assert r["status"] == 4 and 0 < r["iterations"] <= 2 * 200, r


def test_server_stop_callback():
    """The stop callback (reference pdlp.cpp:2136-2140) is polled once per
    step: while it returns nonzero the running requests end as IterLimit;
    once it returns zero again new requests solve as usual."""
    from mipx_amd._abi import STOP_FN
    flag = {"stop": False, "polls": 0}

    def stop(_user):
        flag["polls"] += 1
        return 1 if flag["stop"] and flag["polls"] >= 3 else 0

    cb = STOP_FN(stop)
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    srv = native.Server(max_slots=4, stop_fn=cb, **TIGHT)
    flag["stop"] = True
    t = srv.submit(g.lp)
    r = _drain(srv, [t])[t][0]
    assert r["status"] == 3 and 0 < r["iterations"] <= 4 * 200, r
    flag["stop"] = False
    t = srv.submit(g.lp)
    r2 = _drain(srv, [t])[t][0]
    srv.close()
    assert r2["status"] == 0 and r2["iterations"] > r["iterations"], r2


@pytest.mark.parametrize("precision", [0, 1])
def test_server_warm_variants_identical_to_solo_batch(precision):
    """Warm-started bounds variants (children of a solved root) in a mix with
    cold variants and a fresh LP: each result is bit-identical to the
    warm-started shared batch of that request alone."""
    opts = dict(precision=precision, max_iter=200000, **TIGHT)
    g = lpgen.c2_shared(seed=0xC2, m=300, n=800)
    r0, x0, y0 = native.solve(g.lp, **TIGHT)
    bounds = [g.branch_bounds(lpgen.c2_instance_seed(i)) for i in range(6)]
    srv = native.Server(max_slots=16, **opts)
    base = srv.add_base(g.lp)
    warm_t = {srv.submit_bounds(base, *bounds[i], warm=(x0, y0)): i for i in (0, 2, 4)}
    cold_t = {srv.submit_bounds(base, *bounds[i]): i for i in (1, 3, 5)}
    fresh = lpgen.c3_instance(7, 120).lp
    tf = srv.submit(fresh)
    got = _drain(srv, list(warm_t) + list(cold_t) + [tf])
    srv.close()
    for t, i in warm_t.items():
        res, xs, ys = native.solve_batch_shared(g.lp, bounds[i][0][None], bounds[i][1][None],
                                                warm=(x0[None], y0[None]), **opts)
        _same(got[t], (res[0], xs[0], ys[0]), f"warm{i}")
    for t, i in cold_t.items():
        _same(got[t], _alone_shared(g, bounds[i], opts), f"cold{i}")


def test_server_refuses_base_beyond_lane_offsets():
    """A base's lane-per-instance slices are addressed with 31-bit byte
    offsets: a base with more than 4,194,303 columns is refused with
    MPX_ERR_UNSUPPORTED (code 4) before anything is allocated, and the
    server keeps working."""
    from mipx_amd.lp import LpProblem
    n = 4_194_304
    big = LpProblem(1, n, np.array([0, 1]), np.array([0]), np.array([1.0]), np.zeros(n),
                    np.zeros(n), np.ones(n), np.array([0.0]), np.array([1.0]))
    srv = native.Server(max_slots=4, **TIGHT)
    with pytest.raises(RuntimeError, match=r"code 4\)"):
        srv.add_base(big)
    g = lpgen.c2_shared(seed=0xC2, m=60, n=150)
    base = srv.add_base(g.lp)
    t = srv.submit_bounds(base, g.lp.col_lower, g.lp.col_upper)
    r = _drain(srv, [t])[t][0]
    srv.close()
    assert r["status"] == 0, r
