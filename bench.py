#!/usr/bin/env python3
"""LP solves/s of batched PDLP on one MI355X.

Workload (default `c2`): 4,096 branch-and-bound node relaxations of one MIP
that share A (m=3,000, n=8,000, 6 nonzeros per column, integer coefficients
in [-5,5]\\{0}), each branching a seeded 2-10% of the 40% integer columns,
solved to 1e-6 relative KKT (primal, dual, gap) in one batched launch
sequence.  A step is one batched solve of all instances (device scaling,
power iteration, PDHG, polish, recovery) with the problem data already
resident in HBM.  Beside it the line reports the plain single-instance FP64
GPU PDLP on a sample of the same instances (the speedup denominator), p50 /
p99 per-instance completion latency, the roofline of the dominant kernel
(the primal iteration kernel, timed launch by launch with HIP events on the
batch's stream) and the reference CPU PDLP on the host cores.

  python bench.py [--gpus 1] [--steps 2] [--warmup 1] [--workload c2]
                  [--count 4096] [--precision fp64|mixed]

Workload `c4` drives the continuous-batching server instead: Poisson
arrivals (--rate per second for --duration seconds), 70% branch-and-bound
bounds variants of the c2 matrix and 30% fresh LPs with lognormal nnz
(median 5e4); it reports requests/s served and p50 / p99 latency from each
request's scheduled arrival to its completion.

  python bench.py --workload c4 [--rate 200] [--duration 60] [--slots 2048]
"""
import argparse
import json
import os
import queue
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LP solves/s at 1e-6 rel. KKT, batched vs plain PDLP; p99 latency; HBM GB/s"
PEAK_GBS = 8000.0  # MI355X HBM3E
TOL = dict(primal_tol=1e-6, dual_tol=1e-6, optimality_tol=1e-6, verbose=0)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def c2_workload(count):
    from mipx_amd import lpgen
    g = lpgen.c2_shared()
    n = g.lp.num_cols
    lo = np.empty((count, n))
    hi = np.empty((count, n))
    for i in range(count):
        lo[i], hi[i] = g.branch_bounds(lpgen.c2_instance_seed(i))
    return g, lo, hi


def c3_workload(count):
    from mipx_amd import lpgen
    sizes = lpgen.c3_sizes(count)
    return [lpgen.c3_instance(i, m).lp for i, m in enumerate(sizes)]


def packed_primal_bytes(lps, sv):
    """Algorithmic bytes of one launch of the packed primal kernel with every
    LP active: per LP x, x0, l, u read and x, x_r written, c and tau read
    (8n values), y gathered once (m values), its A' in CSR (nnz values +
    nnz indices + n+1 starts)."""
    return sum((8 * lp.num_cols + lp.num_rows) * sv + lp.nnz * (sv + 4) + (lp.num_cols + 1) * 4
               for lp in lps)


def primal_kernel_bytes(lp, slots, sv):
    """Algorithmic bytes of one launch of the primal iteration kernel for
    `slots` instances sharing A: per instance x, x0, l, u read, x and x_r
    written (6n values) and y read once (m values); shared once per launch:
    c, tau (2n values) and A' in CSR (nnz values + nnz indices + n+1 starts)."""
    m, n, nnz = lp.num_rows, lp.num_cols, lp.nnz
    return slots * (6 * n + m) * sv + 2 * n * sv + nnz * (sv + 4) + (n + 1) * 4


def refine_primal_bytes(lp, slots):
    """Algorithmic bytes of one launch of the refinement's primal kernel
    (FP32) for `slots` instances sharing A: per instance x, x0, l', u', c'
    read, x and x_r written (7n values) and y read once (m values); shared
    once per launch: tau (n values) and A' in CSR."""
    m, n, nnz = lp.num_rows, lp.num_cols, lp.nnz
    return slots * (7 * n + m) * 4 + n * 4 + nnz * (4 + 4) + (n + 1) * 4


def read_traffic(workload, precision, kernel=None):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE, WRITE_SIZE), committed under profiles/ when available."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{precision}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if "kernels" in d:
        return d["kernels"].get(kernel)
    return d.get("bytes_per_launch")


def host_info():
    """CPU model, logical CPUs usable by this process and physical cores of
    the host (lscpu), recorded beside the CPU baselines."""
    import subprocess
    info = {"nproc": len(os.sched_getaffinity(0)), "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict(l.split(":", 1) for l in out.splitlines() if ":" in l)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name")
        sockets = int(kv.get("Socket(s)", "1") or 1)
        cores = int(kv.get("Core(s) per socket", "0") or 0)
        info["physical_cores"] = sockets * cores if cores else None
    except (OSError, ValueError, subprocess.SubprocessError):
        info["model"] = info["physical_cores"] = None
    return info


def _pin(slot):
    """Pin this process to the slot-th CPU it is allowed to run on."""
    cpus = sorted(os.sched_getaffinity(0))
    os.sched_setaffinity(0, {cpus[slot % len(cpus)]})
    return cpus[slot % len(cpus)]


def cpu_baseline(lp, lo, hi, budget_s):
    """The CPU restatement of the reference's PDLP (the oracle, bit-identical
    to the reference's CPU path), single-threaded and pinned to one core, on
    instances of the same workload until `budget_s` seconds are used.  The
    instances it solves are compared with the GPU results afterwards."""
    import threading
    from oracle import binding as ob
    out = {}

    def work():
        out["cpu"] = _pin(0)
        t0 = time.perf_counter()
        done = []
        while time.perf_counter() - t0 < budget_s and len(done) < lo.shape[0]:
            i = len(done)
            r, _, _ = ob.solve(lp.with_col_bounds(lo[i], hi[i]), **TOL)
            done.append(r)
        out["el"] = time.perf_counter() - t0
        out["done"] = done
    th = threading.Thread(target=work)
    th.start()
    th.join()
    done, el = out["done"], out["el"]
    return dict(value=len(done) / el, unit="LP solves/s", cores=1, kind="port",
                sample=f"the first {len(done)} C2 instances solved back to back on one host "
                       f"core (cpu {out['cpu']}) in {el:.1f} s"), done


def _cpu_worker(args):
    """One host process of cpu_parallel, pinned to its own core: C2
    instances wid, wid + nw, ... solved back to back until budget_s seconds
    are used."""
    wid, nw, count, budget_s = args
    _pin(wid)
    from oracle import binding as ob
    from mipx_amd import lpgen
    g = lpgen.c2_shared()
    t0 = time.perf_counter()
    done = 0
    i = wid
    while time.perf_counter() - t0 < budget_s and i < count:
        lo, hi = g.branch_bounds(lpgen.c2_instance_seed(i))
        ob.solve(g.lp.with_col_bounds(lo, hi), **TOL)
        done += 1
        i += nw
    return done, time.perf_counter() - t0


def cpu_parallel(count, budget_s, workers):
    """The same CPU PDLP on `workers` host processes at once, each pinned to
    its own core (one LP each, as the reference would run independent node
    LPs): aggregate solves/s of the host's share of cores, beside the
    single-core cpu_baseline."""
    import multiprocessing as mp
    workers = max(1, min(workers, len(os.sched_getaffinity(0))))
    with mp.get_context("spawn").Pool(workers) as pool:
        res = pool.map(_cpu_worker, [(w, workers, count, budget_s) for w in range(workers)])
    done = sum(d for d, _ in res)
    el = max(e for _, e in res)
    return dict(value=done / el, unit="LP solves/s", cores=workers, kind="port",
                sample=f"{done} C2 instances on {workers} host processes, one pinned core "
                       f"each, in {el:.1f} s")


def copy_ceiling(repeats=5):
    """Best-of rate of a plain 1 GiB device-to-device streaming copy (read +
    write bytes / kernel time): the measured bandwidth ceiling beside the
    8 TB/s nominal peak."""
    from mipx_amd import native
    best = 0.0
    for _ in range(repeats):
        nb = native.copy_bench(1 << 30)
        best = max(best, 2.0 * nb / (native.copy_bench_ms * 1e-3) / 1e9)
    return round(best, 1)


def iter_bytes(m, n, nnz, sv):
    """Algorithmic bytes of one PDHG iteration of one LP:
    2 nnz (s_v + 4) + (m + n + 2) 4 + 9 n s_v + 7 m s_v."""
    return 2 * nnz * (sv + 4) + (m + n + 2) * 4 + 9 * n * sv + 7 * m * sv


def whole_solve_bytes_shared(lp, res, rounds, N):
    """Algorithmic bytes of a whole shared-A batch solve: per instance and
    iteration (7 n + 4 m) s_v at the precision it ran in (plain FP32
    iterations at 4 bytes, FP64 at 8), the refinement's FP32 iterations
    (8 n + 7 m) 4 (its objective and row data are per instance: c', rl', ru',
    q), plus per batched iteration the matrix in both orientations and the
    shared vectors (c, tau, sigma, rl, ru) once, counted at FP32 width (a
    lower bound)."""
    m, n, nnz = lp.num_rows, lp.num_cols, lp.nnz
    per = lambda sv: (7 * n + 4 * m) * sv
    ref = (8 * n + 7 * m) * 4
    inst = sum((r["iterations_fp32"] - r.get("iterations_refine", 0)) * per(4)
               + r.get("iterations_refine", 0) * ref
               + (r["iterations"] - r["iterations_fp32"]) * per(8) for r in res)
    shared = rounds * N * (2 * nnz * 8 + (m + n + 2) * 4 + (2 * n + 3 * m) * 4)
    return inst + shared


def run_c4(a, prec):
    """Serving benchmark: Poisson arrivals into mpx_server_*; a producer
    thread generates requests ahead of their arrival times."""
    from mipx_amd import lpgen, native
    sched = lpgen.c4_schedule(a.rate, a.duration)
    g = lpgen.c2_shared()
    srv = native.Server(hbm_budget_bytes=a.budget_gb * 1e9, max_slots=a.slots,
                        precision=prec, max_iter=a.max_iter, **TOL)
    base = srv.add_base(g.lp)
    q = queue.Queue(maxsize=512)

    def produce():
        for i, (t, variant, nnz) in enumerate(sched):
            if variant:
                q.put((i, t, None, g.branch_bounds(lpgen.c4_variant_seed(i))))
            else:
                q.put((i, t, lpgen.c4_fresh(i, nnz).lp, None))
        q.put(None)

    threading.Thread(target=produce, daemon=True).start()
    info = {}
    lat = {"variant": [], "fresh": []}
    status = [0] * 7
    iters = []
    worst_kkt = 0.0
    pending, exhausted = None, False
    late = 0.0
    t_submit = t_result = 0.0
    t0 = time.perf_counter()
    last_log = t0
    while True:
        now = time.perf_counter() - t0
        while not exhausted:
            if pending is None:
                try:
                    pending = q.get_nowait()
                except queue.Empty:
                    break
                if pending is None:
                    exhausted = True
                    break
            i, t_arr, lp, bnd = pending
            if t_arr > now:
                break
            late = max(late, now - t_arr)
            ts = time.perf_counter()
            tk = srv.submit(lp) if lp is not None else srv.submit_bounds(base, *bnd)
            t_submit += time.perf_counter() - ts
            info[tk] = (t_arr, "fresh" if lp is not None else "variant")
            pending = None
        active = srv.step()
        t_done = time.perf_counter() - t0
        ts = time.perf_counter()
        for tk in srv.poll():
            r = srv.result(tk, want_vectors=False)[0]
            t_arr, kind = info.pop(tk)
            lat[kind].append(1000.0 * (t_done - t_arr))
            status[r["status"]] += 1
            iters.append(r["iterations"])
            if r["status"] == 0:
                worst_kkt = max(worst_kkt, r["rel_primal"], r["rel_dual"], r["rel_gap"])
        t_result += time.perf_counter() - ts
        if time.perf_counter() - last_log > 5.0:
            last_log = time.perf_counter()
            st = srv.stats()
            log(f"t={t_done:.1f}s submitted {st['submitted']} finished {st['finished']} "
                f"queued {st['queued']} preparing {st['preparing']} running {st['running']} "
                f"rounds {st['rounds']} iterate {st['iterate_ms']:.0f} admit {st['admit_ms']:.0f} "
                f"activate {st['activate_ms']:.0f} retire {st['retire_ms']:.0f} ms, "
                f"max step {st['max_step_ms']:.0f} ms, submit {1000 * t_submit:.0f} ms")
        if exhausted and pending is None and active == 0 and not info:
            break
        if t_done > a.duration + a.drain_seconds:
            log("drain limit reached")
            break
        if active == 0:
            time.sleep(0.0005)
    elapsed = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    # Plain sequential single-LP FP64 GPU PDLP on a sample of the same stream.
    sample = np.linspace(0, len(sched) - 1, a.plain_sample).astype(int).tolist()
    tp = time.perf_counter()
    for i in sample:
        _, variant, nnz = sched[i]
        one = (g.lp.with_col_bounds(*g.branch_bounds(lpgen.c4_variant_seed(i))) if variant
               else lpgen.c4_fresh(i, nnz).lp)
        native.solve(one, max_iter=a.max_iter, **TOL)
    plain = len(sample) / (time.perf_counter() - tp)
    log(f"plain fp64 sequential: {plain:.2f} requests/s")
    all_lat = lat["variant"] + lat["fresh"]
    pc = lambda v, p: round(float(np.percentile(v, p)), 2) if v else None
    done = sum(status)
    return {
        "metric": "LP requests/s served at 1e-6 rel. KKT (Poisson arrivals); p50/p99 latency",
        "value": round(done / elapsed, 3),
        "unit": "requests/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(1000.0 * elapsed, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if prec == 1 else "f64",
        "data": "synthetic (seeded generator: mipx_amd/csrc/lpgen.cpp, Poisson stream seed 0xC4)",
        "config": {"workload": f"c4: Poisson {a.rate:g} requests/s for {a.duration:g} s, 70% "
                               "bounds variants of the c2 matrix (m=3000 n=8000), 30% fresh LPs "
                               "nnz~lognormal(5e4, 1) in [1e3, 5e6], m=nnz/15, n=2.5m",
                   "offered_requests": len(sched), "precision": a.precision, "slots": a.slots,
                   "hbm_budget_gb": a.budget_gb},
        "offered_rate": round(len(sched) / a.duration, 2),
        "plain_fp64_gpu_requests_per_s": round(plain, 3),
        "speedup_vs_plain": round(done / elapsed / plain, 2),
        "finished": done,
        "unfinished": len(sched) - done,
        "latency_ms": {"p50": pc(all_lat, 50), "p99": pc(all_lat, 99), "max": pc(all_lat, 100),
                       "variant_p50": pc(lat["variant"], 50), "variant_p99": pc(lat["variant"], 99),
                       "fresh_p50": pc(lat["fresh"], 50), "fresh_p99": pc(lat["fresh"], 99)},
        "status_counts": status,
        "mean_iterations": float(np.mean(iters)) if iters else None,
        "max_rel_kkt": worst_kkt,
        "max_submit_lag_s": round(late, 3),
        "server": {"rounds": st["rounds"], "iterate_ms": round(st["iterate_ms"], 1),
                   "admit_ms": round(st["admit_ms"], 1), "activate_ms": round(st["activate_ms"], 1),
                   "retire_ms": round(st["retire_ms"], 1), "max_step_ms": round(st["max_step_ms"], 1),
                   "setup_query_ms": round(st["setup_query_ms"], 1),
                   "setup_wait_ms": round(st["setup_wait_ms"], 1),
                   "host_submit_ms": round(1000 * t_submit, 1),
                   "host_result_ms": round(1000 * t_result, 1),
                   "pool_peak_gb": round(st["pool_peak_bytes"] / 1e9, 3)},
    }


def cpu_baseline_lps(lps, budget_s):
    """As cpu_baseline, on the LPs of a heterogeneous batch in order."""
    import threading
    from oracle import binding as ob
    out = {}

    def work():
        out["cpu"] = _pin(0)
        t0 = time.perf_counter()
        done = []
        while time.perf_counter() - t0 < budget_s and len(done) < len(lps):
            r, _, _ = ob.solve(lps[len(done)], **TOL)
            done.append(r)
        out["el"] = time.perf_counter() - t0
        out["done"] = done
    th = threading.Thread(target=work)
    th.start()
    th.join()
    done, el = out["done"], out["el"]
    return dict(value=len(done) / el, unit="LP solves/s", cores=1, kind="port",
                sample=f"the first {len(done)} C3 LPs solved back to back on one host core "
                       f"(cpu {out['cpu']}) in {el:.1f} s"), done


def run_c1(a):
    """One large LP (m=1e6, n=2e6, power-law columns): the single-LP path in
    FP64 and in mixed precision (FP32 iterations, FP64 checks, refinement to
    1e-8), each for a fixed number of PDHG iterations; reports iterations/s
    and the algorithmic HBM rate of an iteration."""
    from mipx_amd import lpgen, native
    g = lpgen.c1_large()
    lp = g.lp
    m, n, nnz = lp.num_rows, lp.num_cols, lp.nnz
    log(f"c1: m={m} n={n} nnz={nnz}")
    runs = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        opts = dict(TOL, max_iter=a.c1_iters, precision=prec)
        if prec == 1:
            opts["refine_tol"] = 1e-8
        native.solve(lp, **dict(opts, max_iter=200))  # warm-up: code load, graph build
        t = time.perf_counter()
        r, _, _ = native.solve(lp, **opts)
        el = time.perf_counter() - t
        it = r["iterations"]
        sv = 4 if (prec == 1 and r["iterations_fp32"] == it) else 8
        b_iter = 2 * nnz * (sv + 4) + (m + n + 2) * 4 + 9 * n * sv + 7 * m * sv
        per = (el - r["setup_seconds"]) / max(1, it)
        runs[name] = {"iterations": it, "iterations_fp32": r["iterations_fp32"],
                      "status": r["status"], "rel_primal": r["rel_primal"],
                      "rel_dual": r["rel_dual"], "rel_gap": r["rel_gap"],
                      "setup_s": round(r["setup_seconds"], 3), "wall_s": round(el, 3),
                      "us_per_iteration": round(1e6 * per, 2),
                      "iterations_per_s": round(1.0 / per, 1),
                      "algorithmic_GBs": round(b_iter / per / 1e9, 1),
                      "bytes_per_iteration": b_iter}
        log(f"{name}: {runs[name]}")
    best = runs["mixed"]
    return {
        "metric": "PDHG iterations/s on one large LP (single-LP path); HBM GB/s",
        "value": best["iterations_per_s"], "unit": "iterations/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": round(best["wall_s"] * 1000, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic (seeded generator: mipx_amd/csrc/lpgen.cpp, seed 0xC1)",
        "config": {"workload": f"c1: m={m} n={n} nnz={nnz}, power-law columns 1..200 mean 6, "
                               f"{a.c1_iters} PDHG iterations per run"},
        "roofline": {"bound": "hbm", "achieved": best["algorithmic_GBs"], "peak": PEAK_GBS,
                     "unit": "GB/s", "frac": round(best["algorithmic_GBs"] / PEAK_GBS, 4),
                     "traffic": None, "kernel": "one PDHG iteration (primal + dual fused SpMV steps)"},
        "runs": runs,
    }


def run_c1_solve(a, prec):
    """C1 solved to 1e-6 relative KKT in one precision (mixed: FP32
    iterations, FP64 checks, then FP32 refinement around an FP64 base, to
    --c1-refine-tol when given), under a wall-clock limit; the library's
    progress lines (every 2,000 iterations) go to stdout.  The returned
    point is certified on the device (mpx_kkt_check_device) in original
    space; like the reference, a solve stopped by a limit returns no point."""
    from mipx_amd import lpgen, native
    g = lpgen.c1_large()
    lp = g.lp
    m, n, nnz = lp.num_rows, lp.num_cols, lp.nnz
    log(f"c1 solve: m={m} n={n} nnz={nnz} precision={'mixed' if prec else 'fp64'} "
        f"limit {a.c1_solve:.0f} s")
    opts = dict(TOL, precision=prec, max_iter=4000000, max_solve_seconds=a.c1_solve, verbose=1)
    if prec == 1 and a.c1_refine_tol > 0:
        opts["refine_tol"] = a.c1_refine_tol
    t = time.perf_counter()
    r, x, y = native.solve(lp, **opts)
    el = time.perf_counter() - t
    sys.stdout.flush()
    k = native.kkt_check_device(lp, x, y)
    names = {0: "Optimal", 1: "Infeasible", 2: "Unbounded", 3: "IterLimit", 4: "TimeLimit"}
    out = {"status": names.get(r["status"], str(r["status"])), "iterations": r["iterations"],
           "iterations_fp32": r["iterations_fp32"],
           "iterations_refine": r.get("iterations_refine", 0),
           "objective": r["objective"], "rel_primal": r["rel_primal"], "rel_dual": r["rel_dual"],
           "rel_gap": r["rel_gap"], "setup_s": round(r["setup_seconds"], 3),
           "wall_s": round(el, 2),
           "us_per_iteration": round(1e6 * (el - r["setup_seconds"]) / max(1, r["iterations"]), 1),
           "device_kkt": {kk: k[kk] for kk in ("rel_primal", "rel_dual", "rel_gap",
                                                 "bound_viol", "dual_sign_viol")}}
    log(f"c1 solve: {out}")
    return {
        "metric": "seconds to solve one large LP to 1e-6 relative KKT (single-LP path)",
        "value": out["wall_s"], "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(el * 1000, 1), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if prec else "f64",
        "data": "synthetic (seeded generator: mipx_amd/csrc/lpgen.cpp, seed 0xC1)",
        "config": {"workload": f"c1: m={m} n={n} nnz={nnz}, power-law columns 1..200 mean 6, "
                               f"solved to 1e-6 rel KKT" +
                               (f", refined to {a.c1_refine_tol:g}" if prec and a.c1_refine_tol > 0
                                else "")},
        "solve": out,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4"])
    ap.add_argument("--c1-iters", type=int, default=4000)
    ap.add_argument("--c1-solve", type=float, default=0.0,
                    help="c1: solve to tolerance in --precision under this wall-clock limit "
                         "(seconds) instead of timing --c1-iters iterations")
    ap.add_argument("--c1-refine-tol", type=float, default=0.0,
                    help="c1 solve, mixed precision: refine to this tolerance (0: the solve's)")
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--slots", type=int, default=2048)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--drain-seconds", type=float, default=120.0)
    ap.add_argument("--max-iter", type=int, default=400000)
    ap.add_argument("--count", type=int, default=None,
                    help="instances (default: 4096 for c2, 1024 for c3)")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--handover", type=float, default=0.0,
                    help="mixed precision: FP32 -> FP64 handover KKT ratio (0: library default)")
    ap.add_argument("--refine", type=int, default=1, choices=[0, 1],
                    help="mixed precision: FP32 refinement around an FP64 base after the plain "
                         "FP32 phase (1, default) or FP64 iterations (0)")
    ap.add_argument("--spmv-group", type=int, default=0,
                    help="lanes per column / row of the packed kernels (0: per-LP choice)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="host processes of the parallel CPU figure (C2)")
    ap.add_argument("--profile-only", action="store_true",
                    help="upload the workload and time --launches launches of the iteration "
                         "kernels one by one (for rocprofv3), then a 1 GiB device copy")
    ap.add_argument("--launches", type=int, default=20)
    ap.add_argument("--profile-refine", action="store_true",
                    help="with --profile-only (mixed precision): time the refinement's "
                         "correction kernels instead, every instance moved into that phase")
    a = ap.parse_args()
    if a.count is None:
        a.count = 1024 if a.workload == "c3" else 4096
    if a.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path here "
              "(batched and served instances would shard across GPUs with no exchange)",
              file=sys.stderr)
        sys.exit(2)

    from mipx_amd import native
    prec = 1 if a.precision == "mixed" else 0
    if a.workload == "c4":
        print(json.dumps(run_c4(a, prec)), flush=True)
        return
    if a.workload == "c1":
        print(json.dumps(run_c1_solve(a, prec) if a.c1_solve > 0 else run_c1(a)), flush=True)
        return
    log(f"building workload {a.workload} count={a.count}")
    lps = None
    # --profile-only: eleven rounds cover the columns-per-wave autotuning
    # (two passes over five candidates) that precedes the profiled launches.
    cap = dict(max_iter=11 * 200) if a.profile_only else {}
    if a.handover > 0:
        cap["mixed_handover"] = a.handover
    if a.spmv_group > 0:
        cap["spmv_group"] = a.spmv_group
    cap["mixed_refine"] = a.refine
    if a.workload == "c3":
        lps = c3_workload(a.count)
        lp = lps[0]
        log(f"c3: {len(lps)} LPs, sum nnz {sum(x.nnz for x in lps)}")
        batch = native.ResidentBatch(lps=lps, precision=prec, **TOL, **cap)
    else:
        g, lo, hi = c2_workload(a.count)
        lp = g.lp
        log("uploading (resident batch)")
        batch = native.ResidentBatch(lp, lo, hi, precision=prec, **TOL, **cap)
    if a.profile_only:
        # A capped solve first: it autotunes the lane-per-instance kernels'
        # columns per wave, which the profiled launches then use (as in the
        # bench).  Packed batches have nothing to tune.
        if lps is None:
            batch.run()
        prof = batch.profile(launches=a.launches, refine=a.profile_refine)
        sv = 4 if prec == 1 else 8
        bl = (packed_primal_bytes(lps, sv) if lps is not None
              else refine_primal_bytes(lp, prof["profiled_slots"]) if a.profile_refine
              else primal_kernel_bytes(lp, prof["profiled_slots"], sv))
        print(json.dumps({"columns_per_wave": prof["columns_per_wave"],
                          "rows_per_wave": prof["rows_per_wave"],
                          "primal_kernel_us": prof["primal_kernel_us"],
                          "dual_kernel_us": prof["dual_kernel_us"],
                          "slots": prof["profiled_slots"], "launches": prof["profiled_launches"],
                          "algorithmic_bytes_per_launch": bl,
                          "achieved_GBs": bl / (prof["primal_kernel_us"] * 1e-6) / 1e9}), flush=True)
        batch.close()
        copied = native.copy_bench(1 << 30)
        log(f"device copy of {copied} bytes done (counter calibration)")
        return
    for w in range(a.warmup):
        t = time.perf_counter()
        batch.run()
        log(f"warmup {w}: {time.perf_counter() - t:.2f} s")
    native.synchronize()
    t0 = time.perf_counter()
    for s in range(a.steps):
        batch.run()
        log(f"step {s} done at {time.perf_counter() - t0:.2f} s")
    native.synchronize()
    elapsed = time.perf_counter() - t0
    st = batch.stats()
    want_vec = not a.no_cpu
    res, X, Y = batch.results(want_vectors=want_vec)
    status = np.bincount([r["status"] for r in res], minlength=7).tolist()
    iters = np.array([r["iterations"] for r in res])
    kkt = max(max(r["rel_primal"], r["rel_dual"], r["rel_gap"]) for r in res if r["status"] == 0)
    value = a.count * a.steps / elapsed
    log(f"batched: {value:.1f} solves/s, status counts {status}, mean iters {iters.mean():.0f}")

    # Whole-solve algorithmic bytes per second (one step = one full solve).
    N = 200
    if lps is not None:
        ws_bytes = sum(r["iterations_fp32"] * iter_bytes(l.num_rows, l.num_cols, l.nnz, 4) +
                       (r["iterations"] - r["iterations_fp32"]) *
                       iter_bytes(l.num_rows, l.num_cols, l.nnz, 8) for l, r in zip(lps, res))
    else:
        ws_bytes = whole_solve_bytes_shared(lp, res, st["rounds"], N)
    ws_gbs = ws_bytes / (elapsed / a.steps) / 1e9

    prof = batch.profile(launches=50)
    sv = 4 if prec == 1 else 8
    bytes_launch = (packed_primal_bytes(lps, sv) if lps is not None
                    else primal_kernel_bytes(lp, prof["profiled_slots"], sv))
    achieved = bytes_launch / (prof["primal_kernel_us"] * 1e-6) / 1e9
    primal_kernel = (None if lps is not None else
                     f"kILPrimal<{'float' if prec == 1 else 'double'},{prof['columns_per_wave']}>")
    log(f"primal kernel {prof['primal_kernel_us']:.1f} us for {prof['profiled_slots']} instances "
        f"-> {achieved:.0f} GB/s")
    batch.close()
    kernels = {primal_kernel or "kPackedPrimal": {
        "us": round(prof["primal_kernel_us"], 2), "GBs": round(achieved, 1),
        "frac": round(achieved / PEAK_GBS, 4)}}
    mean_plain = float(np.mean([r["iterations_fp32"] - r["iterations_refine"] for r in res]))
    mean_ref = float(np.mean([r["iterations_refine"] for r in res]))
    mean_64 = float(np.mean([r["iterations"] - r["iterations_fp32"] for r in res]))
    # Kernel time of a whole solve, kernel by kernel: mean iterations in the
    # kernel's phase x its launch time at full occupancy (an upper bound).
    # (packed batches run the refinement in the same FP32 packed kernel)
    first = (mean_64 if prec == 0 else mean_plain + (mean_ref if lps is not None else 0.0))
    shares = {primal_kernel or "kPackedPrimal": first * prof["primal_kernel_us"]}
    if prec == 1 and lps is None and a.refine:
        b2 = native.ResidentBatch(lp, lo, hi, precision=1, **TOL)
        pr = b2.profile(launches=50, refine=True)
        b2.close()
        byr = refine_primal_bytes(lp, pr["profiled_slots"])
        gr = byr / (pr["primal_kernel_us"] * 1e-6) / 1e9
        kernels["kILPrimal<float,1,true> (refinement)"] = {
            "us": round(pr["primal_kernel_us"], 2), "GBs": round(gr, 1),
            "frac": round(gr / PEAK_GBS, 4), "dual_us": round(pr["dual_kernel_us"], 2),
            "algorithmic_bytes_per_launch": int(byr), "instances_in_launch": pr["profiled_slots"]}
        shares["kILPrimal<float,1,true> (refinement)"] = mean_ref * pr["primal_kernel_us"]
        log(f"refinement primal kernel {pr['primal_kernel_us']:.1f} us -> {gr:.0f} GB/s")
    if prec == 1 and lps is None:
        # The FP64 primal kernel of the same instances, timed the same way
        # (an FP64 batch, every instance active).
        b64 = native.ResidentBatch(lp, lo, hi, precision=0, max_iter=11 * N, **TOL)
        b64.run()
        p64 = b64.profile(launches=50)
        b64.close()
        by64 = primal_kernel_bytes(lp, p64["profiled_slots"], 8)
        g64 = by64 / (p64["primal_kernel_us"] * 1e-6) / 1e9
        kernels[f"kILPrimal<double,{p64['columns_per_wave']}>"] = {
            "us": round(p64["primal_kernel_us"], 2), "GBs": round(g64, 1),
            "frac": round(g64 / PEAK_GBS, 4), "dual_us": round(p64["dual_kernel_us"], 2)}
        shares[f"kILPrimal<double,{p64['columns_per_wave']}>"] = mean_64 * p64["primal_kernel_us"]
        log(f"fp64 primal kernel {p64['primal_kernel_us']:.1f} us -> {g64:.0f} GB/s")
    # The roofline line's kernel: the primal kernel with the largest share of
    # a solve's kernel time.
    dom_key = max(shares, key=shares.get)
    if dom_key in kernels and "algorithmic_bytes_per_launch" in kernels[dom_key]:
        kd = kernels[dom_key]
        dom_gbs, dom_us = kd["GBs"], kd["us"]
        dom_bytes, dom_slots = kd["algorithmic_bytes_per_launch"], kd["instances_in_launch"]
    elif dom_key.startswith("kILPrimal<double"):
        kd = kernels[dom_key]
        dom_gbs, dom_us = kd["GBs"], kd["us"]
        dom_bytes, dom_slots = primal_kernel_bytes(lp, a.count, 8), a.count
    else:
        dom_key = primal_kernel or "kPackedPrimal"
        dom_gbs, dom_us, dom_bytes, dom_slots = (achieved, prof["primal_kernel_us"], bytes_launch,
                                                 prof["profiled_slots"])
    log(f"dominant primal kernel: {dom_key} ({shares})")
    ceiling = copy_ceiling()
    log(f"streaming copy ceiling {ceiling} GB/s")

    # Plain single-instance FP64 GPU PDLP on a sample of the same instances.
    t = time.perf_counter()
    plain_iters = []
    # C3 sizes are log-uniform: sample across the batch, not its head.
    sample = (np.linspace(0, len(lps) - 1, a.plain_sample).astype(int).tolist() if lps is not None
              else list(range(a.plain_sample)))
    for i in sample:
        one = lps[i] if lps is not None else lp.with_col_bounds(lo[i], hi[i])
        r, _, _ = native.solve(one, **TOL)
        plain_iters.append(r["iterations"])
    plain = a.plain_sample / (time.perf_counter() - t)
    log(f"plain fp64 single-instance: {plain:.2f} solves/s")

    cpu = cpu_par = vs_oracle = None
    if not a.no_cpu:
        cpu, cpu_res = (cpu_baseline_lps(lps, a.cpu_seconds) if lps is not None
                        else cpu_baseline(lp, lo, hi, a.cpu_seconds))
        cpu["host"] = host_info()
        log(f"cpu baseline: {cpu}")
        # Each GPU result of the instances the CPU leg solved, against the
        # oracle's FP64 solve of the same LP and the long-double certificate.
        from oracle import binding as ob
        errs, certs = [], []
        for i, ro in enumerate(cpu_res):
            one = lps[i] if lps is not None else lp.with_col_bounds(lo[i], hi[i])
            if ro["status"] == 0 and res[i]["status"] == 0:
                errs.append(abs(res[i]["objective"] - ro["objective"]) / (1 + abs(ro["objective"])))
                k = ob.kkt_check(one, X[i], Y[i])
                certs.append(max(k["rel_primal"], k["rel_dual"], k["rel_gap"]))
        vs_oracle = {"instances": len(errs),
                     "max_rel_objective_error": max(errs) if errs else None,
                     "max_certified_rel_kkt": max(certs) if certs else None}
        log(f"vs oracle: {vs_oracle}")
        if lps is None and a.cpu_workers > 1:
            cpu_par = cpu_parallel(a.count, a.cpu_seconds, a.cpu_workers)
            log(f"cpu parallel: {cpu_par}")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "LP solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1000.0 * elapsed / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if prec == 1 else "f64",
        "data": "synthetic (seeded generator: mipx_amd/csrc/lpgen.cpp, base seed "
                + ("0xC3)" if lps is not None else "0xC2)"),
        "config": {"workload": ("c3: 1024 independent LPs, m log-uniform on [200, 5e4], n=2.5m, "
                                "4-8 nnz per column, tol 1e-6 rel KKT" if lps is not None else
                                "c2: 4096 B&B node relaxations sharing A, m=3000 n=8000 "
                                "nnz=48000, tol 1e-6 rel KKT"),
                   "instances": a.count, "precision": a.precision,
                   "kernels": "lane-per-instance" if st["interleaved"] else "packed"},
        "roofline": {"bound": "hbm", "achieved": round(dom_gbs, 1), "peak": PEAK_GBS,
                     "unit": "GB/s", "frac": round(dom_gbs / PEAK_GBS, 4),
                     "traffic": read_traffic(a.workload, a.precision, dom_key.split(" ")[0]),
                     "kernel": (("kPackedPrimal (primal half step, packed tiles)" if lps is not None
                                 else f"{dom_key} (primal half step, lane-per-instance)")
                                + ": the primal kernel with the largest share of a solve's "
                                  "kernel time"),
                     "kernel_us": round(dom_us, 2),
                     "algorithmic_bytes_per_launch": int(dom_bytes),
                     "instances_in_launch": dom_slots,
                     "kernels": kernels,
                     "whole_solve": {"achieved": round(ws_gbs, 1), "frac": round(ws_gbs / PEAK_GBS, 4),
                                     "algorithmic_bytes": int(ws_bytes),
                                     "definition": "sum over instances of iterations x bytes "
                                                   "per iteration at the precision run / wall "
                                                   "time of one solve"},
                     "copy_ceiling_GBs": ceiling},
        "cpu_baseline": cpu,
        "cpu_parallel": cpu_par,
        "vs_oracle": vs_oracle,
        "plain_fp64_gpu_solves_per_s": round(plain, 3),
        "speedup_vs_plain": round(value / plain, 2),
        "latency_ms": {"p50": round(st["latency_p50_ms"], 2), "p99": round(st["latency_p99_ms"], 2),
                       "max": round(st["latency_max_ms"], 2)},
        "status_counts": status,
        "mean_iterations": float(iters.mean()),
        "mean_fp32_iterations": float(np.mean([r["iterations_fp32"] for r in res])),
        "mean_iterations_by_phase": {"fp32": round(mean_plain, 1), "fp32_refinement": round(mean_ref, 1),
                                     "fp64": round(mean_64, 1)},
        "primal_kernel_time_per_solve_us": {k: round(v, 1) for k, v in shares.items()},
        "plain_mean_iterations": float(np.mean(plain_iters)),
        "max_rel_kkt": kkt,
        "setup_ms": round(st["setup_ms"], 2),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
