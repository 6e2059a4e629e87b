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
"""
import argparse
import json
import os
import sys
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


def primal_kernel_bytes(lp, slots, sv):
    """Algorithmic bytes of one launch of the primal iteration kernel for
    `slots` instances sharing A: per instance x, x0, l, u read, x and x_r
    written (6n values) and y read once (m values); shared once per launch:
    c, tau (2n values) and A' in CSR (nnz values + nnz indices + n+1 starts)."""
    m, n, nnz = lp.num_rows, lp.num_cols, lp.nnz
    return slots * (6 * n + m) * sv + 2 * n * sv + nnz * (sv + 4) + (n + 1) * 4


def read_traffic(workload, precision):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE), committed under profiles/ when available."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{precision}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def cpu_baseline(lp, lo, hi, budget_s):
    """The reference's own CPU PDLP (oracle/_ref, built from its sources) or,
    when that was not built, the oracle restatement, single-threaded, on
    instances of the same workload until `budget_s` seconds are used."""
    from oracle import binding as ob
    kind = "reference" if ob.ref_available() else "port"
    solve = ob.ref_solve if kind == "reference" else ob.solve
    t0 = time.perf_counter()
    done = 0
    while time.perf_counter() - t0 < budget_s and done < lo.shape[0]:
        r, _, _ = solve(lp.with_col_bounds(lo[done], hi[done]), **TOL)
        done += 1
    el = time.perf_counter() - t0
    return dict(value=done / el, unit="LP solves/s", cores=1, kind=kind,
                sample=f"{done} C2 instances solved back to back on one host core in {el:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c2"])
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="upload the workload and time --launches launches of the iteration "
                         "kernels one by one (for rocprofv3), then a 1 GiB device copy")
    ap.add_argument("--launches", type=int, default=20)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(instances shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from mipx_amd import native
    prec = 1 if a.precision == "mixed" else 0
    log(f"building workload {a.workload} count={a.count}")
    g, lo, hi = c2_workload(a.count)
    lp = g.lp
    log("uploading (resident batch)")
    batch = native.ResidentBatch(lp, lo, hi, precision=prec, **TOL)
    if a.profile_only:
        prof = batch.profile(launches=a.launches)
        sv = 4 if prec == 1 else 8
        bl = primal_kernel_bytes(lp, prof["profiled_slots"], sv)
        print(json.dumps({"primal_kernel_us": prof["primal_kernel_us"],
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
    res, _, _ = batch.results(want_vectors=False)
    status = np.bincount([r["status"] for r in res], minlength=7).tolist()
    iters = np.array([r["iterations"] for r in res])
    kkt = max(max(r["rel_primal"], r["rel_dual"], r["rel_gap"]) for r in res if r["status"] == 0)
    value = a.count * a.steps / elapsed
    log(f"batched: {value:.1f} solves/s, status counts {status}, mean iters {iters.mean():.0f}")

    prof = batch.profile(launches=50)
    sv = 4 if prec == 1 else 8
    bytes_launch = primal_kernel_bytes(lp, prof["profiled_slots"], sv)
    achieved = bytes_launch / (prof["primal_kernel_us"] * 1e-6) / 1e9
    log(f"primal kernel {prof['primal_kernel_us']:.1f} us for {prof['profiled_slots']} instances "
        f"-> {achieved:.0f} GB/s")
    batch.close()

    # Plain single-instance FP64 GPU PDLP on a sample of the same instances.
    t = time.perf_counter()
    plain_iters = []
    for i in range(a.plain_sample):
        r, _, _ = native.solve(lp.with_col_bounds(lo[i], hi[i]), **TOL)
        plain_iters.append(r["iterations"])
    plain = a.plain_sample / (time.perf_counter() - t)
    log(f"plain fp64 single-instance: {plain:.2f} solves/s")

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(lp, lo, hi, a.cpu_seconds)
        log(f"cpu baseline: {cpu}")

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
        "data": "synthetic (seeded generator: mipx_amd/csrc/lpgen.cpp, base seed 0xC2)",
        "config": {"workload": "c2: 4096 B&B node relaxations sharing A, m=3000 n=8000 "
                               "nnz=48000, tol 1e-6 rel KKT",
                   "instances": a.count, "precision": a.precision,
                   "kernels": "lane-per-instance" if st["interleaved"] else "packed"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_GBS, 4),
                     "traffic": read_traffic(a.workload, a.precision),
                     "kernel": "kILPrimal (primal half step, lane-per-instance)",
                     "kernel_us": round(prof["primal_kernel_us"], 2),
                     "algorithmic_bytes_per_launch": int(bytes_launch),
                     "instances_in_launch": prof["profiled_slots"]},
        "cpu_baseline": cpu,
        "plain_fp64_gpu_solves_per_s": round(plain, 3),
        "speedup_vs_plain": round(value / plain, 2),
        "latency_ms": {"p50": round(st["latency_p50_ms"], 2), "p99": round(st["latency_p99_ms"], 2),
                       "max": round(st["latency_max_ms"], 2)},
        "status_counts": status,
        "mean_iterations": float(iters.mean()),
        "plain_mean_iterations": float(np.mean(plain_iters)),
        "max_rel_kkt": kkt,
        "setup_ms": round(st["setup_ms"], 2),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
