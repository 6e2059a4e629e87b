"""Autotuner of the packed iteration kernels' lane groups (tiles and rows per
wave) for batches of independent LPs.

For each LP-shape bucket (b_c, b_r) = (round(log2 mean entries per column),
round(log2 mean entries per row)) it builds a seeded training batch of LPs of
that shape, times the packed primal + dual kernels of every preset (lanes per
column, lanes per row) in launch order on the batch's stream
(mpx_batch_profile, every LP active), and writes the fastest preset per
bucket into mipx_amd/csrc/packed_presets.h, which the library compiles in.

The choice is a function of the LP's shape only, so a batched or served LP's
arithmetic never depends on timing or on what else runs with it.  Run on the
target GPU, then rebuild:

    python -m mipx_amd.autotune --write      # measure and rewrite the table
    python -c "import __graft_entry__ as g; g.build()"
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(PKG, "csrc", "packed_presets.h")
PRESETS = [(1, 1), (2, 2), (2, 4), (4, 4), (4, 8), (8, 8), (8, 16), (16, 16)]
COL_BUCKETS, ROW_BUCKETS = 6, 8
TOL = dict(primal_tol=1e-6, dual_tol=1e-6, optimality_tol=1e-6, verbose=0)


def training_batch(bc, br, count, seed=0x7E5E):
    """`count` LPs of the bucket's shape: per column U{k-2..k+2} entries
    around k = 2^bc (at least 1), n = m * 2^br / 2^bc, m log-uniform on
    [800, 8000], magnitudes 10^U(-3, 3) (the heterogeneous-batch family)."""
    from mipx_amd import lpgen
    rng = np.random.Generator(np.random.PCG64(seed + 97 * bc + br))
    k = 2 ** bc
    lo, hi = max(1, k - 2), k + 2
    ratio = 2 ** br / 2 ** bc
    lps = []
    for i in range(count):
        m = int(round(math.exp(rng.uniform(math.log(800), math.log(8000)))))
        n = max(2, int(round(m * ratio)))
        p = lpgen.params(m, n, 1, lo, min(hi, m), value_kind=1, vlo=-3.0, vhi=3.0,
                         seed=(seed << 20) + (bc << 12) + (br << 8) + i)
        lps.append(lpgen.Generated(p).lp)
    return lps


def measure(lps, launches, precision):
    from mipx_amd import native
    out = []
    for p in range(len(PRESETS)):
        b = native.ResidentBatch(lps=lps, precision=precision, packed_preset=p, **TOL)
        prof = b.profile(launches=launches)
        b.close()
        out.append(prof["primal_kernel_us"] + prof["dual_kernel_us"])
    return out


def write_table(table, notes, device):
    rows = []
    for bc in range(COL_BUCKETS):
        rows.append("    {" + ", ".join(f"{table[bc][br]:2d}" for br in range(ROW_BUCKETS)) + "},")
    text = f"""// packed_presets.h -- lanes per column / row of the packed iteration kernels
// by LP shape, as measured by the project's autotuner (mipx_amd/autotune.py,
// which rewrote this file from its measurements on {device}).
// Buckets: b_c = round(log2(mean entries per column)) in 0..5,
// b_r = round(log2(mean entries per row)) in 0..7; -1 = not tuned (the
// nearest-shape heuristic decides).  Presets 0..7 are (lanes per column,
// lanes per row) = {", ".join(f"({a},{b})" for a, b in PRESETS)}.
// Measured (primal + dual kernel microseconds per iteration, every LP of the
// training batch active, presets 0..7):
{notes}#pragma once

#include <cmath>

namespace mpx {{

constexpr int kTuneColBuckets = {COL_BUCKETS}, kTuneRowBuckets = {ROW_BUCKETS};

constexpr signed char kTunedPreset[kTuneColBuckets][kTuneRowBuckets] = {{
{chr(10).join(rows)}
}};

inline int tuneBucket(double mean, int buckets) {{
    if (!(mean > 0.0)) return 0;
    const long b = std::lround(std::log2(mean));
    return static_cast<int>(b < 0 ? 0 : (b >= buckets ? buckets - 1 : b));
}}

// The tuned preset for a matrix with nnz entries, m rows, n columns (-1: none).
inline int tunedPreset(long long nnz, int m, int n) {{
    if (m <= 0 || n <= 0) return -1;
    const int bc = tuneBucket(static_cast<double>(nnz) / n, kTuneColBuckets);
    const int br = tuneBucket(static_cast<double>(nnz) / m, kTuneRowBuckets);
    return kTunedPreset[bc][br];
}}

}}  // namespace mpx
"""
    with open(TABLE, "w") as f:
        f.write(text)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--buckets", default="2:4,3:4,3:3,2:3,3:5",
                    help="b_c:b_r pairs to tune (default: the shapes of the seeded "
                         "heterogeneous and serving families and their neighbours)")
    ap.add_argument("--lps", type=int, default=48)
    ap.add_argument("--launches", type=int, default=20)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--write", action="store_true", help="rewrite csrc/packed_presets.h")
    ap.add_argument("--json", default=None, help="also write the measurements here")
    ap.add_argument("--from-json", default=None,
                    help="rewrite the table from measurements saved by --json (no GPU work)")
    a = ap.parse_args()
    sys.path.insert(0, os.path.dirname(PKG))
    if a.from_json:
        with open(a.from_json) as f:
            saved = json.load(f)
        table = [[-1] * ROW_BUCKETS for _ in range(COL_BUCKETS)]
        notes = ""
        for spec, r in saved["buckets"].items():
            bc, br = (int(v) for v in spec.split(":"))
            table[bc][br] = r["best"]
            notes += (f"//   ({bc},{br}): " + " ".join(f"{u:.0f}" for u in r["us"]) +
                      f" -> preset {r['best']} {tuple(r['preset'])}\n")
        write_table(table, notes, saved["device"])
        return
    import torch
    device = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "unknown device"
    table = [[-1] * ROW_BUCKETS for _ in range(COL_BUCKETS)]
    notes = ""
    results = {}
    prec = 1 if a.precision == "mixed" else 0
    for spec in a.buckets.split(","):
        bc, br = (int(v) for v in spec.split(":"))
        t = time.perf_counter()
        lps = training_batch(bc, br, a.lps)
        us = measure(lps, a.launches, prec)
        best = int(np.argmin(us))
        table[bc][br] = best
        results[spec] = {"us": [round(u, 2) for u in us], "best": best, "preset": PRESETS[best],
                         "lps": len(lps), "nnz": int(sum(l.nnz for l in lps))}
        notes += (f"//   ({bc},{br}): " + " ".join(f"{u:.0f}" for u in us) +
                  f" -> preset {best} {PRESETS[best]}\n")
        print(f"[autotune] bucket ({bc},{br}): {[round(u, 1) for u in us]} -> preset {best} "
              f"{PRESETS[best]} ({time.perf_counter() - t:.1f} s)", file=sys.stderr, flush=True)
    if a.write:
        write_table(table, notes, device)
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"device": device, "precision": a.precision, "launches": a.launches,
                       "buckets": results}, f, indent=1)
    print(json.dumps(results))


if __name__ == "__main__":
    main()
