"""Builds the native libraries in-tree.

  libmipx_pdlp.so   HIP library for gfx950 (hipcc), the C-ABI of include/mipx_pdlp.h
  libmipx_lpgen.so  host-only synthetic LP generator (g++)

Run `python -m mipx_amd.build`.  Device code is compiled with
-ffp-contract=off (reference-order arithmetic) and with packed FP32
instructions disabled.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
ROOT = os.path.dirname(PKG)

HIP_SOURCES = ["pdlp_solver.hip", "pdlp_batch.hip"]
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build_hip(jobs=8):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [hipcc] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
        if out.strip():
            print(out)
        if p.returncode != 0:
            raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o",
          os.path.join(PKG, "libmipx_pdlp.so")] + objs)
    for o in objs:
        os.remove(o)


def build_lpgen():
    _run(["g++", "-std=c++17", "-O3", "-fPIC", "-shared", "-Wall", "-o",
          os.path.join(PKG, "libmipx_lpgen.so"), os.path.join(CSRC, "lpgen.cpp")])


def build_all():
    build_lpgen()
    build_hip()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
