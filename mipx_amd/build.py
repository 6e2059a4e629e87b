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

HIP_SOURCES = ["pdlp_solver.hip", "pdlp_batch.hip", "pdlp_kkt.hip"]
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


def build_host():
    """C++ mirror of mipx::PdlpSolver (libmipx_host.so), the CLI and the C++
    host tests, linked against the HIP library with an $ORIGIN rpath."""
    inc = ["-I" + os.path.join(PKG, "host", "include"), "-I" + os.path.join(ROOT, "include")]
    src = os.path.join(PKG, "host", "src")
    flags = ["-std=c++20", "-O2", "-Wall", "-fPIC"]
    _run(["g++"] + flags + inc + ["-shared", "-o", os.path.join(PKG, "libmipx_host.so"),
                                  os.path.join(src, "pdlp.cpp"), os.path.join(src, "sparse_matrix.cpp"),
                                  os.path.join(src, "mps_io.cpp"),
                                  os.path.join(src, "presolve.cpp"),
                                  os.path.join(src, "presolve_c.cpp"),
                                  "-L" + PKG, "-lmipx_pdlp", "-lz", "-Wl,-rpath,$ORIGIN"])
    bind = os.path.join(PKG, "bin")
    os.makedirs(bind, exist_ok=True)
    for exe, extra in (("mipx_solve", ["-lmipx_lpgen"]), ("host_tests", []), ("io_tests", ["-lz"])):
        _run(["g++"] + flags + inc + ["-o", os.path.join(bind, exe), os.path.join(src, exe + ".cpp"),
                                      "-L" + PKG, "-lmipx_host", "-lmipx_pdlp"] + extra +
             ["-Wl,-rpath,$ORIGIN/.."])


def build_all():
    build_lpgen()
    build_hip()
    build_host()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
