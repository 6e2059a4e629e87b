"""Seeded synthetic LP workloads (C++ generator in csrc/lpgen.cpp).

Workloads, all with 0 <= x <= 10, half equality / half <= rows and a
feasible primal (x0, b = A x0 plus <= slack) and dual (c = A' y0 + s0):

  c0_plain      m=2,000      n=5,000      5 nnz/col, values +-U[0.1, 1]
  c1_large      m=1,000,000  n=2,000,000  power-law col counts on [1, 200],
                mean 6; values +-10^U[-2, 2]
  c2_shared     m=3,000      n=8,000      6 nnz/col, integers in [-5,5]\\{0};
                per instance a B&B branching of 2-10% of the 40% integer cols
  c3_hetero     m = exp(U[ln 200, ln 5e4]), n = 2.5 m, U{4..8} nnz/col,
                values +-10^U[-3, 3]
  c4 serving    Poisson arrivals; 70% bounds variants of the c2 matrix, 30%
                fresh c3-style LPs with nnz ~ lognormal(median 5e4, sigma 1)
                clipped to [1e3, 5e6] and m = nnz / 15
"""
import ctypes as C
import math
import os

import numpy as np

from .lp import LpProblem
from ._loader import load_library


class GenParams(C.Structure):
    _fields_ = [
        ("m", C.c_int32), ("n", C.c_int32),
        ("colnnz_kind", C.c_int32), ("colnnz_lo", C.c_int32), ("colnnz_hi", C.c_int32),
        ("colnnz_mean", C.c_double),
        ("value_kind", C.c_int32), ("vlo", C.c_double), ("vhi", C.c_double),
        ("xhi", C.c_double), ("slack", C.c_double), ("eq_fraction", C.c_double),
        ("seed", C.c_uint64),
        ("free_fraction", C.c_double), ("lower_only_fraction", C.c_double),
        ("upper_only_fraction", C.c_double),
    ]


_lib = None


def _gen():
    global _lib
    if _lib is None:
        lib = load_library("libmipx_lpgen.so")
        lib.mpxgen_create.restype = C.c_void_p
        lib.mpxgen_create.argtypes = [C.POINTER(GenParams)]
        lib.mpxgen_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int64)]
        lib.mpxgen_get.argtypes = [C.c_void_p] + [C.c_void_p] * 9
        lib.mpxgen_branch_bounds.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.c_double,
                                             C.c_double, C.c_void_p, C.c_void_p]
        lib.mpxgen_destroy.argtypes = [C.c_void_p]
        _lib = lib
    return _lib


class Generated:
    """A generated LP plus the generator handle (for instance bounds)."""

    def __init__(self, params):
        lib = _gen()
        self.params = params
        self._h = lib.mpxgen_create(C.byref(params))
        if not self._h:
            raise ValueError("invalid generator parameters")
        m, n, nnz = C.c_int32(), C.c_int32(), C.c_int64()
        lib.mpxgen_dims(self._h, C.byref(m), C.byref(n), C.byref(nnz))
        m, n, nnz = m.value, n.value, nnz.value
        if nnz >= 2 ** 31:
            raise ValueError("nnz exceeds int32 indices")
        rs = np.empty(m + 1, np.int32)
        ci = np.empty(nnz, np.int32)
        va = np.empty(nnz, np.float64)
        vecs = [np.empty(k, np.float64) for k in (n, n, n, m, m, n)]
        ptrs = [a.ctypes.data for a in [rs, ci, va] + vecs]
        lib.mpxgen_get(self._h, *ptrs)
        obj, cl, cu, rl, ru, x0 = vecs
        self.x0 = x0
        self.lp = LpProblem(m, n, rs, ci, va, obj, cl, cu, rl, ru, 0, 0.0,
                            name=f"gen{params.seed:#x}")

    def branch_bounds(self, seed, int_fraction=0.4, frac_lo=0.02, frac_hi=0.10):
        n = self.lp.num_cols
        lo = np.empty(n, np.float64)
        hi = np.empty(n, np.float64)
        _gen().mpxgen_branch_bounds(self._h, C.c_uint64(seed), int_fraction, frac_lo,
                                    frac_hi, lo.ctypes.data, hi.ctypes.data)
        return lo, hi

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mpxgen_destroy(self._h)
            self._h = None


def params(m, n, colnnz_kind=0, lo=5, hi=5, mean=0.0, value_kind=0, vlo=0.1, vhi=1.0,
           seed=0, xhi=10.0, slack=1.0, eq_fraction=0.5, free=0.0, lower_only=0.0,
           upper_only=0.0):
    p = GenParams()
    p.free_fraction, p.lower_only_fraction, p.upper_only_fraction = free, lower_only, upper_only
    p.m, p.n = int(m), int(n)
    p.colnnz_kind, p.colnnz_lo, p.colnnz_hi, p.colnnz_mean = colnnz_kind, lo, hi, mean
    p.value_kind, p.vlo, p.vhi = value_kind, vlo, vhi
    p.xhi, p.slack, p.eq_fraction = xhi, slack, eq_fraction
    p.seed = seed
    return p


def c0_plain(seed=0xC0, m=2000, n=5000, per_col=5, free=0.0, lower_only=0.0, upper_only=0.0):
    """C0 family; `free`, `lower_only`, `upper_only` turn that fraction of
    the columns into free / 0 <= x / x <= 10 columns (dual kept feasible)."""
    return Generated(params(m, n, 0, per_col, per_col, value_kind=0, vlo=0.1, vhi=1.0,
                            seed=seed, free=free, lower_only=lower_only,
                            upper_only=upper_only))


def c1_large(seed=0xC1, m=1_000_000, n=2_000_000):
    return Generated(params(m, n, 2, 1, 200, 6.0, value_kind=1, vlo=-2.0, vhi=2.0,
                            seed=seed))


def c2_shared(seed=0xC2, m=3000, n=8000):
    return Generated(params(m, n, 0, 6, 6, value_kind=2, vlo=0.0, vhi=5.0, seed=seed))


def c2_instance_seed(i, base=0xC2):
    return (base << 32) + int(i)


def c3_sizes(count=1024, seed=0xC3, m_lo=200, m_hi=50_000):
    """m_i = round(exp(U(ln m_lo, ln m_hi))) from a numpy PCG stream."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.uniform(math.log(m_lo), math.log(m_hi), size=count)
    return np.rint(np.exp(u)).astype(np.int64)


def c3_instance(i, m, base=0xC3):
    m = int(m)
    return Generated(params(m, int(round(2.5 * m)), 1, 4, 8, value_kind=1, vlo=-3.0, vhi=3.0,
                            seed=(base << 32) + int(i)))


def c4_schedule(rate=200.0, duration=60.0, seed=0xC4, variant_fraction=0.7,
                nnz_median=5e4, nnz_sigma=1.0, nnz_lo=1e3, nnz_hi=5e6):
    """[(arrival seconds, is_variant, nnz)] of a Poisson stream of requests,
    from a numpy PCG stream (every request draws all three numbers)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        variant = bool(rng.uniform() < variant_fraction)
        nnz = float(np.clip(rng.lognormal(math.log(nnz_median), nnz_sigma), nnz_lo, nnz_hi))
        if t >= duration:
            return out
        out.append((t, variant, nnz))


def c4_fresh(i, nnz, base=0xC4):
    """Fresh request i: a c3-style LP with about `nnz` entries (m = nnz / 15)."""
    m = max(20, int(round(nnz / 15.0)))
    return Generated(params(m, int(round(2.5 * m)), 1, 4, 8, value_kind=1, vlo=-3.0, vhi=3.0,
                            seed=(base << 32) + int(i)))


def c4_variant_seed(i, base=0xC4):
    return (base << 32) + int(i)
