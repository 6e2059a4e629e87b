"""The allocation-free entry point (mpx_kkt_check_device_noalloc) against
the allocating one (mpx_kkt_check_device) through the C-ABI: the same eight
outputs bit for bit on two calls in a row, the second on a smaller LP (the
fixed scratch keeps the first call's partials), and a clear error for an LP
beyond the fixed scratch."""
import ctypes as C

import numpy as np
import pytest

from conftest import TIGHT
from lp_cases import one_sided_lp
from mipx_amd import lpgen, native
from mipx_amd._abi import MpxLp, c_double_p, c_int32_p
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def test_noalloc_matches_allocating_entry_point_twice():
    big = lpgen.c0_plain(seed=0xC0, m=2000, n=5000).lp
    small = one_sided_lp(seed=0x5D0, m=200, n=500)
    for lp in (big, small):
        r, x, y = ob.solve(lp, **TIGHT)
        rng = np.random.default_rng(lp.num_rows)
        for px, py in ((x, y), (x + 0.01 * rng.standard_normal(lp.num_cols), y)):
            want = native.kkt_check_device(lp, px, py)
            got = native.kkt_check_device(lp, px, py, entry="mpx_kkt_check_device_noalloc")
            print(lp.num_rows, got)
            assert got == want, (lp.num_rows, got, want)
            ref = ob.kkt_check(lp, px, py)
            for k, v in ref.items():
                assert got[k] == pytest.approx(v, rel=1e-9, abs=1e-14), k


def test_noalloc_rejects_lp_beyond_fixed_scratch():
    """Refused on the host before any device work: 4e6 rows and 4e6 columns
    need 31,250 blocks of partial sums, above the 26,214 the scratch holds.
    The pointers are never dereferenced."""
    L = native.lib()
    dummy = (C.c_double * 1)()
    di = C.cast(dummy, c_int32_p)
    dd = C.cast(dummy, c_double_p)
    lp = MpxLp(4_000_000, 4_000_000, 1, 0, di, di, dd, dd, dd, dd, dd, dd, 0.0)
    rc = L.mpx_kkt_check_device_noalloc(C.byref(lp), C.addressof(dummy), C.addressof(dummy),
                                        C.addressof(dummy), C.addressof(dummy),
                                        C.addressof(dummy), C.addressof(dummy), None)
    assert rc == 1
    assert "fixed scratch" in native.last_error()
