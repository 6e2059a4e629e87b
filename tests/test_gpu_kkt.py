"""The device KKT certificate (mpx_kkt_check_device) against the oracle's
long-double certificate of the same (x, y), through the C-ABI."""
import numpy as np
import pytest

from conftest import TIGHT
from lp_cases import closed_form_cases, one_sided_lp
from mipx_amd import lpgen, native
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def _close(dev, ref, tag):
    for k, v in ref.items():
        assert dev[k] == pytest.approx(v, rel=1e-9, abs=1e-14), (tag, k, dev[k], v)


@pytest.mark.parametrize("kind", ["c0", "one_sided", "maximize"])
def test_kkt_device_matches_oracle_certificate(kind):
    """At a solver's optimum, at planted errors (dual sign flip, perturbed
    primal, scaled dual) and at random points the eight fields agree with the
    long-double certificate to 1e-9."""
    lp = (one_sided_lp(seed=0x5D0, m=200, n=500) if kind == "one_sided"
          else lpgen.c0_plain(seed=77, m=150, n=400).lp)
    if kind == "maximize":
        lp.sense = 1
        lp.obj = -lp.obj
    r, x, y = ob.solve(lp, **TIGHT)
    assert r["status"] == 0
    rng = np.random.default_rng(5)
    xb = x.copy()
    xb[0] += 1.0
    points = [(x, y), (x, -y), (xb, y), (x, 1.01 * y),
              (rng.uniform(-1, 11, lp.num_cols), rng.standard_normal(lp.num_rows))]
    for t, (px, py) in enumerate(points):
        ref = ob.kkt_check(lp, px, py)
        dev = native.kkt_check_device(lp, px, py)
        _close(dev, ref, (kind, t))
    good = native.kkt_check_device(lp, x, y)
    assert good["rel_primal"] <= 1.001e-6 and good["rel_gap"] <= 1.001e-6


def test_kkt_device_closed_forms_and_empty_rows():
    """The reference's closed-form LPs (tests/test_pdlp.cpp) at the oracle's
    answers (optima, Farkas rays, primal rays), including LPs with no rows
    and with infinite bounds on either side."""
    for c, lp in closed_form_cases():
        if lp.num_cols == 0:
            continue
        opts = dict(verbose=0, max_iter=200000)
        opts.update(c.get("options", {}))
        _, x, y = ob.solve(lp, **opts)
        ref = ob.kkt_check(lp, x, y)
        dev = native.kkt_check_device(lp, x, y)
        _close(dev, ref, c["name"])
