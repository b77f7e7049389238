"""Heliostats whose shading walks are skipped (CPU).

The library marks a heliostat "clear" when no ray toward the sun from its mirror, within the
sunshape's largest angle, can meet another heliostat (a projection test with margins above
the float error of the intersection); the trace kernels then skip its shading walks, which
would all return -1.  Pinned here against the oracle, which walks every ray: no ray of a
clear heliostat is shaded in the oracle's records, on fields and suns where shading is
common, and nothing is marked when the sun is too low for the test.
"""
import numpy as np
import pytest

import oracle
import onemirror_amd as om
from onemirror_amd import fields as F


def _clear(d, p):
    return om.clear_heliostats(d, p)


def _cases():
    d, p = F.config_mid()
    yield "mid", F.subset(d, slice(2000, 2300)), p
    yield "mid low sun", dict(F.subset(d, slice(2000, 2300)), sun_dir=F.sun_direction(18.0, 200.0)), p
    yield "mid lattice", F.subset(d, slice(4000, 4200)), F.with_lattice(p)
    d, p = F.config_small()
    yield "small", d, p
    yield "small low sun", dict(d, sun_dir=F.sun_direction(12.0, 150.0)), p
    d, p = F.config_small_cylinder()
    yield "cylinder pillbox", d, p


@pytest.mark.parametrize("case", range(6))
def test_clear_heliostats_are_never_shaded(case):
    label, d, p = list(_cases())[case]
    c = _clear(d, p)
    o = oracle.OracleScene(d, p)
    rec = o.trace_f32(records=True)
    off = o._range(0, None)[2]
    shaded = np.zeros(len(c), bool)
    hs = np.repeat(np.arange(len(c)), np.diff(off))
    shaded[np.unique(hs[rec["shade"] >= 0])] = True
    print(label, "clear", int(c.sum()), "of", len(c), "shaded heliostats", int(shaded.sum()))
    assert not np.any(shaded & (c == 1)), label
    if "low sun" in label:
        assert shaded.sum() > 0  # the case has shading to miss


def test_nothing_is_cleared_with_the_sun_at_the_horizon():
    d, p = F.config_small()
    assert _clear(dict(d, sun_dir=F.sun_direction(2.0, 180.0)), p).sum() == 0


def test_most_heliostats_are_clear_on_the_bench_fields():
    d, p = F.config_mid()
    assert _clear(d, p).mean() > 0.5
