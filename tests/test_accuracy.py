"""FP32 path against the FP64 model of the same rays (oracle, CPU).

The FP64 model traces the same Philox draws with double geometry and sums energies in
double.  Two parts of the FP32 flux error are checked separately:
  * accumulation: integer fixed-point sums converted once to float, so for the same hits
    the FP32 map is within 2^-24 relative of the exact sum of the per-ray energies;
  * geometry: FP32 and FP64 rays that land on neighbouring pixels (a few in 10^4) move
    single-ray amounts between pixels; total power still agrees to ~1e-6.
"""
import numpy as np
import pytest

import oracle
from onemirror_amd import fields as F


@pytest.mark.parametrize("lattice", [False, True])
def test_small_fp32_against_fp64(lattice):
    d, p = F.config_small()
    if lattice:
        p = F.with_lattice(p)
    o = oracle.OracleScene(d, p)
    scale, _ = o.flux_scale()
    a = o.trace_f32(records=True)
    b = o.trace_f64(records=True)
    e_exact = a["energy"].astype(np.float64) * 2.0 ** -31     # exact sum of FP32 ray energies
    f32 = (oracle.energy_to_float(a["energy"]) * np.float32(scale)).astype(np.float64)
    f64 = b["energy"] * float(scale)
    # accumulation: the float map against the exact sum of the same hits
    raw = oracle.energy_to_float(a["energy"]).astype(np.float64)
    m = e_exact > 0
    assert np.max(np.abs(raw[m] - e_exact[m]) / e_exact[m]) <= 2.0 ** -24
    # geometry: rays whose FP32 and FP64 pixel differ, blockers that differ
    nr = a["pixel"].size
    flips = int((a["pixel"] != b["pixel"]).sum())
    assert 0 < flips < 1e-3 * nr
    assert int((a["shade"] != b["shade"]).sum()) <= 1e-5 * nr
    assert int((a["block"] != b["block"]).sum()) <= 1e-5 * nr
    # total intercepted power
    assert abs(f32.sum() - f64.sum()) / f64.sum() < 1e-5
    # most pixels agree to 1e-5; the rest differ by whole rays
    mm = f64 > 0
    rel = np.abs(f32[mm] - f64[mm]) / f64[mm]
    assert np.median(rel) < 1e-6
    assert (rel > 1e-5).mean() < 0.15
