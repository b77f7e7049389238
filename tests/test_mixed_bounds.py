"""The FP32 sampling of the mixed-precision trace (OM_PREC_MIXED) against the FP64 model,
exhaustively over every draw (CPU; no GPU needed).

The mixed trace certifies each ray's walk and receiver pixel with a bound on its direction's
deviation from the FP64 model's, 2^-19 (theta + 4 theta_s) + 2^-26 per ray (om_trace64.hip,
trace_ray_mixed), built from bounds on its FP32 sampling.  Each draw of a ray is one of 2^23
values u = (2k + 1) 2^-24, so those bounds are checked here on every one of them, with room
to spare (a factor of 4 on the slope error and the disk, 2 on the tail):
  * the slope-error deviate sqrt(2) erfinv(2u - 1) of the FP32 erfinv (the oracle's
    restatement of the device function, bit-exact with it) within 2^-21 relative of scipy's;
  * the Buie sunshape angle from the float image of the model's table (the largest bin whose
    lower edge is at or below the value, the value held below the table's top) within
    2^-21 theta + 2^-28 rad of the model's angle, on the draws below the tail (k < tail_k0,
    the model's choice);
  * the tail angle from the value's excess over int465 formed without cancellation,
    (k - tail_k0) 2^-23 A + (u(tail_k0) A - int465), and the float power law (the oracle's
    restated log and exp), within 2^-20 theta of the model's (the float log's error, scaled
    by |log x| / (gamma + 1), reaches about 2^-21); the direct float form
    (u A - int465) B + C cancels near the tail's start and misses that bound.
The model's angles (oracle buie_theta64) are pinned here too: below the tail they invert the
model's piecewise-linear CDF (a brute-force search over its bins), and the tail's power law
meets the table's last edge (4.65 mrad at int465) and reaches 9.3 mrad at u = 1.
"""
import math

import numpy as np
import pytest
from scipy import special

import oracle

K = np.arange(1 << 23, dtype=np.int64)
U = (2.0 * K + 1.0) * 2.0 ** -24
CASES = [(0.1, 2048), (0.05, 2048), (0.1, 1024)]  # circumsolar ratio, bins


def test_float_gauss_within_2m21_relative():
    g = oracle.math("gauss", U.astype(np.float32)).astype(np.float64)
    ref = math.sqrt(2.0) * special.erfinv(2.0 * U - 1.0)
    rel = np.abs(g - ref) / np.abs(ref)
    print(f"float gauss: max relative error {rel.max():.3e} over all 2^23 draws")
    assert rel.max() <= 2.0 ** -21


@pytest.mark.parametrize("csr,ng", CASES)
def test_model_angle_inverts_the_model_cdf(csr, ng):
    t = oracle.buie_table64(csr, ng)
    cx, cy, sl = t["cx"], t["cy"], t["slope"]
    assert np.all(np.diff(cy) > 0) and np.allclose(sl, np.diff(cx) / np.diff(cy), rtol=1e-14)
    th = oracle.buie_theta64(csr, ng, U)
    v = U * t["A"]
    disk = v < t["int465"]
    # a brute-force search of every value's bin (a dense comparison against all edges)
    sub = np.flatnonzero(disk)[:: 97]
    lo = (cy[None, :ng] <= v[sub, None]).sum(axis=1) - 1
    want = cx[lo] + (v[sub] - cy[lo]) * (cx[lo + 1] - cx[lo]) / (cy[lo + 1] - cy[lo])
    assert np.allclose(th[sub], want, rtol=1e-13, atol=0)
    assert np.all(np.diff(th) >= 0)  # monotone in the draw, across the tail's start
    # the power law meets the table's edge and the profile's 9.3 mrad end
    tail = lambda x: ((x - t["int465"]) * t["B"] + t["C"]) ** t["inv_gp1"]
    assert tail(t["int465"]) == pytest.approx(4.65, rel=1e-12)
    assert tail(t["A"]) == pytest.approx(9.3, rel=1e-12)
    assert th[~disk].min() >= 4.65 - 1e-9 and th[-1] <= 9.3


@pytest.mark.parametrize("csr,ng", CASES)
def test_buie_float_angle_within_bound(csr, ng):
    t = oracle.buie_table64(csr, ng)
    cx, cy, sl, int465, A = t["cx"], t["cy"], t["slope"], t["int465"], t["A"]
    th_model = oracle.buie_theta64(csr, ng, U) / 1000.0
    k0 = int(np.argmax(U * A >= int465))  # tail_k0: the model's first tail draw
    disk = K < k0
    # the disk: the float image of the table (the float value held below its top)
    fy, fs, fx = cy.astype(np.float32), sl.astype(np.float32), cx.astype(np.float32)
    vf = np.minimum(U.astype(np.float32) * np.float32(A), np.nextafter(fy[ng], np.float32(0)))
    lf = np.clip(np.searchsorted(fy, vf, side="right") - 1, 0, ng - 1)
    th_disk = (((vf - fy[lf]) * fs[lf] + fx[lf]) * np.float32(1e-3)).astype(np.float64)
    err = np.abs(th_disk - th_model)[disk]
    print(f"csr {csr}, {ng} bins: disk angle error max {err.max():.3e} rad over {k0} draws")
    assert np.all(err <= th_model[disk] * 2.0 ** -21 + 2.0 ** -28)
    # the tail: the excess over int465 from the draw's offset from tail_k0
    kt = K[~disk]
    du = np.float32(2.0 ** -23 * A)
    dv0 = np.float32(U[k0] * A - int465)
    dv = (kt - k0).astype(np.float32) * du + dv0
    x = dv * np.float32(t["B"]) + np.float32(t["C"])
    th_tail = (oracle.math("exp", oracle.math("log", x) * np.float32(t["inv_gp1"])) * np.float32(1e-3))
    et = np.abs(th_tail.astype(np.float64) - th_model[~disk])
    print(f"csr {csr}, {ng} bins: tail angle error max {et.max():.3e} rad over {kt.size} draws")
    assert np.all(et <= th_model[~disk] * 2.0 ** -20)
    # the direct float form loses the excess to cancellation
    xd = (U[~disk].astype(np.float32) * np.float32(A) - np.float32(int465)) * np.float32(t["B"]) + np.float32(t["C"])
    th_d = np.abs(xd.astype(np.float64)) ** t["inv_gp1"] / 1000.0
    assert np.any(np.abs(th_d - th_model[~disk]) > th_model[~disk] * 2.0 ** -20)
