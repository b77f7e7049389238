"""Annual time-step helpers on the host (no GPU): the product's om_sun_direction and
om_dni_hottel against the oracle's restatement of main.cpp:71-73 and SPA.cpp:1174-1189, the
restatement pinned by closed forms and invariants, the schedule of main.cpp:59-61, and the
approximate solar position used for synthetic schedules."""
import math

import numpy as np
import pytest

import oracle.annual as OA
import onemirror_amd as om
from onemirror_amd import annual as A


def test_schedule_is_the_reference_loop():
    s = A.reference_schedule()
    assert s == OA.schedule()
    assert len(s) == 12 * 10 * 7
    assert s[0] == (1, 1, 6) and s[-1] == (12, 28, 18)
    assert sorted({d for _, d, _ in s}) == [1, 4, 7, 10, 13, 16, 19, 22, 25, 28]


def test_oracle_sun_direction_closed_forms():
    # zenith 0: straight down; unit length everywhere
    assert np.allclose(OA.sun_direction(0.0, 123.0), [0.0, -1.0, 0.0])
    z, a = np.meshgrid(np.linspace(0, 89, 30), np.linspace(0, 359, 40))
    d = OA.sun_direction(z, a)
    assert np.allclose(np.linalg.norm(d, axis=-1), 1.0)
    assert np.allclose(d[..., 1], -np.cos(np.deg2rad(z)))
    # the horizontal part points away from the sun's azimuth (north = azimuth 0 = -z side,
    # east = azimuth 90 = +x side): light from the east travels west (-x)
    assert np.allclose(OA.sun_direction(90.0, 90.0), [-1.0, 0.0, 0.0], atol=1e-15)
    assert np.allclose(OA.sun_direction(90.0, 0.0), [0.0, 0.0, 1.0], atol=1e-15)
    assert np.allclose(OA.sun_direction(90.0, 180.0), [0.0, 0.0, -1.0], atol=1e-15)


def test_oracle_dni_invariants():
    # day numbers of a 365-day year
    assert OA.day_number(1, 1) == 1 and OA.day_number(3, 1) == 60 and OA.day_number(12, 31) == 365
    # extraterrestrial term: 1353 +- 45, largest on day 355, smallest half a year later,
    # mean 1353 over a year
    n = np.arange(1, 366)
    s = OA.extraterrestrial(n)
    assert abs(s.max() - 1398.0) < 1e-9 and int(n[np.argmax(s)]) == 355
    assert abs(s.min() - 1308.0) < 0.01 and int(n[np.argmin(s)]) in (172, 173)
    assert abs(s.mean() - 1353.0) < 1e-9
    # Hottel coefficients of a sea-level site: a0 = 0.4237 - 0.00821 * 36 etc.
    a0, a1, k = OA.hottel_coefficients(0.0)
    assert abs(a0 - 0.12814) < 1e-9 and abs(a1 - 0.7568875) < 1e-9 and abs(k - 0.387225) < 1e-9
    # transmittance: a0 + a1 e^-k overhead, falling to a0 at the horizon, monotone in zenith
    t0 = OA.dni_hottel(0.0, 6, 1, 0.0) / OA.extraterrestrial(OA.day_number(6, 1))
    assert abs(t0 - (a0 + a1 * math.exp(-k))) < 1e-12
    th = OA.dni_hottel(89.999, 6, 1, 0.0) / OA.extraterrestrial(OA.day_number(6, 1))
    assert abs(th - a0) < 1e-6
    z = np.linspace(0, 89, 90)
    assert np.all(np.diff(OA.dni_hottel(z, 6, 1, 0.0)) < 0)
    # a higher site sees a clearer sky
    assert OA.dni_hottel(30.0, 6, 1, 2000.0) > OA.dni_hottel(30.0, 6, 1, 0.0)


def test_product_helpers_match_the_oracle():
    rng = np.random.default_rng(5)
    for _ in range(200):
        # the C entry points take float arguments: compare at the float values
        z, a = float(np.float32(rng.uniform(0, 89))), float(np.float32(rng.uniform(0, 360)))
        got = om.sun_direction(z, a)
        want = OA.sun_direction(z, a)
        assert np.allclose(got, want, atol=2e-7), (z, a, got, want)
        m, d, e = int(rng.integers(1, 13)), int(rng.integers(1, 29)), float(np.float32(rng.uniform(0, 3000)))
        g, w = om.dni_hottel(z, m, d, e), float(OA.dni_hottel(z, m, d, e))
        assert abs(g - w) <= 2e-7 * w, (z, m, d, e, g, w)
    assert math.isnan(om.dni_hottel(10.0, 13, 1, 0.0))


def test_solar_position_textbook_cases():
    # equinox and solstices at solar noon: zenith = |latitude - declination|, sun due south
    lat = 30.3
    for (m, d), decl in [((3, 21), 0.0), ((6, 21), 23.44), ((12, 21), -23.44)]:
        # local solar noon at longitude 120 in zone 8 is about 12:00 clock time (+- equation
        # of time, at most 16 minutes): search the hour of the smallest zenith
        hours = np.linspace(11.0, 13.0, 241)
        zs = [A.solar_position(lat, 120.0, 8.0, m, d, h)[0] for h in hours]
        i = int(np.argmin(zs))
        assert abs(zs[i] - abs(lat - decl)) < 0.6, (m, d, zs[i])
        az = A.solar_position(lat, 120.0, 8.0, m, d, hours[i])[1]
        assert abs(az - 180.0) < 2.0, (m, d, az)
    # morning sun in the east, afternoon in the west
    assert 45 < A.solar_position(lat, 120.0, 8.0, 6, 21, 8.0)[1] < 135
    assert 225 < A.solar_position(lat, 120.0, 8.0, 6, 21, 16.0)[1] < 315


def test_annual_steps_skip_low_sun_and_carry_dni():
    steps = A.annual_steps()
    assert 400 < len(steps) < 840
    assert all(s["zenith"] <= A.ZENITH_MAX for s in steps)
    s = steps[len(steps) // 2]
    assert np.allclose(s["sun_dir"], OA.sun_direction(s["zenith"], s["azimuth"]), atol=2e-7)
    assert abs(s["dni"] - OA.dni_hottel(s["zenith"], s["month"], s["day"], 40.0)) < 1e-3
    # explicit angles override the approximate positions
    sched = [(6, 1, 12), (6, 1, 14), (6, 1, 16)]
    st = A.annual_steps(sched, angles=[(20.0, 180.0), (80.0, 250.0), (40.0, 260.0)])
    assert [x["hour"] for x in st] == [12, 16]
