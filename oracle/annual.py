"""Oracle of the annual time-step helpers -- TEST INFRASTRUCTURE ONLY (see the package
docstring for who may import it).

Restated in numpy float64 from the reference:
  sun_direction   main.cpp:71-73   (-sin z sin az, -cos z, sin z cos az), z, az in degrees
  dni_hottel      SPA.cpp:1174-1189 getDNI: (1353 + 45 cos(2 pi (n + 10) / 365)) *
                  (a0 + a1 exp(-k / sin(90 deg - z))), n = day of a 365-day year,
                  a0 = 0.4237 - 0.00821 (6 - A)^2, a1 = 0.5055 + 0.00595 (6.5 - A)^2,
                  k = 0.2711 + 0.01858 (2.5 - A)^2, A = site elevation in km (Hottel 1976)
  schedule        main.cpp:59-61   months 1..12, days 1, 4, ..., 28, hours 6, 8, ..., 18
The tests pin these by closed forms and invariants (unit direction, zenith and azimuth
limits, the extraterrestrial extremes, the horizon limit of the transmittance, day counts).
"""
from __future__ import annotations

import numpy as np

MONTH_DAYS = (31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31)


def sun_direction(zenith_deg, azimuth_deg) -> np.ndarray:
    z = np.deg2rad(np.asarray(zenith_deg, np.float64))
    a = np.deg2rad(np.asarray(azimuth_deg, np.float64))
    return np.stack([-np.sin(z) * np.sin(a), -np.cos(z), np.sin(z) * np.cos(a)], axis=-1)


def day_number(month: int, day: int) -> int:
    return int(sum(MONTH_DAYS[: month - 1]) + day)


def hottel_coefficients(elevation_m: float):
    A = elevation_m / 1000.0
    return (0.4237 - 0.00821 * (6.0 - A) ** 2, 0.5055 + 0.00595 * (6.5 - A) ** 2,
            0.2711 + 0.01858 * (2.5 - A) ** 2)


def extraterrestrial(n) -> np.ndarray:
    return 1353.0 + 45.0 * np.cos(2.0 * np.pi * (np.asarray(n, np.float64) + 10.0) / 365.0)


def dni_hottel(zenith_deg, month: int, day: int, elevation_m: float):
    a0, a1, k = hottel_coefficients(elevation_m)
    s = extraterrestrial(day_number(month, day))
    z = np.deg2rad(np.asarray(zenith_deg, np.float64))
    return s * (a0 + a1 * np.exp(-k / np.sin(np.pi / 2 - z)))


def schedule():
    return [(m, d, h) for m in range(1, 13) for d in range(1, 29, 3) for h in range(6, 19, 2)]
