"""Annual time-step driver: the reference's main.cpp:56-116 loop over the days and hours of a
year, run as batches of scene forks on the GPU.

The reference steps month 1..12, day 1, 4, ..., 28 and hour 6, 8, ..., 18 (main.cpp:59-61),
computes the sun position with SPA, skips zenith angles above 75 degrees (main.cpp:68),
sets the sun direction (main.cpp:71-73) and the clear-sky DNI of getDNI (SPA.cpp:1174-1189),
builds the scene and traces it (main.cpp:100-106).  Here each time step is a fork of one
field (shared grid, microsamples and tiles; re-aimed heliostats) and the steps are traced by
the batched launch.  SPA itself is not part of this library: the caller supplies the sun
angles, or ``solar_position`` gives an approximate textbook position (NOAA's general solar
position equations: Spencer's declination and equation of time) for synthetic schedules.
"""
from __future__ import annotations

import math
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import (BATCH_PER_SCENE, BATCH_WINDOW, PREC_FP32, PREC_FP64, Batch, SolarScene, dni_hottel, facet_scene, lib,
               sun_direction, _f32p)

# main.cpp:34-44: the reference's site (Hangzhou) and year
REFERENCE_SITE = {"latitude": 30.3, "longitude": 120.2, "timezone": 8.0, "elevation_m": 40.0}
ZENITH_MAX = 75.0  # main.cpp:68

_MONTH_DAYS = (31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31)


def reference_schedule() -> List[tuple]:
    """(month, day, hour) of main.cpp:59-61: 12 months x days 1, 4, ..., 28 x hours 6..18
    every 2 hours (840 steps)."""
    return [(m, d, h) for m in range(1, 13) for d in range(1, 29, 3) for h in range(6, 19, 2)]


def day_of_year(month: int, day: int) -> int:
    return sum(_MONTH_DAYS[: month - 1]) + day


def solar_position(latitude: float, longitude: float, timezone: float, month: int, day: int,
                   hour: float) -> tuple:
    """Approximate (zenith, azimuth) in degrees, azimuth clockwise from north, at local
    clock time ``hour`` (NOAA general solar position equations; within about 0.5 degree of
    SPA).  For synthetic schedules only."""
    n = day_of_year(month, day)
    g = 2.0 * math.pi / 365.0 * (n - 1 + (hour - 12.0) / 24.0)
    eqtime = 229.18 * (0.000075 + 0.001868 * math.cos(g) - 0.032077 * math.sin(g)
                       - 0.014615 * math.cos(2 * g) - 0.040849 * math.sin(2 * g))
    decl = (0.006918 - 0.399912 * math.cos(g) + 0.070257 * math.sin(g) - 0.006758 * math.cos(2 * g)
            + 0.000907 * math.sin(2 * g) - 0.002697 * math.cos(3 * g) + 0.00148 * math.sin(3 * g))
    tst = hour * 60.0 + eqtime + 4.0 * longitude - 60.0 * timezone   # true solar time, minutes
    ha = math.radians(tst / 4.0 - 180.0)
    lat = math.radians(latitude)
    cz = math.sin(lat) * math.sin(decl) + math.cos(lat) * math.cos(decl) * math.cos(ha)
    cz = max(-1.0, min(1.0, cz))
    z = math.acos(cz)
    sz = math.sin(z)
    if sz < 1e-9:
        return 0.0, 180.0
    ca = (math.sin(lat) * cz - math.sin(decl)) / (math.cos(lat) * sz)
    ca = max(-1.0, min(1.0, ca))
    az = 180.0 - math.degrees(math.acos(ca))   # morning: east of south
    if ha > 0:
        az = 360.0 - az                          # afternoon: west of south
    return math.degrees(z), az % 360.0


def annual_steps(schedule: Optional[Sequence[tuple]] = None, site: Optional[dict] = None,
                 angles: Optional[Sequence[tuple]] = None, zenith_max: float = ZENITH_MAX) -> List[dict]:
    """The time steps of a year that the reference traces: for every (month, day, hour) of
    ``schedule`` (default: main.cpp's) the sun angles (``angles`` if given, one (zenith,
    azimuth) per entry, else ``solar_position`` at ``site``), skipping zenith > zenith_max,
    with the sun direction and the DNI of the site's elevation."""
    schedule = list(schedule or reference_schedule())
    site = dict(REFERENCE_SITE, **(site or {}))
    if angles is not None and len(angles) != len(schedule):
        raise ValueError("one (zenith, azimuth) per schedule entry")
    steps = []
    for i, (m, d, h) in enumerate(schedule):
        if angles is not None:
            z, az = angles[i]
        else:
            z, az = solar_position(site["latitude"], site["longitude"], site["timezone"], m, d, h)
        if z > zenith_max:
            continue
        steps.append({"month": m, "day": d, "hour": h, "zenith": float(z), "azimuth": float(az),
                      "sun_dir": sun_direction(z, az), "dni": dni_hottel(z, m, d, site["elevation_m"])})
    return steps


def _pixel_area(scene: SolarScene) -> float:
    """Receiver pixel area (the Srec of Gtest.cu:40): square pixels of the rectangle, an
    azimuth x height cell of the unrolled cylinder, an equal-area cell of the cone."""
    return float(scene.info["pixel_area"])


class AnnualRun:
    """The time steps of ``steps`` (``annual_steps``) on one field, traced batch by batch.

    Each step is a fork of ``scene`` with its own Philox stream (scene index = step number),
    so its maps are what a scene built for that sun alone gives.  ``run`` returns per-step
    intercept ratios (unsmoothed and smoothed, om_intercept_ratio as results_with_given_normal
    computes them), total and intercepted power, and the annual sums.  ``AnnualRun.facets``
    is the reference driver's own run: the faceted heliostat rebuilt for every step."""

    def __init__(self, scene: Optional[SolarScene], steps: Iterable[dict], batch_size: int = 64,
                 workgroup: int = 0, rays_per_thread: int = 0, mode: str = "batch", scenes=None):
        """``mode``: "batch" traces the steps with the batched launch (hipGraph replays),
        "plain" with one om_trace_plain per step, in the accumulation mode of each scene
        (see SolarScene.autotune_accumulate); the maps are the same.  With ``workgroup`` 0
        a batch of steps is one persistent launch that keeps each step's map window in LDS
        (BATCH_WINDOW), or, for FP64-geometry scenes, which that launch does not trace, one
        window trace after another (BATCH_PER_SCENE).  ``scenes`` (one per step) replaces the
        forks of ``scene``."""
        if mode not in ("batch", "plain"):
            raise ValueError("mode is 'batch' or 'plain'")
        self.scene = scene
        self.mode = mode
        self.steps = list(steps)
        if scenes is not None:
            self.forks = list(scenes)
            if len(self.forks) != len(self.steps):
                raise ValueError("one scene per time step")
        else:
            self.forks = [scene.fork(st["sun_dir"], st["dni"], scene_index=i) for i, st in enumerate(self.steps)]
        self.batch_size = int(batch_size)
        self.workgroup, self.rays_per_thread = workgroup, rays_per_thread
        self._batches = []
        self._maps = []
        if mode == "batch":
            for b in range(0, len(self.forks), self.batch_size):
                part = self.forks[b: b + self.batch_size]
                wg = workgroup
                if wg == 0:
                    fp64 = any(f.params.get("precision", PREC_FP32) == PREC_FP64 for f in part)
                    wg = BATCH_PER_SCENE if fp64 else BATCH_WINDOW
                self._batches.append(Batch(part, workgroup=wg, rays_per_thread=rays_per_thread))
        else:
            self._maps = [f.new_maps(smooth=True) for f in self.forks]
        self.rays = sum(f.info["total_rays"] for f in self.forks)

    @classmethod
    def facets(cls, a: float, b: float, X: float, Y: float, Z: float, steps: Iterable[dict], params: dict,
               ctx=None, **kw) -> "AnnualRun":
        """main.cpp:56-116: for every time step the faceted paraboloid heliostat
        y = x^2/a^2 + z^2/b^2 at (X, Y, Z) turned to that step's master normal
        (main.cpp:84-91) and built into the reference's scene (scene_file_proc.cpp:288-534),
        with the step's sun and DNI; Philox stream = step number."""
        steps = list(steps)
        scenes = [SolarScene(facet_scene(a, b, X, Y, Z, st["sun_dir"], st["dni"]), dict(params, scene_index=i), ctx)
                  for i, st in enumerate(steps)]
        return cls(None, steps, scenes=scenes, **kw)

    def trace(self, stream=None):
        """Trace every step; the maps stay on the device."""
        if self.mode == "batch":
            for b in self._batches:
                b.run(stream)
        else:
            for f, m in zip(self.forks, self._maps):
                f.trace_plain(m, stream)

    def maps(self):
        if self.mode == "batch":
            return [m for b in self._batches for m in b.maps]
        return self._maps

    def results(self, keep_maps: bool = False, nonsmooth_path: Optional[str] = None,
                smooth_path: Optional[str] = None) -> dict:
        """Per-step and annual results from the maps of the last ``trace``.  With paths, each
        step appends its ratio lines to them, as the reference driver's steps do
        (Gtest.cu:47,59 via ImageSaver::savetxt)."""
        from . import save_ratio_txt
        import torch
        torch.cuda.synchronize()
        L = lib()
        rows = []
        flux_maps = []
        for st, f, m in zip(self.steps, self.forks, self.maps()):
            flux = m.flux.float().cpu().numpy().astype(np.float32).ravel()
            smooth = m.flux_smooth.float().cpu().numpy().astype(np.float32).ravel()
            te = float(f.info["total_energy"])
            r = float(L.om_intercept_ratio(flux.ctypes.data_as(_f32p), f.res_x, f.res_y, te))
            rs = float(L.om_intercept_ratio(smooth.ctypes.data_as(_f32p), f.res_x, f.res_y, te))
            if nonsmooth_path:
                save_ratio_txt(nonsmooth_path, flux.reshape(f.res_y, f.res_x), te)
            if smooth_path:
                save_ratio_txt(smooth_path, smooth.reshape(f.res_y, f.res_x), te)
            area = _pixel_area(f)
            rows.append({**{k: st[k] for k in ("month", "day", "hour", "zenith", "azimuth", "dni")},
                         "ratio": r, "ratio_smooth": rs, "total_w": te * area, "intercepted_w": r * te * area})
            if keep_maps:
                flux_maps.append(flux.reshape(f.res_y, f.res_x))
        out = {"steps": rows,
               "intercepted_w_sum": float(sum(r["intercepted_w"] for r in rows)),
               "total_w_sum": float(sum(r["total_w"] for r in rows))}
        out["annual_ratio"] = out["intercepted_w_sum"] / out["total_w_sum"] if out["total_w_sum"] > 0 else 0.0
        if keep_maps:
            out["flux"] = flux_maps
        return out

    def run(self, keep_maps: bool = False) -> dict:
        self.trace()
        return self.results(keep_maps)
