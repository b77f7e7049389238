"""Synthetic heliostat fields and the named workloads of the benchmark.

A scene description is a plain dict with the fields of ``om_scene_desc``
(include/onemirror.h); tracing parameters are a dict with the fields of ``om_params``.

World frame (main.cpp:71-73 of the reference): x = east, y = up, z = south.  The sun
direction is the direction the light travels, (-sin Z sin A, -cos Z, sin Z cos A) for
zenith Z and azimuth A measured from north through east.

Workloads (heliostat counts, sizes, sunshape, microsamples, grid and map sizes):
  small    100 heliostats 2x2 m, radial stagger over radii 40-80 m north of the tower,
           flat 8x8 m receiver at 60 m facing north, pillbox 4.65 mrad, slope 2 mrad,
           sun elevation 45 deg azimuth 180 deg, 16x16 Poisson microsamples x 64 rays,
           64^3 grid, 128x128 map, trimmed mean 5x5 (20 %).
  mid      5,000 heliostats 10x10 m, radii 100-800 m, Buie CSR 0.1 with a 2,048-bin CDF,
           slope 2.5 mrad, 32x32 Poisson microsamples x 40 rays, 128^3 grid, 256x256 map,
           seeded sun position.
  large    25,000 heliostats 12x10 m on a surround field (radii 100-1,500 m), cylinder
           receiver r = 8 m, h = 20 m centred 200 m up, 512x256 map over the unrolled
           surface, Buie CSR 0.05 (2,048-bin CDF), slope 2 mrad, ~1,000 Poisson
           microsamples x 40 rays (about 1e9 rays), 256^3 grid, seeded sun position.
  small_cylinder  100 heliostats 2x2 m around a cylinder r = 2 m, h = 4 m at 30 m
           (128x64 map): the cylinder path at test size.
  small_cone  the same field around a truncated cone (radii 1.5 m top, 2.5 m bottom, 4 m
           high, 128x64 equal-area map).
  batched  64 scenes of 500-20,000 heliostats under seeded suns (config_batched).
  two_grids  two sub-fields of 50 heliostats 2x2 m (north-west and north-east of the tower,
           radii 40-80 m) in two 64^3 grids of their own, the receiver and sun of small:
           the multi-grid scene of the .scn format.
  face2face  the single-heliostat scene of userData/face2face_shadow.scn.
  facet    the annual driver's faceted paraboloid heliostat (nine 0.47 m facets) at the
           position of the face2face heliostat, a = b = 2 sqrt(108.3 m) (focal length the
           distance to the receiver point), with the reference's default parameters
           (lattice microsamples every 0.01 m x 2048 rays: 40.7 million rays a step).
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np

PILLBOX, BUIE = 0, 1
LATTICE, POISSON = 0, 1


def sun_direction(elevation_deg: float, azimuth_deg: float) -> np.ndarray:
    """Direction the light travels for a sun at (elevation, azimuth)."""
    z = math.radians(90.0 - elevation_deg)
    a = math.radians(azimuth_deg)
    return np.array([-math.sin(z) * math.sin(a), -math.cos(z), math.sin(z) * math.cos(a)],
                    dtype=np.float32)


def radial_stagger(n: int, r_min: float, r_max: float, spacing: float,
                   sector_deg: float = 180.0, centre_deg: float = 0.0) -> np.ndarray:
    """Heliostat ground positions (x, z) in rows of constant radius.

    Rows are ``spacing * sqrt(3)/2`` apart; heliostats in a row are ``spacing`` apart along
    the arc, alternate rows offset by half a spacing.  ``centre_deg`` is the azimuth of the
    sector centre measured from north through east (0 = north of the tower).  The inner
    rows are filled first; the last row is filled from the sector centre outwards.
    Returns an (n, 2) float64 array, or fewer rows if the annulus is too small."""
    dr = spacing * math.sqrt(3.0) / 2.0
    out = []
    k = 0
    r = r_min
    half = math.radians(sector_deg) / 2.0
    c0 = math.radians(centre_deg)
    while r <= r_max + 1e-9 and len(out) < n:
        arc = 2.0 * half * r
        cnt = max(1, int(arc // spacing))
        step = 2.0 * half / cnt if sector_deg < 360.0 else 2.0 * math.pi / cnt
        shift = 0.5 * step if (k % 2) else 0.0
        if sector_deg < 360.0:
            angs = -half + (np.arange(cnt) + 0.5) * step
            angs = angs + (shift if k % 2 else 0.0)
            angs = angs[angs < half]
        else:
            angs = np.arange(cnt) * step + shift
        need = n - len(out)
        if len(angs) > need:
            order = np.argsort(np.abs(angs - (0.0 if sector_deg < 360.0 else 0.0)), kind="stable")
            angs = np.sort(angs[order[:need]])
        for a in angs:
            az = c0 + a
            # azimuth from north (-z) through east (+x)
            out.append((r * math.sin(az), -r * math.cos(az)))
        k += 1
        r += dr
    return np.array(out, dtype=np.float64)


def fit_spacing(n: int, r_min: float, r_max: float, sector_deg: float, s_min: float) -> float:
    """Largest spacing (>= s_min) that still fits n heliostats in the annulus sector."""
    lo, hi = s_min, max(s_min * 8.0, r_max)
    if len(radial_stagger(n, r_min, r_max, lo, sector_deg)) < n:
        return lo
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if len(radial_stagger(n, r_min, r_max, mid, sector_deg)) >= n:
            lo = mid
        else:
            hi = mid
    return lo


def _pow2_ceil(x: float) -> float:
    return float(2.0 ** math.ceil(math.log2(x)))


def grid_for(pos: np.ndarray, size: np.ndarray, top: float, n: int) -> dict:
    """Uniform n^3 grid over the heliostat boxes and up to ``top`` metres.

    Cell edges are powers of two, so (p - grid_pos) / interval is exact; the grid is
    centred on the field in x and z and starts at the ground (y = 0)."""
    rad = np.sqrt(np.sqrt((size.astype(np.float64) ** 2).sum(1))) / 2.0
    lo = (pos - rad[:, None]).min(0)
    hi = (pos + rad[:, None]).max(0)
    lo[1] = min(0.0, lo[1])
    hi[1] = max(top, hi[1])
    ext = (hi - lo) * 1.02 + 1.0
    inter = np.array([_pow2_ceil(e / n) for e in ext])
    span = inter * n
    centre = (hi + lo) / 2.0
    gpos = np.round((centre - span / 2.0) / inter) * inter  # multiple of the cell edge
    gpos[1] = np.floor(lo[1] / inter[1]) * inter[1]
    return {"grid_pos": gpos.astype(np.float32), "grid_size": span.astype(np.float32),
            "grid_interval": inter.astype(np.float32)}


def make_field(n: int, width: float, height: float, pivot: float, r_min: float, r_max: float,
               sector_deg: float, jitter: float, seed: int, s_min: float, fit: bool = True) -> dict:
    spacing = fit_spacing(n, r_min, r_max, sector_deg, s_min) if fit else s_min
    xz = radial_stagger(n, r_min, r_max, spacing, sector_deg)[:n]
    if len(xz) < n:
        raise ValueError("field annulus too small for the heliostat count")
    rng = np.random.Generator(np.random.Philox(seed))
    xz = xz + rng.uniform(-jitter, jitter, size=xz.shape)
    pos = np.zeros((n, 3), np.float32)
    pos[:, 0] = xz[:, 0]
    pos[:, 1] = pivot
    pos[:, 2] = xz[:, 1]
    size = np.tile(np.array([width, 0.1, height], np.float32), (n, 1))
    return {"helio_pos": pos, "helio_size": size,
            "helio_gap": np.zeros((n, 2), np.float32),
            "helio_matrix": np.ones((n, 2), np.int32)}


def config_small(seed: int = 1) -> tuple:
    """The 100-heliostat field (1,638,400 rays with a full 16x16 Poisson fill)."""
    f = make_field(100, 2.0, 2.0, 1.2, 40.0, 80.0, 180.0, 0.1, seed, s_min=3.0)
    desc = dict(f)
    desc.update(recv_type=0, recv_pos=np.array([0.0, 60.0, 0.0], np.float32),
                recv_size=np.array([8.0, 8.0, 0.0], np.float32),
                recv_normal=np.array([0.0, 0.0, -1.0], np.float32), recv_face=0,
                recv_pixel_length=np.float32(8.0 / 128.0),
                sun_dir=sun_direction(45.0, 180.0), dni=np.float32(1000.0))
    desc.update(grid_for(f["helio_pos"], f["helio_size"], 64.0, 64))
    prm = dict(seed=0x0123456789ABCDEF, scene_index=0, sunshape=PILLBOX,
               pillbox_half_angle=4.65e-3, csr=0.1, buie_bins=1024, slope_sigma=2e-3,
               micro_mode=POISSON, helio_pixel_length=2.0 / 16.0, poisson_k=8,
               rays_per_sample=64, reflectivity=0.88, smooth_radius=2, smooth_ratio=0.2)
    return desc, prm


def config_two_grids(seed: int = 7, n_each: int = 50) -> tuple:
    """Two sub-fields, each in a grid of its own (om_grid_desc): heliostats 0..n_each-1
    north-west of the tower (azimuths 270-330 deg) in grid 0, the rest north-east (30-90
    deg) in grid 1.  The grids are more than a cell apart, so neither grid bins a
    heliostat of the other."""
    desc, prm = config_small(seed)
    rng = np.random.Generator(np.random.Philox(seed))
    parts = []
    for centre in (300.0, 60.0):
        xz = radial_stagger(n_each, 40.0, 80.0, 3.0, 60.0, centre)[:n_each]
        if len(xz) < n_each:
            raise ValueError("sub-field annulus too small for the heliostat count")
        xz = xz + rng.uniform(-0.1, 0.1, size=xz.shape)
        pos = np.zeros((n_each, 3), np.float32)
        pos[:, 0] = xz[:, 0]
        pos[:, 1] = 1.2
        pos[:, 2] = xz[:, 1]
        parts.append(pos)
    pos = np.concatenate(parts)
    n = pos.shape[0]
    size = np.tile(np.array([2.0, 0.1, 2.0], np.float32), (n, 1))
    desc.update(helio_pos=pos, helio_size=size, helio_gap=np.zeros((n, 2), np.float32),
                helio_matrix=np.ones((n, 2), np.int32))
    grids = []
    for k in range(2):
        sl = slice(k * n_each, (k + 1) * n_each)
        g = grid_for(pos[sl], size[sl], 8.0, 64)
        grids.append({"pos": g["grid_pos"], "size": g["grid_size"], "interval": g["grid_interval"],
                      "first_helio": k * n_each, "n_helio": n_each})
    desc.update(grid_pos=grids[0]["pos"], grid_size=grids[0]["size"], grid_interval=grids[0]["interval"],
                grids=grids)
    return desc, prm


def facet_site() -> dict:
    """(a, b, X, Y, Z) of the faceted heliostat workload: at the face2face heliostat's
    position, focal length a^2 / 4 = its distance to the point (0, 13.01, 0)."""
    X, Y, Z = 12.0, 1.2, 107.0
    f = math.sqrt(X * X + Z * Z + (13.01 - Y) ** 2)
    a = 2.0 * math.sqrt(f)
    return {"a": a, "b": a, "X": X, "Y": Y, "Z": Z}


def config_mid_grids(seed: int = 2) -> tuple:
    """The mid field as a two-grid scene: its heliostats west of the tower (x < 0) first,
    then the east ones, each half in a 128^3 grid of its own."""
    desc, prm = config_mid(seed)
    west = desc["helio_pos"][:, 0] < 0.0
    order = np.concatenate([np.nonzero(west)[0], np.nonzero(~west)[0]])
    d = subset(desc, order)
    n0 = int(west.sum())
    grids = []
    for first, n in ((0, n0), (n0, d["helio_pos"].shape[0] - n0)):
        sl = slice(first, first + n)
        g = grid_for(d["helio_pos"][sl], d["helio_size"][sl], 170.0, 128)
        grids.append({"pos": g["grid_pos"], "size": g["grid_size"], "interval": g["grid_interval"],
                      "first_helio": first, "n_helio": n})
    d.update(grid_pos=grids[0]["pos"], grid_size=grids[0]["size"], grid_interval=grids[0]["interval"],
             grids=grids)
    return d, prm


def mid_sun(seed: int = 2) -> np.ndarray:
    rng = np.random.Generator(np.random.Philox(seed))
    elev = float(rng.uniform(30.0, 70.0))
    azim = float(rng.uniform(120.0, 240.0))
    return sun_direction(elev, azim)


def config_mid(seed: int = 2, n_helio: int = 5000) -> tuple:
    """The 5,000-heliostat field (about 2.05e8 rays)."""
    f = make_field(n_helio, 10.0, 10.0, 6.0, 100.0, 800.0, 180.0, 0.1, seed, s_min=14.5)
    desc = dict(f)
    desc.update(recv_type=0, recv_pos=np.array([0.0, 150.0, 0.0], np.float32),
                recv_size=np.array([25.6, 25.6, 0.0], np.float32),
                recv_normal=np.array([0.0, 0.0, -1.0], np.float32), recv_face=0,
                recv_pixel_length=np.float32(0.1),
                sun_dir=mid_sun(seed), dni=np.float32(1000.0))
    desc.update(grid_for(f["helio_pos"], f["helio_size"], 170.0, 128))
    prm = dict(seed=0x00C0FFEE12345678, scene_index=0, sunshape=BUIE,
               pillbox_half_angle=4.65e-3, csr=0.1, buie_bins=2048, slope_sigma=2.5e-3,
               micro_mode=POISSON, helio_pixel_length=10.0 / 32.0, poisson_k=8,
               rays_per_sample=40, reflectivity=0.88, smooth_radius=2, smooth_ratio=0.2)
    return desc, prm


def seeded_sun(seed: int, elev=(30.0, 70.0), azim=(120.0, 240.0)) -> np.ndarray:
    rng = np.random.Generator(np.random.Philox(seed))
    e = float(rng.uniform(*elev))
    a = float(rng.uniform(*azim))
    return sun_direction(e, a)


def config_large(seed: int = 5, n_helio: int = 25000) -> tuple:
    """The 25,000-heliostat surround field with a cylinder receiver (about 1e9 rays)."""
    f = make_field(n_helio, 12.0, 10.0, 6.0, 100.0, 1500.0, 360.0, 0.1, seed, s_min=16.0)
    desc = dict(f)
    desc.update(recv_type=1, recv_pos=np.array([0.0, 200.0, 0.0], np.float32),
                recv_size=np.array([8.0, 20.0, 0.0], np.float32),
                recv_normal=np.array([0.0, 0.0, -1.0], np.float32), recv_face=0,
                recv_pixel_length=np.float32(0.1), recv_res_x=512, recv_res_y=256,
                sun_dir=seeded_sun(seed), dni=np.float32(1000.0))
    desc.update(grid_for(f["helio_pos"], f["helio_size"], 220.0, 256))
    prm = dict(seed=0x0BADC0DE5EED0001, scene_index=0, sunshape=BUIE,
               pillbox_half_angle=4.65e-3, csr=0.05, buie_bins=2048, slope_sigma=2e-3,
               micro_mode=POISSON, helio_pixel_length=0.343, poisson_k=8,
               rays_per_sample=40, reflectivity=0.88, smooth_radius=2, smooth_ratio=0.2)
    return desc, prm


def config_small_cylinder(seed: int = 6) -> tuple:
    """100 heliostats around a small cylinder: the cylinder receiver at test size."""
    f = make_field(100, 2.0, 2.0, 1.2, 30.0, 60.0, 360.0, 0.1, seed, s_min=3.0)
    desc = dict(f)
    desc.update(recv_type=1, recv_pos=np.array([0.0, 30.0, 0.0], np.float32),
                recv_size=np.array([2.0, 4.0, 0.0], np.float32),
                recv_normal=np.array([0.0, 0.0, -1.0], np.float32), recv_face=0,
                recv_pixel_length=np.float32(0.0625), recv_res_x=128, recv_res_y=64,
                sun_dir=sun_direction(60.0, 180.0), dni=np.float32(1000.0))
    desc.update(grid_for(f["helio_pos"], f["helio_size"], 40.0, 64))
    prm = dict(seed=0x00000000CAFEF00D, scene_index=0, sunshape=PILLBOX,
               pillbox_half_angle=4.65e-3, csr=0.1, buie_bins=1024, slope_sigma=2e-3,
               micro_mode=POISSON, helio_pixel_length=2.0 / 16.0, poisson_k=8,
               rays_per_sample=64, reflectivity=0.88, smooth_radius=2, smooth_ratio=0.2)
    return desc, prm


def config_small_cone(seed: int = 8) -> tuple:
    """100 heliostats around a truncated cone (top radius 1.5 m, bottom 2.5 m, height 4 m,
    centred 30 m up): the cone receiver at test size, 128 x 64 equal-area pixels."""
    desc, prm = config_small_cylinder(seed)
    desc.update(recv_type=2, recv_size=np.array([1.5, 2.5, 4.0], np.float32))
    prm = dict(prm, seed=0x00000000C0DEC0DE)
    return desc, prm


def config_batched(seed: int = 3, n_scenes: int = 64, h_range=(500, 20000),
                   rays_per_helio: float = 20000.0, ray_clip=(1e7, 2e8)) -> list:
    """Batched workload: scenes with heliostat counts log-uniform in h_range, sun elevation
    U[10, 80] deg and azimuth U[90, 270] deg, rays per scene proportional to the field size
    and clipped to ray_clip (the microsample pitch follows), Buie CSR 0.1, slope 2.5 mrad,
    40 rays per microsample, 128^3 grids, flat 25.6 m receivers (256x256 maps).  Returns a
    list of (desc, params); scene i draws Philox stream i."""
    rng = np.random.Generator(np.random.Philox(seed))
    out = []
    s_min = 14.5
    for i in range(n_scenes):
        H = int(round(math.exp(rng.uniform(math.log(h_range[0]), math.log(h_range[1])))))
        elev = float(rng.uniform(10.0, 80.0))
        azim = float(rng.uniform(90.0, 270.0))
        R = min(max(H * rays_per_helio, ray_clip[0]), ray_clip[1])
        L = 40
        pitch = 10.0 / math.sqrt(R / (H * L))
        r_max = math.sqrt(100.0 ** 2 + H * s_min ** 2 * 0.866 / (math.pi / 2.0)) * 1.15 + 2 * s_min
        f = make_field(H, 10.0, 10.0, 6.0, 100.0, r_max, 180.0, 0.1, seed * 1000 + i, s_min, fit=False)
        tower = max(150.0, 0.2 * r_max)
        desc = dict(f)
        desc.update(recv_type=0, recv_pos=np.array([0.0, tower, 0.0], np.float32),
                    recv_size=np.array([25.6, 25.6, 0.0], np.float32),
                    recv_normal=np.array([0.0, 0.0, -1.0], np.float32), recv_face=0,
                    recv_pixel_length=np.float32(0.1), sun_dir=sun_direction(elev, azim),
                    dni=np.float32(1000.0))
        desc.update(grid_for(f["helio_pos"], f["helio_size"], tower + 20.0, 128))
        prm = dict(seed=0x00C0FFEE12345678, scene_index=i, sunshape=BUIE,
                   pillbox_half_angle=4.65e-3, csr=0.1, buie_bins=2048, slope_sigma=2.5e-3,
                   micro_mode=POISSON, helio_pixel_length=pitch, poisson_k=8,
                   rays_per_sample=L, reflectivity=0.88, smooth_radius=2, smooth_ratio=0.2)
        out.append((desc, prm))
    return out


def subset(desc: dict, idx) -> dict:
    """Description restricted to heliostats ``idx`` (same grid, receiver and sun)."""
    d = dict(desc)
    for k in ("helio_pos", "helio_size", "helio_gap", "helio_matrix"):
        d[k] = np.ascontiguousarray(desc[k][idx])
    return d


def with_lattice(prm: dict) -> dict:
    p = dict(prm)
    p["micro_mode"] = LATTICE
    return p
