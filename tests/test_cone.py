"""The truncated-cone receiver (CPU): the oracle's FP32 intersection pinned by the FP64
quadratic, by equal-area rows and by the cylinder it becomes for equal radii; the library's
receiver build against the oracle.

The reference's CircularTruncatedConeReceiver is an empty stub (receiver.cuh:125-140, its
intersection always misses); this receiver is the project's.  Axis vertical through the
receiver position at mid-height, radius r0 + k y at height y in [0, h] above the base (size
= top radius, bottom radius, height: the order of topradius_bottomradius_hight_), outer
surface only; rows hold equal lateral areas.
"""
import ctypes

import numpy as np
import pytest

import oracle
import onemirror_amd as om
from onemirror_amd import fields as F

BASE = np.float32([3.0, 190.0, -2.0])


def _params(r1, r0, h):
    r0, r1, h = np.float32(r0), np.float32(r1), np.float32(h)
    k = np.float32((r1 - r0) / h)
    inv = np.float32(np.float32(1.0) / np.float32(h * (r0 + r1)))
    return r0, k, h, inv, r1


def _cone_f64(o, d, base, r0, r1, h, rx, ry):
    """Nearest root ahead within [0, h] of |p_xz|^2 = (r0 + k p_y)^2 in double, entering the
    surface; pixel with equal-area rows.  Also how close to a pixel edge, the silhouette or
    a rim (in pixels, or relative for the silhouette)."""
    o = o.astype(np.float64) - base.astype(np.float64)
    d = d.astype(np.float64)
    k = (r1 - r0) / h
    w0, wd = r0 + k * o[:, 1], k * d[:, 1]
    a = d[:, 0] ** 2 + d[:, 2] ** 2 - wd ** 2
    b = 2 * (o[:, 0] * d[:, 0] + o[:, 2] * d[:, 2] - w0 * wd)
    c = o[:, 0] ** 2 + o[:, 2] ** 2 - w0 ** 2
    disc = b * b - 4 * a * c
    sq = np.sqrt(np.maximum(disc, 0))
    roots = np.sort(np.stack([(-b - sq) / (2 * a), (-b + sq) / (2 * a)], 1), 1)
    n = o.shape[0]
    t = np.full(n, -1.0)
    for j in (1, 0):
        y = o[:, 1] + roots[:, j] * d[:, 1]
        ok = (disc >= 0) & (roots[:, j] >= 1e-6) & (y >= 0) & (y <= h)
        t = np.where(ok, roots[:, j], t)
    hp = o + t[:, None] * d
    enter = d[:, 0] * hp[:, 0] + d[:, 2] * hp[:, 2] - d[:, 1] * k * (r0 + k * hp[:, 1]) < 0
    ok = (t >= 0) & enter
    th = np.mod(np.arctan2(hp[:, 2], hp[:, 0]), 2 * np.pi)
    cu = th / (2 * np.pi) * rx
    rv = hp[:, 1] * (2 * r0 + k * hp[:, 1]) / (h * (r0 + r1)) * ry
    px = np.where(ok, np.minimum(rv.astype(np.int64), ry - 1) * rx + np.minimum(cu.astype(np.int64), rx - 1), -1)
    edge = np.minimum(np.abs(cu - np.round(cu)), np.abs(rv - np.round(rv)))
    edge = np.minimum(edge, np.abs(disc) / np.maximum(b * b, 1e-30) * 1e3)
    for j in (0, 1):  # either root at a rim (the top or bottom edge of the surface)
        y = o[:, 1] + roots[:, j] * d[:, 1]
        edge = np.minimum(edge, np.minimum(np.abs(y), np.abs(y - h)) / h * ry)
    return px, edge


def _rays(n, seed, r_aim=10.0):
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, 2 * np.pi, n)
    rad = rng.uniform(50, 1500, n)
    o = np.stack([BASE[0] + rad * np.cos(ang), rng.uniform(0, 10, n), BASE[2] + rad * np.sin(ang)], 1)
    aim = np.stack([BASE[0] + rng.uniform(-r_aim, r_aim, n), BASE[1] + rng.uniform(-5, 25, n),
                    BASE[2] + rng.uniform(-r_aim, r_aim, n)], 1)
    d = aim - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    return o.astype(np.float32), d.astype(np.float32)


@pytest.mark.parametrize("r1,r0", [(6.0, 9.0), (9.0, 6.0), (0.0, 8.0), (8.0, 0.5)])
def test_cone_pixel_against_quadratic(r1, r0):
    """FP32 pixels agree with the FP64 quadratic except within FP32 rounding of an edge."""
    h, rx, ry = 20.0, 512, 256
    r0f, k, hf, inv, r1f = _params(r1, r0, h)
    o, d = _rays(200000, 8)
    got = oracle.cone_pixel(o, d, BASE, r0f, k, hf, inv, rx, ry)
    want, edge = _cone_f64(o, d, BASE, float(r0f), float(r1f), float(hf), rx, ry)
    assert (want >= 0).mean() > 0.2 and (want < 0).mean() > 0.1
    bad = got != want
    assert bad.mean() < 1e-3
    assert np.all(edge[bad] < 1e-2), edge[bad].max()
    # entering side: where the surface faces up (r1 <= r0), rays from below enter it against
    # the horizontal part of its normal, so the hit faces the ray's origin
    if r1 <= r0:
        hit = got >= 0
        col = (got[hit] % rx + 0.5) / rx * 2 * np.pi
        side = np.stack([np.cos(col), np.sin(col)], 1)
        toward = np.stack([o[hit, 0] - BASE[0], o[hit, 2] - BASE[2]], 1)
        assert np.all((side * toward).sum(1) > 0)


def test_rows_hold_equal_areas():
    """Points spread uniformly over the lateral surface, each hit head-on from outside,
    fill every row and column alike (chi-square), and land in the row of their area
    fraction."""
    r1, r0, h, rx, ry = 3.0, 7.0, 12.0, 64, 32
    r0f, k, hf, inv, r1f = _params(r1, r0, h)
    rng = np.random.default_rng(3)
    n = 400000
    u = rng.uniform(0, 1, n)
    kk = float(k)
    # y (2 r0 + k y) = u h (r0 + r1): the height below which the area fraction is u
    y = (-2 * r0 + np.sqrt(4 * r0 * r0 + 4 * kk * u * h * (r0 + r1))) / (2 * kk)
    phi = rng.uniform(0, 2 * np.pi, n)
    rad = r0 + kk * y
    p = np.stack([rad * np.cos(phi), y, rad * np.sin(phi)], 1)
    nout = np.stack([np.cos(phi), -kk * np.ones(n), np.sin(phi)], 1)
    nout /= np.linalg.norm(nout, axis=1)[:, None]
    o = (BASE + p + 5.0 * nout).astype(np.float32)
    d = (-nout).astype(np.float32)
    got = oracle.cone_pixel(o, d, BASE, r0f, k, hf, inv, rx, ry)
    assert (got >= 0).all()
    rows, cols = got // rx, got % rx
    exp = n / ry
    chi = ((np.bincount(rows, minlength=ry) - exp) ** 2 / exp).sum()
    assert chi < 80, chi                    # 31 degrees of freedom
    exp = n / rx
    chi = ((np.bincount(cols, minlength=rx) - exp) ** 2 / exp).sum()
    assert chi < 130, chi                   # 63 degrees of freedom
    want = np.minimum((u * ry).astype(np.int64), ry - 1)
    assert (rows != want).mean() < 1e-3


def test_equal_radii_give_the_cylinder():
    h, rx, ry, r = 20.0, 512, 256, 8.0
    r0f, k, hf, inv, _ = _params(r, r, h)
    assert k == 0
    o, d = _rays(200000, 9)
    cone = oracle.cone_pixel(o, d, BASE, r0f, k, hf, inv, rx, ry)
    cyl = oracle.cylinder_pixel(o, d, BASE, np.float32(r) * np.float32(r), h, rx, ry)
    assert (cone >= 0).mean() > 0.2
    assert (cone != cyl).mean() < 1e-4


def test_inside_and_through_the_open_top_miss():
    r0f, k, hf, inv, _ = _params(4.0, 6.0, 10.0)
    # from the axis outwards: inside the cone
    n = 1000
    ang = np.linspace(0, 2 * np.pi, n, endpoint=False)
    o = np.tile(BASE + np.float32([0, 5, 0]), (n, 1)).astype(np.float32)
    d = np.stack([np.cos(ang), np.zeros(n), np.sin(ang)], 1).astype(np.float32)
    assert (oracle.cone_pixel(o, d, BASE, r0f, k, hf, inv, 64, 32) == -1).all()
    # steeply down through the open top onto the inner wall
    o = np.tile(BASE + np.float32([0, 30, 0]), (n, 1)).astype(np.float32)
    d = np.stack([0.3 * np.cos(ang), -np.ones(n), 0.3 * np.sin(ang)], 1)
    d = (d / np.linalg.norm(d, axis=1)[:, None]).astype(np.float32)
    assert (oracle.cone_pixel(o, d, BASE, r0f, k, hf, inv, 64, 32) == -1).all()


def test_library_cone_receiver_matches_the_oracle():
    d, p = F.config_small_cone()
    keep = []
    dd = om._desc_struct(d, keep)
    pp = om._params_struct(F.with_lattice(p))
    h = ctypes.c_void_p()
    assert om.lib().om_scene_create(None, ctypes.byref(dd), ctypes.byref(pp), ctypes.byref(h)) == 0, \
        om.lib().om_last_error()
    try:
        info = om.SceneInfo()
        om.lib().om_scene_get_info(h, ctypes.byref(info))
        o = oracle.OracleScene(d, F.with_lattice(p))
        assert (info.res_x, info.res_y) == (o.res_x, o.res_y) == (128, 64)
        scale, total = o.flux_scale()
        assert np.float32(info.flux_scale) == np.float32(scale)
        assert np.float32(info.total_energy) == np.float32(total)
        r1, r0, hh = 1.5, 2.5, 4.0
        area = np.pi * (r0 + r1) * np.hypot(hh, r1 - r0) / (128 * 64)
        assert abs(info.pixel_area - area) < 1e-6 * area
        n = np.zeros(300, np.float32)
        om.lib().om_scene_export(h, b"helio_normal", n.ctypes.data, n.nbytes)
        assert np.array_equal(n.reshape(-1, 3), o.helio()["normal"])
    finally:
        om.lib().om_scene_destroy(h)
    # default resolution from the pixel length: mean circumference and slant height
    d2 = dict(d, recv_res_x=0, recv_res_y=0, recv_pixel_length=np.float32(0.05))
    o2 = oracle.OracleScene(d2, F.with_lattice(p))
    assert (o2.res_x, o2.res_y) == (int(np.float32(3.14159265) * np.float32(4.0) * np.float32(20)),
                                    int(np.float32(np.sqrt(np.float32(17.0))) * np.float32(20)))
    # radii and height are checked
    for bad in ([1.0, -1.0, 4.0], [0.0, 0.0, 4.0], [1.0, 2.0, 0.0]):
        dd = om._desc_struct(dict(d, recv_size=np.float32(bad)), keep)
        assert om.lib().om_scene_create(None, ctypes.byref(dd), ctypes.byref(pp), ctypes.byref(h)) == 1
