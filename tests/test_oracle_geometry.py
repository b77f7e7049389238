"""Pins of the oracle's geometry and sampling against closed forms, brute force and
invariants (the reference ships no tests or golden outputs)."""
import numpy as np
import pytest
from scipy import integrate, special, stats

import oracle
from onemirror_amd import fields as F


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _angle(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.arctan2(np.linalg.norm(np.cross(a, b)), np.dot(a, b))


# --------------------------------------------------------------------------------------
# frames, rotation, intersection
# --------------------------------------------------------------------------------------
def test_local2world_frame():
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = _unit(rng.normal(size=3)).astype(np.float32)
        if abs(n[0]) < 1e-3 and abs(n[2]) < 1e-3:
            continue
        ey = oracle.local2world([0, 1, 0], n)
        ex = oracle.local2world([1, 0, 0], n)
        ez = oracle.local2world([0, 0, 1], n)
        assert np.allclose(ey, n, atol=1e-7)
        M = np.stack([ex, ey, ez]).astype(np.float64)
        assert np.allclose(M @ M.T, np.eye(3), atol=3e-7)
        # u = (0,1,0) x n points along +x rotated about y: horizontal
        assert abs(ex[1]) < 1e-7
    # identity frame near +-Y (global_function.cuh:29-30)
    d = np.float32([0.3, 0.4, 0.5])
    assert np.array_equal(oracle.local2world(d, [0, 1, 0]), d)
    assert np.array_equal(oracle.local2world(d, [5e-7, -1, 0]), d)


def test_rotate_y():
    o = np.float32([1.0, 2.0, 3.0])
    # (0,0,1) -> (1,0,0): quarter turn about y
    r = oracle.rotate_y(o, [0, 0, 1], [1, 0, 0])
    assert np.allclose(r, [3.0, 2.0, -1.0], atol=1e-6)
    r = oracle.rotate_y(o, [0, 0, 1], [0, 0, -1])
    assert np.allclose(r, [-1.0, 2.0, -3.0], atol=1e-6)
    r = oracle.rotate_y(o, [0, 0, 1], [0, 0, 1])
    assert np.allclose(r, o)


def test_parallelogram_constructed_hits():
    rng = np.random.default_rng(1)
    for _ in range(500):
        A = rng.uniform(-50, 50, 3)
        E1 = rng.normal(size=3) * 3
        E2 = rng.normal(size=3) * 3
        u, v = rng.uniform(0.05, 0.95, 2)
        t = rng.uniform(1, 100)
        d = _unit(rng.normal(size=3))
        P = A + u * E1 + v * E2
        o = P - t * d
        ok, tuv = oracle.ray_parallelogram(o, d, A, A + E1, A + E2)
        n = np.cross(E1, E2)
        if abs(np.dot(n, d)) / np.linalg.norm(n) < 1e-2:
            continue
        assert ok
        assert np.allclose(tuv, [t, u, v], rtol=2e-4, atol=2e-4)
    # misses outside [0,1], behind the origin, parallel
    A, B, C = np.zeros(3), np.float32([1, 0, 0]), np.float32([0, 0, 1])
    assert not oracle.ray_parallelogram([1.5, 1, 0.5], [0, -1, 0], A, B, C)[0]
    assert not oracle.ray_parallelogram([0.5, -1, 0.5], [0, -1, 0], A, B, C)[0]
    assert not oracle.ray_parallelogram([0.5, 1, 0.5], [1, 0, 0], A, B, C)[0]
    ok, tuv = oracle.ray_parallelogram([1, 1, 1], [0, -1, 0], A, B, C)  # corner u = v = 1
    assert ok and tuv[1] == 1 and tuv[2] == 1


# --------------------------------------------------------------------------------------
# DDA against the exact cell sequence
# --------------------------------------------------------------------------------------
def _empty_scene(gpos, ginter, gn):
    d = {
        "helio_pos": np.zeros((0, 3), np.float32), "helio_size": np.zeros((0, 3), np.float32),
        "helio_gap": np.zeros((0, 2), np.float32), "helio_matrix": np.zeros((0, 2), np.int32),
        "recv_type": 0, "recv_pos": [0, 50, 0], "recv_size": [8, 8, 0], "recv_normal": [0, 0, -1],
        "recv_face": 0, "recv_pixel_length": 0.0625, "grid_pos": gpos,
        "grid_size": (np.asarray(ginter) * np.asarray(gn)).astype(np.float32),
        "grid_interval": ginter, "sun_dir": [0, -1, -1], "dni": 1000.0,
    }
    p = dict(seed=1, sunshape=0, slope_sigma=0.0, micro_mode=0, helio_pixel_length=0.1,
             rays_per_sample=1)
    return oracle.OracleScene(d, p)


def _exact_cells(o, d, gpos, ginter, gn):
    """Cells entered after the start cell, by exact (double) crossing times; also the
    smallest relative gap between consecutive distinct crossing times seen."""
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    g = np.trunc((o - gpos) / ginter).astype(int)
    cells, gaps = [], []
    tnext = np.empty(3)
    step = np.where(d >= 0, 1, -1)
    for a in range(3):
        if abs(d[a]) <= 1e-6:
            tnext[a] = np.inf
        else:
            b = gpos[a] + (g[a] + (1 if d[a] >= 0 else 0)) * ginter[a]
            tnext[a] = (b - o[a]) / d[a]
    while True:
        srt = np.sort(tnext)
        gaps.append((srt[1] - srt[0]) / max(abs(srt[0]), 1e-9))
        a = int(np.argmin(tnext))
        g[a] += step[a]
        if g[a] < 0 or g[a] >= gn[a]:
            return cells, min(gaps)
        cells.append(int(g[0] * gn[1] * gn[2] + g[1] * gn[2] + g[2]))
        tnext[a] += ginter[a] / abs(d[a])


def test_atan2_against_libm():
    """The project's atan2 (cylinder azimuth) against double-precision atan2, over all
    octants, the axes and tiny/huge ratios."""
    rng = np.random.default_rng(7)
    y = np.concatenate([rng.normal(size=200000), [0.0, 1.0, -1.0, 0.0, 1e-30, -1e-30, 3.0]])
    x = np.concatenate([rng.normal(size=200000), [1.0, 0.0, 0.0, -1.0, -2.0, 5.0, 1e-30]])
    y, x = y.astype(np.float32), x.astype(np.float32)
    pairs = np.empty(2 * y.size, np.float32)
    pairs[0::2], pairs[1::2] = y, x
    got = oracle.math("atan2", pairs).astype(np.float64)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(got - want)) < 5e-7
    assert got[-4] == np.float32(np.pi) and got[-6] == np.float32(np.pi / 2) and got[-7] == 0.0


def _cylinder_pixel_f64(o, d, base, r, h, res_x, res_y):
    """Entering root of the ray-cylinder quadratic in double, and the unrolled pixel; also
    returns how close the hit is to a pixel edge (in pixels) or to the silhouette."""
    o = o.astype(np.float64) - base.astype(np.float64)
    d = d.astype(np.float64)
    a = d[:, 0] ** 2 + d[:, 2] ** 2
    b = 2 * (o[:, 0] * d[:, 0] + o[:, 2] * d[:, 2])
    c = o[:, 0] ** 2 + o[:, 2] ** 2 - r * r
    disc = b * b - 4 * a * c
    ok = (a >= 1e-6) & (disc > 0)
    t = np.where(ok, (-b - np.sqrt(np.maximum(disc, 0))) / (2 * a), -1.0)
    ok &= t >= 1e-6
    hp = o + t[:, None] * d
    ok &= (hp[:, 1] >= 0) & (hp[:, 1] <= h)
    th = np.mod(np.arctan2(hp[:, 2], hp[:, 0]), 2 * np.pi)
    cu, rv = th / (2 * np.pi) * res_x, hp[:, 1] / h * res_y
    px = np.where(ok, np.minimum(rv.astype(np.int64), res_y - 1) * res_x +
                  np.minimum(cu.astype(np.int64), res_x - 1), -1)
    edge = np.minimum(np.minimum(np.abs(cu - np.round(cu)), np.abs(rv - np.round(rv))),
                      np.abs(disc) / np.maximum(b * b, 1e-30) * 1e3)
    edge = np.minimum(edge, np.minimum(np.abs(hp[:, 1]), np.abs(hp[:, 1] - h)) / h * res_y)
    return px, edge


def test_cylinder_pixel_against_quadratic():
    """Cylinder receiver pixels of the FP32 restatement agree with the FP64 quadratic,
    except for hits within FP32 rounding of a pixel edge, the silhouette or a rim."""
    rng = np.random.default_rng(8)
    n = 200000
    base, r, h, rx, ry = np.float32([3.0, 190.0, -2.0]), 8.0, 20.0, 512, 256
    ang = rng.uniform(0, 2 * np.pi, n)
    rad = rng.uniform(50, 1500, n)
    o = np.stack([base[0] + rad * np.cos(ang), rng.uniform(0, 10, n), base[2] + rad * np.sin(ang)], 1)
    aim = np.stack([base[0] + rng.uniform(-10, 10, n), base[1] + rng.uniform(-5, 25, n),
                    base[2] + rng.uniform(-10, 10, n)], 1)
    d = aim - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    o, d = o.astype(np.float32), d.astype(np.float32)
    got = oracle.cylinder_pixel(o, d, base, np.float32(r) * np.float32(r), h, rx, ry)
    want, edge = _cylinder_pixel_f64(o, d, base, r, h, rx, ry)
    assert (want >= 0).mean() > 0.3 and (want < 0).mean() > 0.1
    bad = got != want
    assert bad.mean() < 1e-3
    assert np.all(edge[bad] < 1e-2), edge[bad].max()
    # entering side: the hit faces the ray's origin
    hit = got >= 0
    col = (got[hit] % rx + 0.5) / rx * 2 * np.pi
    side = np.stack([np.cos(col), np.sin(col)], 1)
    toward = np.stack([o[hit, 0] - base[0], o[hit, 2] - base[2]], 1)
    assert np.all((side * toward).sum(1) > 0)


def test_dda_matches_exact_cell_sequence():
    gpos = np.float32([-10.0, -3.0, 5.0])
    ginter = np.float32([3.0, 1.75, 2.5])
    gn = np.array([9, 11, 7])
    s = _empty_scene(gpos, ginter, gn)
    rng = np.random.default_rng(5)
    agree = ties = 0
    for _ in range(600):
        o = (gpos + rng.uniform(0.05, 0.95, 3) * ginter * gn).astype(np.float32)
        d = _unit(rng.normal(size=3)).astype(np.float32)
        got, hit = s.dda_probe(o, d)
        assert hit == -1
        want, mingap = _exact_cells(o, d, gpos.astype(np.float64), ginter.astype(np.float64), gn)
        if list(got) == want:
            agree += 1
        else:
            assert mingap < 1e-4, (o, d, list(got), want)
            ties += 1
    assert agree >= 590


def test_dda_axis_aligned_and_tie_break():
    gpos = np.float32([0, 0, 0])
    ginter = np.float32([1, 1, 1])
    gn = np.array([4, 4, 4])
    s = _empty_scene(gpos, ginter, gn)
    # +x ray from cell (0,1,2): visits x = 1,2,3 (start cell never tested, Q1)
    got, _ = s.dda_probe([0.5, 1.5, 2.5], [1, 0, 0])
    assert list(got) == [(x * 16 + 1 * 4 + 2) for x in (1, 2, 3)]
    # exact diagonal tie tMax.x == tMax.y == tMax.z: the reference steps z first (Q4)
    got, _ = s.dda_probe([0.5, 0.5, 0.5], _unit([1, 1, 1]).astype(np.float32))
    assert got[0] == 0 * 16 + 0 * 4 + 1
    # tie between x and y with z parallel: y is taken (tMax.x >= tMax.y and tMax.y < tMax.z)
    got, _ = s.dda_probe([0.5, 0.5, 0.5], _unit([1, 1, 0]).astype(np.float32))
    assert got[0] == 0 * 16 + 1 * 4 + 0


# --------------------------------------------------------------------------------------
# grid binning: every heliostat box lands in exactly the cells it overlaps
# --------------------------------------------------------------------------------------
def test_grid_binning_matches_box_overlap():
    d, p = F.config_small()
    s = oracle.OracleScene(d, F.with_lattice(p))
    g = s.grid()
    gn = g["grid_num"]
    gpos = d["grid_pos"].astype(np.float64)
    gi = d["grid_interval"].astype(np.float64)
    cs, ci = g["cell_start"], g["cell_items"]
    owner = {}
    for c in range(len(cs) - 1):
        items = ci[cs[c]:cs[c + 1]]
        assert np.all(np.diff(items) > 0), "items of a cell are in heliostat order"
        for h in items:
            owner.setdefault(int(h), set()).add(c)
    for h in range(s.n_helio):
        pos = d["helio_pos"][h].astype(np.float64)
        r = np.sqrt(np.linalg.norm(d["helio_size"][h].astype(np.float64))) / 2  # Q6
        lo = np.floor((pos - r - gpos) / gi).astype(int)
        hi = np.floor((pos + r - gpos) / gi).astype(int)
        want = {int(x * gn[1] * gn[2] + y * gn[2] + z)
                for x in range(lo[0], hi[0] + 1) for y in range(lo[1], hi[1] + 1)
                for z in range(lo[2], hi[2] + 1)}
        assert owner.get(h, set()) == want, h


# --------------------------------------------------------------------------------------
# samplers against closed-form CDFs
# --------------------------------------------------------------------------------------
def _sample(scene, h, n):
    sun = scene.sun()["sun_dir"]
    nrm = scene.helio()["normal"][h]
    ths, tns = np.empty(n), np.empty(n)
    L = scene.prm["rays_per_sample"]
    for i in range(n):
        a, b = scene.sample_ray(h, i // L, i % L)
        ths[i] = _angle(-a.astype(np.float64), sun)
        tns[i] = _angle(b, nrm)
    return ths, tns


def test_pillbox_and_slope_error_distributions():
    d, p = F.config_small()
    p = F.with_lattice(p)
    s = oracle.OracleScene(d, p)
    th, tn = _sample(s, 3, 30000)
    tmax = p["pillbox_half_angle"]
    assert th.max() <= tmax * (1 + 1e-4)
    assert stats.kstest(th, lambda x: np.clip(x / tmax, 0, 1) ** 2).pvalue > 1e-3
    sig = p["slope_sigma"]
    # polar deviation |N(0, sigma)| (Q13): half-normal
    assert stats.kstest(tn, lambda x: special.erf(x / (sig * np.sqrt(2)))).pvalue > 1e-3


def _buie_cdf_factory(csr):
    k = 0.9 * np.log(13.5 * csr) * csr ** -0.3
    g = 2.2 * np.log(0.52 * csr) * csr ** 0.43 - 0.1
    disk = lambda t: np.cos(0.326 * t) / np.cos(0.308 * t)
    I465 = integrate.quad(disk, 0, 4.65, epsabs=1e-13)[0]
    tail = lambda t: np.exp(k) / (g + 1) * (t ** (g + 1) - 4.65 ** (g + 1))
    total = I465 + tail(9.3)

    def cdf(t_rad):
        t = np.asarray(t_rad) * 1000.0
        out = np.empty_like(t)
        for i, x in enumerate(t):
            if x <= 4.65:
                out[i] = integrate.quad(disk, 0, x)[0] / total
            else:
                out[i] = (I465 + tail(min(x, 9.3))) / total
        return out
    return cdf, I465, total, k, g


def test_buie_table_against_quadrature():
    d, p = F.config_mid()
    p = F.with_lattice(p)
    s = oracle.OracleScene(F.subset(d, slice(0, 10)), p)
    tb = s.buie_table()
    cdf, I465, total, k, g = _buie_cdf_factory(p["csr"])
    disk = lambda t: np.cos(0.326 * t) / np.cos(0.308 * t)
    for i in (1, 100, 1000, 2048):
        x = float(tb["cdf_x"][i])
        ref = integrate.quad(disk, 0, x)[0]
        assert abs(tb["cdf_y"][i] - ref) <= 2e-5 * ref, (i, tb["cdf_y"][i], ref)
    sun = s.sun()
    assert abs(sun["A"] - total) <= 2e-5 * total
    assert abs(sun["int465"] - I465) <= 2e-5 * I465


def test_buie_sampler_distribution():
    d, p = F.config_mid()
    p = F.with_lattice(p)
    s = oracle.OracleScene(F.subset(d, slice(0, 10)), p)
    th, _ = _sample(s, 2, 20000)
    cdf, *_ = _buie_cdf_factory(p["csr"])
    assert th.max() <= 9.3e-3 * (1 + 1e-4)
    assert (th > 4.65e-3).mean() > 0.01  # the circumsolar tail is sampled
    assert stats.kstest(th, cdf).pvalue > 1e-3


# --------------------------------------------------------------------------------------
# whole path on scenes with analytic answers
# --------------------------------------------------------------------------------------
def _one_helio_scene(recv_size=8.0, pix=0.0625, sun=(45.0, 180.0), size=2.0):
    desc = {
        "helio_pos": np.float32([[3.0, 1.5, -40.0]]),
        "helio_size": np.float32([[size, 0.1, size]]),
        "helio_gap": np.zeros((1, 2), np.float32), "helio_matrix": np.ones((1, 2), np.int32),
        "recv_type": 0, "recv_pos": np.float32([0, 30, 0]),
        "recv_size": np.float32([recv_size, recv_size, 0]), "recv_normal": np.float32([0, 0, -1]),
        "recv_face": 0, "recv_pixel_length": np.float32(pix),
        "grid_pos": np.float32([-64, -8, -64]), "grid_size": np.float32([128, 64, 128]),
        "grid_interval": np.float32([2, 2, 2]), "sun_dir": F.sun_direction(*sun),
        "dni": np.float32(1000.0),
    }
    prm = dict(seed=3, sunshape=0, pillbox_half_angle=0.0, slope_sigma=0.0, micro_mode=0,
               helio_pixel_length=size / 16, rays_per_sample=4, reflectivity=0.88)
    return desc, prm


def test_specular_footprint_exact():
    """Zero sunshape and zero slope error: every ray of a microsample lands on the pixel of
    the mirror reflection of its centre, computed here in double precision."""
    desc, prm = _one_helio_scene()
    s = oracle.OracleScene(desc, prm)
    r = s.trace_f32(records=True)
    sun = _unit(desc["sun_dir"])
    rv = s.receiver()["vertices"].astype(np.float64)
    focus = (rv[0] + rv[2]) / 2
    pos = desc["helio_pos"][0].astype(np.float64)
    n = _unit(_unit(focus - pos) - sun)
    u = _unit(np.cross([0, 1, 0], n))
    v = _unit(np.cross(u, n))
    refl = sun - 2 * np.dot(n, sun) * n
    E1, E2 = rv[1] - rv[0], rv[3] - rv[0]
    nr = np.cross(E1, E2)
    res = s.res_x
    L = prm["rays_per_sample"]
    p = prm["helio_pixel_length"]
    checked = 0
    for m in range(s.micro()["count"][0]):
        row, col = divmod(m, 16)
        loc = np.array([col * p + p / 2 - 1.0, 0.05, row * p + p / 2 - 1.0])
        o = pos + loc[0] * u + loc[1] * n + loc[2] * v
        t = np.dot(rv[0] - o, nr) / np.dot(refl, nr)
        hit = o + t * refl - rv[0]
        uu = np.dot(hit, E1) / np.dot(E1, E1)
        vv = np.dot(hit, E2) / np.dot(E2, E2)
        fu, fv = uu * res, vv * res
        if min(abs(fu - round(fu)), abs(fv - round(fv))) < 1e-3:
            continue
        want = int(fu) * res + int(fv)
        got = r["pixel"][m * L:(m + 1) * L]
        assert np.all(got == want), (m, got, want)
        checked += 1
    assert checked > 200
    assert r["stats"]["hit"] == r["stats"]["rays"]
    # energy per ray |s.n| (calEnergy, eta = 1)
    e = r["energy"].sum() * 2.0 ** -31
    assert abs(e - r["stats"]["rays"] * abs(np.dot(sun, n))) < 1e-5 * e


def test_intercept_ratio_one_for_large_receiver():
    desc, prm = _one_helio_scene(recv_size=40.0, pix=0.5)
    s = oracle.OracleScene(desc, prm)
    r = s.trace_f32()
    scale, total = s.flux_scale()
    flux = oracle.energy_to_float(r["energy"]) * np.float32(scale)
    assert abs(oracle.intercept_ratio(flux, total) - 1.0) < 1e-5


def test_full_shading_by_heliostat_on_the_sun_line():
    """Heliostat B (6x6 m) sits 5 m from A (2x2 m) along the direction towards the sun:
    every sun ray of A is stopped by B, and none of B's is stopped."""
    sun = F.sun_direction(40.0, 170.0)
    a = np.float64([3.0, 1.5, -40.0])
    b = a - 5.0 * _unit(sun)
    desc = {
        "helio_pos": np.float32([a, b]), "helio_size": np.float32([[2, 0.1, 2], [6, 0.1, 6]]),
        "helio_gap": np.zeros((2, 2), np.float32), "helio_matrix": np.ones((2, 2), np.int32),
        "recv_type": 0, "recv_pos": np.float32([0, 30, 0]), "recv_size": np.float32([8, 8, 0]),
        "recv_normal": np.float32([0, 0, -1]), "recv_face": 0, "recv_pixel_length": np.float32(0.0625),
        "grid_pos": np.float32([-64, -8, -64]), "grid_size": np.float32([128, 64, 128]),
        "grid_interval": np.float32([1, 1, 1]), "sun_dir": sun, "dni": np.float32(1000.0),
    }
    prm = dict(seed=9, sunshape=0, pillbox_half_angle=4.65e-3, slope_sigma=2e-3, micro_mode=0,
               helio_pixel_length=0.25, rays_per_sample=8)
    s = oracle.OracleScene(desc, prm)
    r = s.trace_f32(records=True)
    off = s.micro()["ray_offset"]
    assert np.all(r["shade"][off[0]:off[1]] == 1)
    assert np.all(r["shade"][off[1]:off[2]] == -1)
    # the FP64 evaluation agrees ray for ray
    r64 = s.trace_f64(records=True)
    assert np.array_equal(r["shade"], r64["shade"])


# --------------------------------------------------------------------------------------
# trimmed mean (image_smooth.cu:10-58, heap.cuh)
# --------------------------------------------------------------------------------------
def test_trimmed_mean_hand_worked_heap_example():
    """Row [1,5,4,3,2], radius 2, ratio 0.4.  x=2: cnt 5, k 2.  The two-heap update keeps
    max-heap {2,1} and min-heap {1,5} (the first k values enter both heaps; 4, 3, 2 each
    replace the max-heap top), so the result is (15-3-6)/1 = 6, where the textbook
    trimmed mean is 3 (Q16).  Other pixels have k = 1, where the update is exact."""
    img = np.float32([[1, 5, 4, 3, 2]])
    out = oracle.trimmed_mean(img, 2, 0.4)
    assert np.array_equal(out, np.float32([[4, 3.5, 6, 3.5, 3]]))


def test_trimmed_mean_textbook_when_k_le_1():
    rng = np.random.default_rng(7)
    img = rng.gamma(2.0, 1.0, (23, 31)).astype(np.float32)
    out = oracle.trimmed_mean(img, 1, 0.15)
    H, W = img.shape
    for y in range(H):
        for x in range(W):
            w = img[max(0, y - 1):min(H, y + 2), max(0, x - 1):min(W, x + 2)].ravel().astype(np.float64)
            k = int(np.float32(w.size) * np.float32(0.15))
            srt = np.sort(w)
            want = srt[k:w.size - k].mean()
            assert abs(out[y, x] - want) <= 2e-6 * abs(want)
    # k = 0: plain mean
    out0 = oracle.trimmed_mean(img, 2, 0.0)
    assert abs(out0[10, 10] - img[8:13, 8:13].astype(np.float64).mean()) < 1e-5
