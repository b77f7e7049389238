"""GPU parity: the HIP path against the CPU oracle on the same Philox streams.

Bar: bit-exact per-ray shade/block/pixel indices, integer pixel counts and fixed-point
energies; FP32 flux maps equal to the oracle's conversion of the same integers.
"""
import os

import numpy as np
import pytest

import oracle
from onemirror_amd import fields as F

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def om():
    import onemirror_amd
    return onemirror_amd


def _records_equal(g, o, label):
    for k in ("shade", "block", "pixel"):
        a, b = g[k], o[k]
        assert a.shape == b.shape, (label, k, a.shape, b.shape)
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"{label}: {k} differs at {bad.size} rays, first {bad[:5]}: " \
                              f"gpu {a[bad[:5]]} oracle {b[bad[:5]]}"


@pytest.mark.parametrize("fn", range(10))
def test_device_math_bitexact(om, fn):
    rng = np.random.default_rng(fn)
    if fn == 0:
        x = rng.integers(0, 2**32, 100000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    elif fn in (1, 2, 8):
        x = oracle.math("u01", rng.integers(0, 2**32, 100000, dtype=np.uint64).astype(np.uint32))
    elif fn in (3, 4):
        x = np.concatenate([rng.normal(0, 0.01, 50000), rng.uniform(-4, 4, 50000)]).astype(np.float32)
    elif fn == 5:
        x = np.exp(rng.uniform(-40, 40, 100000)).astype(np.float32)
    elif fn == 6:
        x = rng.uniform(-30, 30, 100000).astype(np.float32)
    elif fn == 7:
        x = rng.uniform(-0.9999999, 0.9999999, 100000).astype(np.float32)
    else:
        x = np.abs(rng.normal(0, 100, 100000)).astype(np.float32)
    names = ["u01", "sin2pi", "cos2pi", "sin_small", "cos_small", "log", "exp", "erfinv",
             "gauss", "sqrt"]
    want = oracle.math(names[fn], x.view(np.uint32) if fn == 0 else x)
    got = om.default_context().math_eval(fn, x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
        f"{names[fn]}: {(got != want).sum()} mismatches"


def test_atomic_probes_add_every_value(om):
    """The atomic throughput probes (om_math_eval 17 and 18) add 64 values of 64 bits per
    thread into device memory at hashed addresses: the words' total equals the sum of every
    add, returned or not, so each probe does the n x 64 atomics its rate is counted on.  The
    LDS probe (19) runs to completion and leaves each thread's digest (24 bits)."""
    import ctypes
    import torch
    ctx = om.Context(0)
    n = 1 << 16
    x = torch.rand(n, device="cuda")
    v = (x.cpu().numpy().view(np.uint32) & 0xff).astype(np.uint64) + 1
    want = int(v.sum()) * 64
    for fn in (17, 18):
        y = torch.zeros(n, device="cuda")
        assert om.lib().om_math_eval(ctx._h, fn, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                     n, None) == 0
        torch.cuda.synchronize()
        assert int(y.cpu().numpy().view(np.uint64).sum()) == want, fn
    y = torch.full((n,), -1.0, device="cuda")
    assert om.lib().om_math_eval(ctx._h, 19, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                 n, None) == 0
    torch.cuda.synchronize()
    d = y.cpu().numpy()
    assert np.all((d >= 0) & (d < 2 ** 24))


def test_device_sqrt_exhaustive(om):
    """The kernel's short square root equals IEEE sqrt on every float in [1, 4); an input
    scaled by 4^k (no denormal or overflow on either side) scales its root by 2^k exactly,
    so this covers the whole range the kernel uses it on."""
    x = np.arange(0x3F800000, 0x40800000, dtype=np.uint32).view(np.float32)
    got = om.default_context().math_eval(9, x)
    want = np.sqrt(x)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first x {x[bad[:4]]}"


def test_device_division_matches_ieee(om):
    """The kernel's short division equals IEEE division: every reciprocal 1/b for b in
    [1, 2) (the mantissa decides the result bits, exponents only scale them), and random
    quotients over the operand ranges the kernel divides in (zero and 2^-90..2^60
    numerators, 2^-60..2^60 denominators, and the walk's |dir| in (1e-6, 1])."""
    ctx = om.default_context()
    b = np.arange(0x3F800000, 0x40000000, dtype=np.uint32).view(np.float32)
    pairs = np.empty(2 * b.size, np.float32)
    pairs[0::2] = 1.0
    pairs[1::2] = b
    got = ctx.math_eval(10, pairs)[: b.size]
    want = np.float32(1.0) / b
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
        f"{(got != want).sum()} reciprocal mismatches"
    rng = np.random.default_rng(12)
    n = 1 << 22
    sgn = lambda k: rng.choice(np.array([-1.0, 1.0]), k)
    num = np.concatenate([2.0 ** rng.uniform(-90, 60, n) * sgn(n),
                          2.0 ** rng.uniform(-44, 22, n) * sgn(n),
                          np.zeros(4096)]).astype(np.float32)
    den = np.concatenate([2.0 ** rng.uniform(-60, 60, n) * sgn(n),
                          rng.uniform(1e-6, 1.0, n) * sgn(n),
                          rng.uniform(1e-6, 1.0, 4096)]).astype(np.float32)
    pairs = np.empty(2 * num.size, np.float32)
    pairs[0::2] = num
    pairs[1::2] = den
    got = ctx.math_eval(10, pairs)[: num.size]
    want = num / den
    # a zero numerator may give a zero of either sign: the walk takes |q|, and log's result
    # does not depend on the sign of (m - 1) / (m + 1) when that is zero
    ok = (got.view(np.uint32) == want.view(np.uint32)) | ((num == 0) & (got == 0))
    bad = np.nonzero(~ok)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first {num[bad[:3]]} / {den[bad[:3]]}"


@pytest.fixture(scope="module")
def small(om):
    d, p = F.config_small()
    return d, p, om.SolarScene(d, p), oracle.OracleScene(d, p)


def test_small_scene_counts(small):
    d, p, g, o = small
    mc = g.export("micro_count", 4 * o.n_helio, np.int32)
    assert np.array_equal(mc, o.micro()["count"])
    assert g.info["total_rays"] == o.total_rays


def test_small_poisson_points_bitexact(small):
    d, p, g, o = small
    cap = g.info["max_micro"]
    pts = g.export("poisson", 8 * cap * o.n_helio, np.float32).reshape(o.n_helio, cap, 2)
    mc = o.micro()["count"]
    for h in range(o.n_helio):
        want = o.poisson(h)
        assert np.array_equal(pts[h, : mc[h]].view(np.uint32), want.view(np.uint32)), h


def test_small_records_bitexact(small):
    d, p, g, o = small
    gr = g.trace_records()
    orc = o.trace_f32(records=True)
    _records_equal(gr, orc, "small/poisson")


def test_small_lattice_records_bitexact(om):
    d, p = F.config_small()
    p = F.with_lattice(p)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    _records_equal(g.trace_records(), o.trace_f32(records=True), "small/lattice")


def test_small_maps_bitexact(small):
    import torch
    d, p, g, o = small
    maps = g.trace_plain()
    torch.cuda.synchronize()
    orc = o.trace_f32()
    counts = maps.counts.cpu().numpy().view(np.uint32)
    energy = maps.energy.cpu().numpy().view(np.uint64)
    assert np.array_equal(counts, orc["counts"])
    assert np.array_equal(energy, orc["energy"])
    scale, total = o.flux_scale()
    raw = oracle.energy_to_float(orc["energy"])
    flux = (raw * np.float32(scale)).astype(np.float32)
    assert np.array_equal(maps.flux.cpu().numpy().view(np.uint32), flux.view(np.uint32))
    sm = oracle.trimmed_mean(raw, p["smooth_radius"], p["smooth_ratio"]) * np.float32(scale)
    assert np.array_equal(maps.flux_smooth.cpu().numpy().view(np.uint32),
                          sm.astype(np.float32).view(np.uint32))


def test_trace_is_deterministic(small):
    import torch
    d, p, g, o = small
    a = g.trace_plain()
    b = g.trace_plain()
    torch.cuda.synchronize()
    assert torch.equal(a.energy, b.energy) and torch.equal(a.counts, b.counts)


def test_heliostat_accumulation_equals_plain(om, small):
    import torch
    d, p, g, o = small
    full = g.trace_plain()
    acc = g.new_maps(smooth=False)
    for h0 in range(0, o.n_helio, 7):
        g.trace_heliostats(h0, min(h0 + 7, o.n_helio), acc)
    torch.cuda.synchronize()
    assert torch.equal(full.energy, acc.energy) and torch.equal(full.counts, acc.counts)


def test_results_with_given_normal(om, small):
    d, p, g, o = small
    res = om.results_with_given_normal(g)
    orc = o.trace_f32()
    scale, total = o.flux_scale()
    raw = oracle.energy_to_float(orc["energy"])
    flux = (raw * np.float32(scale)).astype(np.float32)
    sm = (oracle.trimmed_mean(raw, p["smooth_radius"], p["smooth_ratio"]) * np.float32(scale)).astype(np.float32)
    assert np.array_equal(res["flux"], flux)
    assert res["ratio"] == oracle.intercept_ratio(flux, total)
    assert res["ratio_smooth"] == oracle.intercept_ratio(sm, total)
    assert 0.5 < res["ratio"] <= 1.0


def test_smoother_matches_oracle(om):
    import torch
    rng = np.random.default_rng(3)
    img = rng.gamma(2.0, 1.0, (61, 77)).astype(np.float32)
    img[rng.random(img.shape) < 0.3] = 0.0
    for r, ratio in ((2, 0.2), (5, 0.02), (1, 0.0), (3, 0.3)):
        t = torch.as_tensor(img).cuda()
        om.ImageSmoother.image_smooth(t, r, ratio, 77, 61)
        torch.cuda.synchronize()
        want = oracle.trimmed_mean(img, r, ratio)
        assert np.array_equal(t.cpu().numpy().view(np.uint32), want.view(np.uint32)), (r, ratio)


def test_face2face_scene(om):
    path = os.path.join(HERE, "golden", "face2face_shadow.scn")
    sun = F.sun_direction(40.0, 160.0)
    p = dict(om.default_params())
    p.update(rays_per_sample=256, helio_pixel_length=0.05, slope_sigma=1e-3, seed=7)
    g = om.SolarScene.from_scn(path, sun, 1000.0, p)
    d = om.scn_parse(path)
    d["sun_dir"] = sun
    o = oracle.OracleScene(d, p)
    assert g.info["res_x"] == 400 and g.info["res_y"] == 400
    _records_equal(g.trace_records(), o.trace_f32(records=True), "face2face")


def test_set_sun_reaims(om, small):
    d, p, _, _ = small
    g = om.SolarScene(d, F.with_lattice(p))
    sun = F.sun_direction(30.0, 200.0)
    g.set_sun(sun, 900.0)
    d2 = dict(d)
    d2["sun_dir"] = sun
    d2["dni"] = np.float32(900.0)
    o = oracle.OracleScene(d2, F.with_lattice(p))
    _records_equal(g.trace_records(), o.trace_f32(records=True), "set_sun")
    assert g.info["flux_scale"] == o.flux_scale()[0]


@pytest.fixture(scope="module")
def mid(om):
    d, p = F.config_mid()
    return d, p, om.SolarScene(d, p), oracle.OracleScene(d, p)


def test_mid_poisson_counts(mid):
    d, p, g, o = mid
    mc = g.export("micro_count", 4 * o.n_helio, np.int32)
    assert np.array_equal(mc, o.micro()["count"])


@pytest.mark.parametrize("elev", [18.0, 35.0])
def test_clear_heliostats_keep_records_and_maps(om, elev):
    """Low suns shade much of a field: the kernels skip the shading walks of the heliostats
    marked clear and walk the rest; records and maps stay the oracle's."""
    import torch
    d, p = F.config_mid()
    d = dict(F.subset(d, slice(2000, 2250)), sun_dir=F.sun_direction(elev, 200.0))
    clear = om.clear_heliostats(d, p)
    print(f"clear heliostats elevation {elev}: {int(clear.sum())} of {clear.size}", flush=True)
    assert 0 < clear.sum() < clear.size
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    orc = o.trace_f32(records=True)
    _records_equal(g.trace_records(), orc, f"clear/{elev}")
    assert (orc["shade"] >= 0).sum() > 0
    maps = g.trace_plain()
    torch.cuda.synchronize()
    assert np.array_equal(maps.counts.cpu().numpy().view(np.uint32), orc["counts"])
    assert np.array_equal(maps.energy.cpu().numpy().view(np.uint64), orc["energy"])


@pytest.mark.parametrize("h0", [0, 2400, 4900])
def test_mid_records_bitexact(mid, h0):
    d, p, g, o = mid
    h1 = h0 + 40
    _records_equal(g.trace_records(h0, h1), o.trace_f32(h0, h1, records=True), f"mid[{h0}:{h1}]")


@pytest.mark.parametrize("mode,shift,compact", [(0, 0, 0), (1, 0, 0), (0, 32, 0), (0, 33, 0), (0, 62, 0),
                                                (0, 61, 0), (2, 0, 0), (2, 32, 0), (2, 33, 0), (2, 62, 0),
                                                (2, 61, 0), (2, 0, 1), (2, 33, 1), (2, 62, 1)])
def test_accumulation_modes_bitexact(om, mode, shift, compact):
    """Packed single-atomic accumulation (with its carry and wrap bookkeeping driven hard by
    narrow fields: E = 32/33 carries out of the energy field every hit or two, E = 61/62
    wraps the count field every few hits), the same words kept in an LDS window of the map
    by persistent workgroups (mode 2: hits inside and outside the window, LDS words that
    carry and wrap; compact: with the rays that pass the blocking test packed into full
    waves for the receiver stage) and two device-scope atomics per hit all reproduce the
    oracle's integer maps."""
    import torch
    d, p = F.config_mid()
    d = F.subset(d, slice(1000, 1400))
    p = dict(p, accumulate=mode, pack_shift=shift, ray_compact=compact)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    maps = g.trace_plain()
    torch.cuda.synchronize()
    orc = o.trace_f32()
    assert np.array_equal(maps.counts.cpu().numpy().view(np.uint32), orc["counts"])
    assert np.array_equal(maps.energy.cpu().numpy().view(np.uint64), orc["energy"])
    assert int(orc["counts"].sum()) > 1_000_000


@pytest.mark.parametrize("L,lattice,compact", [(1, False, 0), (7, False, 0), (8, True, 0), (40, True, 0),
                                               (64, False, 0), (65, True, 0), (100, False, 0), (2048, True, 0),
                                               (1, False, 1), (7, True, 1), (65, False, 1), (2048, True, 1)])
def test_window_sample_blocks_bitexact(om, L, lattice, compact):
    """The window kernel traces passes of 64 consecutive rays and holds the starts of 64
    consecutive microsamples, one per lane, refilled when a pass reaches past them: rays per
    microsample below, at and above 64 (a pass within one microsample, across several, and
    refills at every pass for L = 1 or every few for L = 7, 8), and a heliostat's last,
    partial pass all give the oracle's maps, with and without ray compaction (rays held
    across passes, chunks and heliostats until a wave is full, the rest run after the
    wave's last chunk)."""
    import torch
    d, p = F.config_mid()
    n = max(4, min(200, 8000 // L))
    d = F.subset(d, slice(1500, 1500 + n))
    p = dict(F.with_lattice(p) if lattice else p, accumulate=2, rays_per_sample=L, ray_compact=compact)
    print(f"window blocks L={L} lattice={lattice} compact={compact} heliostats={n}", flush=True)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    maps = g.trace_plain()
    torch.cuda.synchronize()
    orc = o.trace_f32()
    assert np.array_equal(maps.counts.cpu().numpy().view(np.uint32), orc["counts"])
    assert np.array_equal(maps.energy.cpu().numpy().view(np.uint64), orc["energy"])
    assert int(orc["counts"].sum()) > 10_000


def test_accumulation_autotune_keeps_the_maps(om):
    """The accumulation autotuner times both atomic strategies on the scene and keeps the
    faster; the maps after it are the oracle's, whichever it kept."""
    import torch
    d, p = F.config_mid()
    d = F.subset(d, slice(2000, 2300))
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    r = g.autotune_accumulate(reps=2)
    assert r["accumulate"] in ("packed", "window")
    assert r["packed_ms"] > 0 and r["window_ms"] > 0
    maps = g.trace_plain()
    torch.cuda.synchronize()
    orc = o.trace_f32()
    assert np.array_equal(maps.counts.cpu().numpy().view(np.uint32), orc["counts"])
    assert np.array_equal(maps.energy.cpu().numpy().view(np.uint64), orc["energy"])


def test_device_atan2_bitexact(om):
    rng = np.random.default_rng(13)
    y = rng.normal(size=200000).astype(np.float32)
    x = rng.normal(size=200000).astype(np.float32)
    pairs = np.empty(2 * y.size, np.float32)
    pairs[0::2], pairs[1::2] = y, x
    got = om.default_context().math_eval(11, pairs)[: y.size]
    want = oracle.math("atan2", pairs)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("lattice", [False, True])
def test_cylinder_small_records_and_maps_bitexact(om, lattice):
    import torch
    d, p = F.config_small_cylinder()
    if lattice:
        p = F.with_lattice(p)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    orc = o.trace_f32(records=True)
    _records_equal(g.trace_records(), orc, f"cylinder/{'lattice' if lattice else 'poisson'}")
    assert (orc["pixel"] >= 0).mean() > 0.2
    maps = g.trace_plain()
    torch.cuda.synchronize()
    assert np.array_equal(maps.counts.cpu().numpy().view(np.uint32).ravel(), orc["counts"].ravel())
    assert np.array_equal(maps.energy.cpu().numpy().view(np.uint64).ravel(), orc["energy"].ravel())
    scale, _ = o.flux_scale()
    flux = (oracle.energy_to_float(orc["energy"]) * np.float32(scale)).astype(np.float32)
    assert np.array_equal(maps.flux.cpu().numpy().view(np.uint32).ravel(), flux.view(np.uint32).ravel())


@pytest.mark.parametrize("compact", [0, 1])
def test_cylinder_sector_windows_bitexact(om, compact):
    """A cylinder map larger than one LDS window (320 x 160 pixels): the window trace gives
    each of its 8 queues one azimuth sector of the heliostats and the window centred on it
    (columns wrapping around the cylinder), and the decode adds each workgroup's window where
    its queue's lies.  The maps equal the oracle's bit for bit, with and without ray
    compaction, on two traces in a row; a trace of a heliostat range (the centred window for
    every queue) in between gives its own maps."""
    import torch
    d, p = F.config_small_cylinder()
    d = dict(d, recv_res_x=320, recv_res_y=160)
    p = dict(p, accumulate=om.ACC_WINDOW, ray_compact=compact)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    assert g.info["res_x"] * g.info["res_y"] == 320 * 160
    orc = o.trace_f32()
    part = o.trace_f32(20, 60)
    for rep in range(2):
        maps = g.trace_plain()
        torch.cuda.synchronize()
        assert np.array_equal(maps.counts.cpu().numpy().view(np.uint32).ravel(), orc["counts"].ravel()), rep
        assert np.array_equal(maps.energy.cpu().numpy().view(np.uint64).ravel(), orc["energy"].ravel()), rep
        m = g.new_maps(smooth=False)
        g.trace_heliostats(20, 60, m)
        torch.cuda.synchronize()
        assert np.array_equal(m.counts.cpu().numpy().view(np.uint32).ravel(), part["counts"].ravel()), rep
        assert np.array_equal(m.energy.cpu().numpy().view(np.uint64).ravel(), part["energy"].ravel()), rep
    assert int(orc["counts"].sum()) > 100_000


@pytest.fixture(scope="module")
def large(om):
    d, p = F.config_large()
    return d, p, om.SolarScene(d, p), oracle.OracleScene(d, p)


def test_large_poisson_counts(large):
    d, p, g, o = large
    mc = g.export("micro_count", 4 * o.n_helio, np.int32)
    assert np.array_equal(mc, o.micro()["count"])
    assert g.info["total_rays"] == o.total_rays > 9.9e8


@pytest.mark.parametrize("h0", [0, 12000, 24980])
def test_large_records_bitexact(large, h0):
    d, p, g, o = large
    _records_equal(g.trace_records(h0, h0 + 20), o.trace_f32(h0, h0 + 20, records=True),
                   f"large[{h0}:{h0 + 20}]")
