"""The faceted heliostat of the annual driver on the GPU (bit-exact against the oracle).

The scene of scene_file_proc.cpp:288-534 (nine facets with given normals, its receiver and
grid) traced with the reference's default parameters, its per-ray records, and a stretch of
the annual driver (main.cpp:56-116: one faceted scene per time step) batched and plain.
"""
import numpy as np
import pytest

import oracle
from onemirror_amd import fields as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def om():
    import onemirror_amd
    return onemirror_amd


def _desc(om, sun, dni=950.0):
    s = F.facet_site()
    return om.facet_scene(s["a"], s["b"], s["X"], s["Y"], s["Z"], sun, dni)


@pytest.mark.parametrize("el,az", [(50.0, 160.0), (20.0, 100.0), (70.0, 250.0)])
def test_facet_scene_maps_bitexact(om, el, az):
    import torch
    d = _desc(om, F.sun_direction(el, az))
    p = om.default_params()
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    print(f"facet scene sun {el}/{az}: {g.info['total_rays']} rays, map {g.res_x}x{g.res_y}", flush=True)
    assert g.info["total_rays"] == o.total_rays == 9 * 47 * 47 * 2048
    m = g.trace_plain()
    torch.cuda.synchronize()
    orc = o.trace_f32()
    assert np.array_equal(m.counts.cpu().numpy().view(np.uint32), orc["counts"])
    assert np.array_equal(m.energy.cpu().numpy().view(np.uint64), orc["energy"])
    assert int(orc["counts"].sum()) > 0.5 * o.total_rays
    scale, total = o.flux_scale()
    raw = oracle.energy_to_float(orc["energy"])
    assert np.array_equal(m.flux.cpu().numpy().view(np.uint32), (raw * np.float32(scale)).astype(np.float32).view(np.uint32))


def test_facet_scene_records_bitexact(om):
    d = _desc(om, F.sun_direction(35.0, 140.0))
    p = dict(om.default_params(), rays_per_sample=64)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    gr, orc = g.trace_records(), o.trace_f32(records=True)
    for k in ("shade", "block", "pixel"):
        assert np.array_equal(gr[k], orc[k]), k
    # the nine facets share one 110 m cell, and the walk never tests the cell a ray starts
    # in (rectgrid_raytracing.cu:69-115): in this scene no ray is shaded or blocked
    assert (gr["shade"] == -1).all() and (gr["block"] == -1).all()
    assert (gr["pixel"] >= 0).mean() > 0.5


def test_facet_annual_batch_equals_plain(om):
    from onemirror_amd.annual import AnnualRun, annual_steps
    steps = annual_steps()[::40]
    s = F.facet_site()
    p = dict(om.default_params(), rays_per_sample=256)
    runs = {mode: AnnualRun.facets(s["a"], s["b"], s["X"], s["Y"], s["Z"], steps, p, mode=mode, batch_size=8)
            for mode in ("batch", "plain")}
    res = {mode: r.run(keep_maps=True) for mode, r in runs.items()}
    for a, b in zip(res["batch"]["flux"], res["plain"]["flux"]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    rows = res["batch"]["steps"]
    assert len(rows) == len(steps) and all(0.3 < r["ratio"] <= 1.0 for r in rows)
    assert 0.3 < res["batch"]["annual_ratio"] <= 1.0


def test_driver_text_outputs_and_loaded_normals(om, tmp_path):
    """The driver's outputs (Gtest.cu:47,59: one ratio line per trace appended to the
    unsmoothed and smoothed result files) for the faceted scene with the normals of the
    reference's userData/normal.txt (InitContent(normals) with load_noraml), against the
    oracle's maps written by the restatement of ImageSaver::savetxt."""
    import os
    import oracle.io as oio
    here = os.path.dirname(os.path.abspath(__file__))
    normals = om.load_normals(os.path.join(here, "golden", "normal.txt"))
    p = dict(om.default_params(), rays_per_sample=128)
    ns, sm = str(tmp_path / "result_nonsmooth.txt"), str(tmp_path / "result_smooth.txt")
    want_ns = want_sm = ""
    for el, az in ((50.0, 160.0), (30.0, 200.0)):
        d = _desc(om, F.sun_direction(el, az))
        d["helio_normal"] = normals
        g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
        res = om.results_with_given_normal(g, ns, sm)
        orc = o.trace_f32()
        scale, total = o.flux_scale()
        raw = oracle.energy_to_float(orc["energy"])
        flux = (raw * np.float32(scale)).astype(np.float32)
        smooth = (oracle.trimmed_mean(raw, p["smooth_radius"], p["smooth_ratio"]) * np.float32(scale)).astype(np.float32)
        assert np.array_equal(res["flux"].view(np.uint32), flux.view(np.uint32))
        assert np.array_equal(res["flux_smooth"].view(np.uint32), smooth.view(np.uint32))
        h, w = flux.shape
        want_ns += oio.savetxt_ratio(w, h, flux, total)
        want_sm += oio.savetxt_ratio(w, h, smooth, total)
    assert open(ns).read() == want_ns
    assert open(sm).read() == want_sm
