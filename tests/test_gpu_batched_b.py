"""The batched workload B on the GPU against the oracle: the first 16 scenes of B (64
scenes in all, heliostats log-uniform in [500, 20,000], suns at elevation 10-80 deg and
azimuth 90-270 deg, 1e7-2e8 rays per scene, 256 x 256 maps; the role of the reference's
annual loop, main.cpp:56-116, which builds one scene per sun position).

Pins:
  * per scene, the persistent batch launches (work items, and LDS windows) and the
    per-scene launch give the maps of a
    plain trace of that scene alone (counts, energy, flux, smoothed flux bit for bit);
  * per scene, three seeded heliostats traced through om_trace_records and
    om_trace_heliostats equal the oracle's records and maps of those heliostats;
  * the whole maps of the three smallest scenes equal the oracle's.
"""
import numpy as np
import pytest

import oracle
from onemirror_amd import fields as F

pytestmark = pytest.mark.gpu

N_SCENES = 16


@pytest.fixture(scope="module")
def om():
    import onemirror_amd
    return onemirror_amd


@pytest.fixture(scope="module")
def batch_b(om):
    cfgs = F.config_batched(n_scenes=N_SCENES)
    scenes = [om.SolarScene(d, p) for d, p in cfgs]
    return cfgs, scenes


def _host(m):
    import torch
    torch.cuda.synchronize()
    return {k: getattr(m, k).cpu().numpy().copy() for k in ("counts", "energy", "flux", "flux_smooth")}


def _bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


def test_b_spans_the_heliostat_range(batch_b):
    cfgs, scenes = batch_b
    hs = [s.info["n_helio"] for s in scenes]
    assert min(hs) < 1000 and max(hs) > 15000, hs
    rays = [s.info["total_rays"] for s in scenes]
    assert min(rays) >= 0.9e7 and max(rays) <= 2.1e8, rays


@pytest.mark.parametrize("wg", [0, 1024, 2048])
def test_b_batch_equals_plain(om, batch_b, wg):
    cfgs, scenes = batch_b
    want = [_host(s.trace_plain()) for s in scenes]
    b = om.Batch(scenes, workgroup=wg)
    for rep in range(2):
        maps = b.run()
        for i, m in enumerate(maps):
            got = _host(m)
            for k in want[i]:
                assert np.array_equal(_bits(got[k]), _bits(want[i][k])), (wg, rep, i, k)
    assert sum(int(w["counts"].sum()) for w in want) > 0


def test_b_heliostat_sample_equals_oracle(om, batch_b):
    import torch
    cfgs, scenes = batch_b
    rng = np.random.default_rng(20)
    for i, ((d, p), g) in enumerate(zip(cfgs, scenes)):
        o = oracle.OracleScene(d, p)
        H = g.info["n_helio"]
        for h in sorted(rng.choice(H, 3, replace=False).tolist()):
            want = o.trace_f32(h, h + 1, records=True)
            rec = g.trace_records(h, h + 1)
            for k in ("shade", "block", "pixel"):
                assert np.array_equal(rec[k], want[k]), (i, h, k)
            maps = g.new_maps(smooth=False)
            g.trace_heliostats(h, h + 1, maps)
            torch.cuda.synchronize()
            assert np.array_equal(maps.counts.cpu().numpy().view(np.uint32).ravel(), want["counts"].ravel()), (i, h)
            assert np.array_equal(maps.energy.cpu().numpy().view(np.uint64).ravel(), want["energy"].ravel()), (i, h)


def test_b_smallest_scenes_whole_maps_equal_oracle(om, batch_b):
    cfgs, scenes = batch_b
    order = np.argsort([s.info["total_rays"] for s in scenes])[:3]
    b = om.Batch([scenes[i] for i in order])
    maps = b.run()
    for j, i in enumerate(order):
        want = oracle.OracleScene(*cfgs[i]).trace_f32()
        got = _host(maps[j])
        assert np.array_equal(got["counts"].view(np.uint32).ravel(), want["counts"].ravel()), i
        assert np.array_equal(got["energy"].view(np.uint64).ravel(), want["energy"].ravel()), i
