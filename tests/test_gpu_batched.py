"""Batched mode on the GPU: one persistent launch over many scenes (work items of 128-512
threads, or OM_BATCH_WINDOW's 1024-thread workgroups with each scene's map window in LDS),
or the window trace of each scene in turn (OM_BATCH_PER_SCENE), replayed from a hipGraph.

Pin: every map of a batch equals the map the plain path gives for that scene alone (same
Philox streams, integer accumulation), bit for bit, for every launch configuration and on
every graph replay (replays are interleaved with other GPU work and synchronisations); and
those plain maps equal the oracle's integer maps of each scene.
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


def _plain(scene):
    import torch
    m = scene.trace_plain()
    torch.cuda.synchronize()
    return {k: getattr(m, k).cpu().numpy().copy() for k in ("counts", "energy", "flux", "flux_smooth")}


def _oracle_equal(scenes, want, descs):
    """The plain maps of each scene equal the oracle's integer maps."""
    for i, (d, p) in enumerate(descs):
        o = oracle.OracleScene(d, p).trace_f32()
        assert np.array_equal(want[i]["counts"].view(np.uint32).ravel(), o["counts"].ravel()), i
        assert np.array_equal(want[i]["energy"].view(np.uint64).ravel(), o["energy"].ravel()), i


def _same(maps, want, label):
    for k, v in want.items():
        got = getattr(maps, k).cpu().numpy()
        assert np.array_equal(got.view(np.uint32) if got.dtype == np.float32 else got,
                              v.view(np.uint32) if v.dtype == np.float32 else v), (label, k)


@pytest.fixture(scope="module")
def pillbox_batch(om):
    """Annual steps of one field (three sun positions), the cylinder field, a second seed."""
    descs = []
    d, p = F.config_small()
    for i, (el, az) in enumerate(((45.0, 180.0), (25.0, 120.0), (70.0, 230.0))):
        dd = dict(d)
        dd["sun_dir"] = F.sun_direction(el, az)
        descs.append((dd, dict(p, scene_index=i)))
    d, p = F.config_small_cylinder()
    descs.append((d, dict(p, scene_index=3)))
    d, p = F.config_small(seed=9)
    descs.append((d, dict(p, scene_index=4)))
    scenes = [om.SolarScene(d, p) for d, p in descs]
    want = [_plain(s) for s in scenes]
    _oracle_equal(scenes, want, descs)
    return scenes, want


@pytest.mark.parametrize("wg,rpt", [(0, 0), (128, 2), (512, 8), (256, 1), (1024, 0), (2048, 0)])
def test_batch_graph_replays_equal_plain(om, pillbox_batch, wg, rpt):
    import torch
    scenes, want = pillbox_batch
    b = om.Batch(scenes, workgroup=wg, rays_per_thread=rpt)
    assert b.workgroup == (wg or 256)
    assert b.items > 0 and b.rays == sum(s.info["total_rays"] for s in scenes)
    for rep in range(3):
        maps = b.run()
        torch.cuda.synchronize()
        for i, m in enumerate(maps):
            _same(m, want[i], f"scene {i} replay {rep}")
        # unrelated GPU work between replays: another trace and a torch reduction
        scenes[0].trace_plain()
        torch.cuda.synchronize()
        assert float(torch.ones(1 << 20, device="cuda").sum()) == float(1 << 20)


def test_trace_batched_one_shot(om, pillbox_batch):
    scenes, want = pillbox_batch
    maps = om.trace_batched(scenes, rays_per_thread=4)
    for i, m in enumerate(maps):
        _same(m, want[i], f"scene {i}")


@pytest.mark.parametrize("wg", [0, 1024, 2048])
def test_batch_buie_fields_equal_plain(om, wg):
    import torch
    d, p = F.config_mid()
    descs = [(F.subset(d, slice(0, 300)), dict(p, scene_index=0))]
    dd = F.subset(d, slice(2000, 2400))
    dd["sun_dir"] = F.sun_direction(20.0, 100.0)
    descs.append((dd, dict(p, scene_index=1)))
    descs.append(F.config_batched(n_scenes=2, h_range=(500, 800))[0])
    scenes = [om.SolarScene(d, p) for d, p in descs]
    want = [_plain(s) for s in scenes]
    if wg == 0:
        _oracle_equal(scenes, want, descs)
    b = om.Batch(scenes, workgroup=wg)
    for rep in range(2):
        maps = b.run()
        torch.cuda.synchronize()
        for i, m in enumerate(maps):
            _same(m, want[i], f"buie scene {i} replay {rep}")


def test_batch_window_two_grids_and_scene_changes(om):
    """OM_BATCH_WINDOW over scenes of two grids each (a workgroup adds its window to the
    maps when the next block belongs to another grid or scene) and over scenes whose maps
    exceed the window (hits outside it go through each wave's buffer): maps equal the plain
    ones on three replays."""
    import torch
    descs = []
    d, p = F.config_two_grids()
    for i, el in enumerate((50.0, 30.0, 65.0)):
        descs.append((dict(d, sun_dir=F.sun_direction(el, 170.0 + 20 * i)),
                      dict(p, scene_index=i, grid_select=om.GRIDS_ALL)))
    scenes = [om.SolarScene(dd, pp) for dd, pp in descs]
    assert all(s.info["n_grids"] == 2 for s in scenes)
    want = [_plain(s) for s in scenes]
    b = om.Batch(scenes, workgroup=om.BATCH_WINDOW)
    for rep in range(3):
        maps = b.run()
        torch.cuda.synchronize()
        for i, m in enumerate(maps):
            _same(m, want[i], f"two-grid scene {i} replay {rep}")
        assert float(torch.ones(1 << 20, device="cuda").sum()) == float(1 << 20)


def test_autotune_launch_table(om, pillbox_batch):
    scenes, _ = pillbox_batch
    r = om.autotune_launch(scenes, reps=1)
    forms = [t[:2] for t in r["trials"]]
    assert len(r["trials"]) == 11 and (om.BATCH_PER_SCENE, 0) in forms and (om.BATCH_WINDOW, 0) in forms
    assert all(ms > 0 for _, _, ms in r["trials"])
    assert (r["workgroup"], r["rays_per_thread"], r["ms"]) in r["trials"]
    assert r["rays_per_s"] > 0


def test_batch_rejects_mixed_sunshapes(om):
    d, p = F.config_small()
    a = om.SolarScene(d, p)
    b = om.SolarScene(d, dict(p, sunshape=F.BUIE))
    with pytest.raises(om.OMError, match="sunshape"):
        om.Batch([a, b])
    with pytest.raises(om.OMError, match="workgroup"):
        om.Batch([a], workgroup=96)


def test_per_scene_batch_survives_context_buffer_growth(om):
    """A per-scene batch (OM_BATCH_PER_SCENE) traces in its own workspace: a later plain
    trace of a larger map on the same context, which reallocates the context's buffers
    (map copies, window slices, smoothing input), leaves the batch's graph replays correct,
    and so does a second per-scene batch run between them."""
    import torch
    ctx = om.Context(0)
    d, p = F.config_small()
    small = [om.SolarScene(d, dict(p, scene_index=i), ctx=ctx) for i in range(2)]
    want = [_plain(s) for s in small]
    b = om.Batch(small, workgroup=om.BATCH_PER_SCENE)
    maps = b.run()
    torch.cuda.synchronize()
    for i, m in enumerate(maps):
        _same(m, want[i], f"scene {i} first run")
    # a larger map on the same context: 300 heliostats of the mid field, 256 x 256 map,
    # window accumulation
    dl, pl = F.config_mid()
    big = om.SolarScene(F.subset(dl, slice(0, 300)), dict(pl, accumulate=om.ACC_WINDOW), ctx=ctx)
    assert big.info["res_x"] * big.info["res_y"] > small[0].info["res_x"] * small[0].info["res_y"]
    big.trace_plain()
    other = om.Batch([big], workgroup=om.BATCH_PER_SCENE)
    other.run()
    for rep in range(2):
        maps = b.run()
        torch.cuda.synchronize()
        for i, m in enumerate(maps):
            _same(m, want[i], f"scene {i} after growth, replay {rep}")
