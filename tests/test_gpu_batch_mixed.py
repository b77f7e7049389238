"""The persistent batch launch with LDS windows (OM_BATCH_WINDOW) over mixed-precision scenes
(om_params.precision = OM_PREC_MIXED) on the GPU.

The launch traces every scene's chunks with the mixed trace's FP32 walks and certified
decisions, lists each ray with an uncertain decision on its scene's (and grid's) redo list,
and after the launch traces every list in FP64.  Pins:
  * every map of the batch equals the plain mixed trace of that scene alone, bit for bit
    (counts, energies, flux and smoothed flux), on three replays of the batch's graph with
    eager traces between them: the same rays reach the FP64 trace, and integer sums do not
    depend on the order;
  * the counts of the small scenes equal the oracle's FP64 model's (trace_f64) on every
    pixel, and their flux is within 1e-5 of the model's on every lit pixel;
  * the same holds over scenes of two grids, a cylinder map larger than the window (hits
    outside it), and Buie slices of the 5,000-heliostat field;
  * a scene whose redo list fills up (redo_cap = 1) is traced wholly in FP64: its maps equal
    the plain FP64 trace's bit for bit, while the other scenes of the batch keep their mixed
    maps; om_batch_redo_count reports the listed rays and the one full scene;
  * scenes of differing precision, and mixed scenes that trace in FP64 (the cone receiver),
    are refused by the persistent launch;
  * the annual driver's batched steps over mixed forks have the FP64 model's counts of each
    step's sun (on the step's own Philox stream) and equal plain traces of the forks.
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


def _mixed(p, om, **kw):
    return dict(p, precision=om.PREC_MIXED, accumulate=om.ACC_WINDOW, **kw)


def _plain(scene):
    import torch
    m = scene.trace_plain()
    torch.cuda.synchronize()
    return {k: getattr(m, k).cpu().numpy().copy() for k in ("counts", "energy", "flux", "flux_smooth")}


def _same(maps, want, label):
    for k, v in want.items():
        got = getattr(maps, k).cpu().numpy()
        assert np.array_equal(got.view(np.uint32) if got.dtype == np.float32 else got,
                              v.view(np.uint32) if v.dtype == np.float32 else v), (label, k)


def _run_and_compare(om, scenes, want, label, reps=3):
    import torch
    b = om.Batch(scenes, workgroup=om.BATCH_WINDOW)
    assert b.workgroup == om.BATCH_WINDOW
    for rep in range(reps):
        maps = b.run()
        torch.cuda.synchronize()
        for i, m in enumerate(maps):
            _same(m, want[i], f"{label} scene {i} replay {rep}")
        # eager work between replays: a plain trace of one scene
        _same(scenes[rep % len(scenes)].trace_plain(), want[rep % len(scenes)], f"{label} eager {rep}")
    return b


def _model_check(scenes, want, descs, label):
    for i, (d, p) in enumerate(descs):
        o = oracle.OracleScene(d, p).trace_f64()
        assert np.array_equal(want[i]["counts"].view(np.uint32).ravel(), o["counts"].ravel()), (label, i)
        scale = float(np.float32(scenes[i].info["flux_scale"]))
        flux64 = o["energy"].ravel() * scale
        lit = flux64 > 0
        rel = np.abs(want[i]["flux"].astype(np.float64).ravel()[lit] - flux64[lit]) / flux64[lit]
        assert rel.size > 0 and rel.max() <= 1e-5, (label, i, rel.max())


def test_small_scenes_equal_plain_and_the_fp64_model(om):
    descs = []
    d, p = F.config_small()
    for i, (el, az) in enumerate(((45.0, 180.0), (25.0, 120.0), (70.0, 230.0))):
        descs.append((dict(d, sun_dir=F.sun_direction(el, az)), _mixed(p, om, scene_index=i)))
    d, p = F.config_small_cylinder()
    descs.append((d, _mixed(p, om, scene_index=3)))
    d, p = F.config_small(seed=9)
    descs.append((d, _mixed(p, om, scene_index=4)))
    scenes = [om.SolarScene(dd, pp) for dd, pp in descs]
    want = [_plain(s) for s in scenes]
    _model_check(scenes, want, descs, "small")
    b = _run_and_compare(om, scenes, want, "small")
    rays, full = b.redo_count()
    print(f"small batch: {b.rays} rays, {rays} traced again in FP64, {full} full lists", flush=True)
    assert full == 0 and rays <= 0.02 * b.rays


def test_two_grid_scenes_equal_plain(om):
    descs = []
    d, p = F.config_two_grids()
    for i, el in enumerate((50.0, 30.0, 65.0)):
        descs.append((dict(d, sun_dir=F.sun_direction(el, 170.0 + 20 * i)),
                      _mixed(p, om, scene_index=i, grid_select=om.GRIDS_ALL)))
    scenes = [om.SolarScene(dd, pp) for dd, pp in descs]
    assert all(s.info["n_grids"] == 2 for s in scenes)
    want = [_plain(s) for s in scenes]
    _model_check(scenes[:1], want, descs[:1], "two grids")
    _run_and_compare(om, scenes, want, "two grids")


def test_buie_field_slices_equal_plain(om):
    d, p = F.config_mid()
    descs = [(F.subset(d, slice(0, 300)), _mixed(p, om, scene_index=0))]
    dd = F.subset(d, slice(2000, 2400))
    dd["sun_dir"] = F.sun_direction(20.0, 100.0)
    descs.append((dd, _mixed(p, om, scene_index=1)))
    db, pb = F.config_batched(n_scenes=2, h_range=(500, 800))[0]
    descs.append((db, _mixed(pb, om)))
    scenes = [om.SolarScene(dd, pp) for dd, pp in descs]
    want = [_plain(s) for s in scenes]
    b = _run_and_compare(om, scenes, want, "buie", reps=2)
    rays, full = b.redo_count()
    assert full == 0 and 0 < rays <= 0.02 * b.rays, (rays, b.rays)


def test_full_redo_list_traces_its_scene_in_fp64(om):
    d, p = F.config_small()
    d2, p2 = F.config_small(seed=9)
    full_p = _mixed(p, om, scene_index=0, redo_cap=1)
    scenes = [om.SolarScene(d, full_p), om.SolarScene(d2, _mixed(p2, om, scene_index=1))]
    want = [_plain(om.SolarScene(d, dict(full_p, precision=om.PREC_FP64))), _plain(scenes[1])]
    import torch
    b = om.Batch(scenes, workgroup=om.BATCH_WINDOW)
    for rep in range(2):
        maps = b.run()
        torch.cuda.synchronize()
        for i, m in enumerate(maps):
            _same(m, want[i], f"full list scene {i} replay {rep}")
        rays, full = b.redo_count()
        assert full == 1 and rays > 1, (rays, full)


def test_persistent_launch_refusals(om):
    d, p = F.config_small()
    mixed = om.SolarScene(d, _mixed(p, om))
    fp32 = om.SolarScene(d, dict(p, scene_index=1))
    with pytest.raises(om.OMError, match="precision"):
        om.Batch([mixed, fp32], workgroup=om.BATCH_WINDOW)
    dc, pc = F.config_small_cone()
    cone = om.SolarScene(dc, _mixed(pc, om))
    with pytest.raises(om.OMError, match="mixed path"):
        om.Batch([cone], workgroup=om.BATCH_WINDOW)
    with pytest.raises(om.OMError, match="mixed"):
        om.Batch([mixed], workgroup=256)


def test_annual_mixed_batch_steps_equal_the_fp64_model(om):
    """The annual driver over mixed-precision forks takes the persistent launch with LDS
    windows: every time step (fork i, Philox stream i) has the counts of the FP64 model of
    that step's sun and DNI, its flux is within 1e-5 of the model's on every lit pixel, and
    the batched maps equal plain traces of the same forks bit for bit."""
    import torch
    from onemirror_amd import annual as A
    d, p = F.config_small()
    p = _mixed(p, om)
    root = om.SolarScene(d, p)
    steps = A.annual_steps([(3, 1, 8), (6, 16, 12), (9, 10, 16), (12, 4, 10)])
    run = A.AnnualRun(root, steps, batch_size=4, mode="batch")
    assert run._batches[0].workgroup == om.BATCH_WINDOW
    run.run(keep_maps=True)
    for i, st in enumerate(steps):
        o = oracle.OracleScene(dict(d, sun_dir=np.float32(st["sun_dir"]), dni=np.float32(st["dni"])), p)
        o.set_ray_stream(i)
        want = o.trace_f64()
        got = run.maps()[i]
        torch.cuda.synchronize()
        counts = got.counts.cpu().numpy().view(np.uint32).ravel()
        assert np.array_equal(counts, want["counts"].ravel()), i
        scale = float(np.float32(run.forks[i].info["flux_scale"]))
        flux64 = want["energy"].ravel() * scale
        lit = flux64 > 0
        rel = np.abs(got.flux.cpu().numpy().astype(np.float64).ravel()[lit] - flux64[lit]) / flux64[lit]
        assert rel.size > 0 and rel.max() <= 1e-5, (i, rel.max())
        plain = run.forks[i].trace_plain()
        torch.cuda.synchronize()
        assert np.array_equal(plain.counts.cpu().numpy().view(np.uint32).ravel(), counts), i
        assert np.array_equal(plain.energy.cpu().numpy().view(np.uint64).ravel(),
                              got.energy.cpu().numpy().view(np.uint64).ravel()), i
