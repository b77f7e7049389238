"""The mixed-precision trace (om_params.precision = OM_PREC_MIXED) on the GPU against the
oracle's FP64 model of the same rays (oracle trace_f64).

The mixed trace samples, walks and meets the receiver in FP32, with each ray's direction an
FP32 deviation from the FP64 model's nominal reflection of its heliostat, and every decision
certified by an error bound; rays with a decision inside its bound are traced again by the
FP64 trace.  Its claim is the FP64 model's outcome for every ray, so
the pins are those of the FP64 trace:
  * per-ray records (shading heliostat, blocking heliostat, receiver pixel) equal the FP64
    model's on the 100-heliostat field (Poisson and lattice microsamples, pillbox sun; 64, 5
    and 1 rays per microsample), the cylinder field, and slices of the 5,000- and
    25,000-heliostat fields (Buie sun);
  * the integer count maps of the window trace equal the FP64 model's counts, on the small
    fields and on the whole 5,000-heliostat field (2e8 rays): a ray whose FP32 decision
    differed from the model's without being sent to FP64 would show as a count off by one;
  * the FP32 flux map is within 1e-5 relative per lit pixel of the FP64 model's flux;
  * the share of rays sent to the FP64 trace is small (printed; bounded at 2 %);
  * a redo list that fills up sends the whole launch to the FP64 trace: the maps are then
    the FP64 trace's, bit for bit, and om_ctx_progress reports that trace's chunks, all
    taken; a list that fits leaves the mixed launch's own chunk counts to om_ctx_progress;
  * a per-scene batch of mixed scenes equals their plain traces on three replays of its graph
    with eager traces between them; the persistent batch launch of 128-512 threads and the
    differentiable mode refuse mixed scenes (the persistent launch with LDS windows takes
    them: test_gpu_batch_mixed.py).
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


def _records_equal(g, o, label):
    for k in ("shade", "block", "pixel"):
        bad = np.nonzero(g[k] != o[k])[0]
        assert bad.size == 0, f"{label}: {k} differs at {bad.size} of {g[k].size} rays, first {bad[:5]}"


def _maps(g):
    import torch
    m = g.trace_plain()
    torch.cuda.synchronize()
    return (m.counts.cpu().numpy().view(np.uint32).ravel().copy(), m.energy.cpu().numpy().view(np.uint64).ravel().copy(),
            m.flux.cpu().numpy().astype(np.float64).ravel())


def _check_against_model(g, want, label):
    counts, energy, flux32 = _maps(g)
    redo = g.redo_count()
    rays = g.info["total_rays"]
    bad = np.nonzero(counts != want["counts"].ravel())[0]
    assert bad.size == 0, f"{label}: counts differ at {bad.size} pixels, first {bad[:5]}"
    scale = float(np.float32(g.info["flux_scale"]))
    flux64 = want["energy"].ravel() * scale
    lit = flux64 > 0
    rel = np.abs(flux32[lit] - flux64[lit]) / flux64[lit]
    print(f"{label}: {rays} rays, {redo} to FP64 ({redo / rays:.2e}), flux rel. error max {rel.max():.2e}", flush=True)
    assert rel.size > 0 and rel.max() <= 1e-5, (label, rel.max())
    assert redo <= 0.02 * rays, (label, redo, rays)
    return redo


@pytest.mark.parametrize("cfg", ["small", "small_lattice", "cylinder", "small_L5", "small_L1", "cylinder_sectors"])
def test_small_fields_records_and_maps(om, cfg):
    """small_L5, small_L1: 5 and 1 rays per microsample, so that a wave's chunk of rays spans
    more than the 64 microsample starts a wave holds at a time (its table is refilled every
    5 passes, and every pass)."""
    if cfg.startswith("cylinder"):
        d, p = F.config_small_cylinder()
        if cfg == "cylinder_sectors":  # a map larger than one window: per-sector windows
            d = dict(d, recv_res_x=320, recv_res_y=160)
    else:
        d, p = F.config_small()
        if cfg == "small_lattice":
            p = F.with_lattice(p)
        elif cfg.startswith("small_L"):
            p = dict(p, rays_per_sample=int(cfg[7:]))
    p = _mixed(p, om)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    want = o.trace_f64(records=True)
    _records_equal(g.trace_records(), want, cfg)
    print(f"{cfg} records: {g.redo_count()} of {g.info['total_rays']} rays to FP64", flush=True)
    _check_against_model(g, want, cfg)


def test_cone_field_traces_in_fp64(om):
    """The cone receiver has no mixed form: the scene traces in FP64, records and maps the
    model's."""
    d, p = F.config_small_cone()
    p = _mixed(p, om)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    want = o.trace_f64(records=True)
    _records_equal(g.trace_records(), want, "cone")
    counts, _, _ = _maps(g)
    assert np.array_equal(counts, want["counts"].ravel())
    assert g.redo_count() == 0


@pytest.mark.parametrize("name,h0,h1", [("mid", 2000, 2040), ("large", 12000, 12008)])
def test_field_slice_records(om, name, h0, h1):
    d, p = (F.config_mid if name == "mid" else F.config_large)()
    p = _mixed(p, om)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    _records_equal(g.trace_records(h0, h1), o.trace_f64(h0, h1, records=True), f"{name} [{h0}, {h1})")


def test_mid_whole_field_counts_equal_the_fp64_model(om):
    d, p = F.config_mid()
    p = _mixed(p, om)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    redo = _check_against_model(g, o.trace_f64(), "mid whole field")
    assert redo > 0  # the bounds do send rays to FP64 on a field of this size
    # the list fitted, so progress reports the mixed launch's own chunks, all taken
    taken, total = g.ctx.progress()
    assert total > 0 and taken == total


def test_full_redo_list_traces_the_launch_in_fp64(om):
    """redo_cap = 1: the first ray sent to FP64 fills the list, so the whole launch is traced
    again by the FP64 window trace: counts and energies equal the FP64 trace's bit for bit."""
    d, p = F.config_small()
    want = _maps(om.SolarScene(d, dict(p, precision=om.PREC_FP64, accumulate=om.ACC_WINDOW)))
    g = om.SolarScene(d, _mixed(p, om, redo_cap=1))
    got = _maps(g)
    assert g.redo_count() > 1
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    # the FP64 retrace took every chunk of the launch's queue again
    taken, total = g.ctx.progress()
    assert total > 0 and taken == total
    # a second trace on the same scene starts from cleared buffers
    again = _maps(g)
    assert np.array_equal(again[0], want[0]) and np.array_equal(again[1], want[1])


def test_mixed_batch_forms(om):
    import torch
    d, p = F.config_small()
    scenes = [om.SolarScene(dict(d, sun_dir=F.sun_direction(el, 180.0)), _mixed(dict(p, scene_index=i), om))
              for i, el in enumerate((45.0, 30.0))]
    want = [_maps(s)[:2] for s in scenes]
    batch = om.Batch(scenes, workgroup=om.BATCH_PER_SCENE)
    # three replays of the batch's graph, with eager traces of the scenes between them (the
    # redo counters and lists are cleared by fill kernels inside the graph)
    for rep in range(3):
        maps = batch.run()
        torch.cuda.synchronize()
        for i, m in enumerate(maps):
            assert np.array_equal(m.counts.cpu().numpy().view(np.uint32).ravel(), want[i][0]), (rep, i)
            assert np.array_equal(m.energy.cpu().numpy().view(np.uint64).ravel(), want[i][1]), (rep, i)
        assert np.array_equal(_maps(scenes[rep % 2])[0], want[rep % 2][0])
    with pytest.raises(om.OMError, match="mixed"):
        om.Batch(scenes, workgroup=256)


def test_differentiable_mode_refuses_mixed(om):
    import torch
    from onemirror_amd import diff
    d, p = F.config_small()
    g = om.SolarScene(d, _mixed(p, om))
    theta = torch.zeros((g.info["n_helio"], 4), dtype=torch.float32, device="cuda")
    with pytest.raises(om.OMError, match="MIXED"):
        diff.forward(g, theta)


def test_workspace_is_clean_after_a_mixed_trace(om):
    """A mixed trace's redo list and the window workgroups' queue numbers lie in the trace
    workspace, past the packed copies of its map.  A later trace of a larger map on the same
    context takes those words as packed copies and adds to them, so they must be left
    cleared: the whole 5,000-heliostat field in mixed precision (its list and sector-free
    window), then the cylinder-map heliostat slice of the 25,000-heliostat field in FP32,
    whose maps must equal the oracle's; and again after a mixed trace of a cylinder field
    with sector windows."""
    import torch
    d, p = F.config_mid()
    g = om.SolarScene(d, _mixed(p, om))
    g.trace_plain()
    torch.cuda.synchronize()
    assert g.redo_count() > 0
    dl, pl = F.config_large()
    big = om.SolarScene(dl, pl, g.ctx)
    want = oracle.OracleScene(dl, pl).trace_f32(12000, 12006)
    for rep in range(2):
        maps = big.new_maps(smooth=False)
        big.trace_heliostats(12000, 12006, maps)
        torch.cuda.synchronize()
        assert np.array_equal(maps.counts.cpu().numpy().view(np.uint32), want["counts"]), rep
        assert np.array_equal(maps.energy.cpu().numpy().view(np.uint64), want["energy"]), rep
        dc, pc = F.config_small_cylinder()
        cyl = om.SolarScene(dict(dc, recv_res_x=320, recv_res_y=160), _mixed(pc, om), g.ctx)
        cyl.trace_plain()
        torch.cuda.synchronize()
        assert cyl.redo_count() > 0
