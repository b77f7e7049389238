"""The reference's literal self-intersection rule on the GPU (om_params.literal_self = 1).

Intersect (rectgrid_raytracing.cu:21-36) tests a ray's own heliostat like every other
candidate and rejects it only through t < 1e-6 (global_function.cuh:102); the ray starts on
that mirror's plane, so FP32 rounding of t decides.  Pin: with the rule switched on, per-ray
records and the integer maps are bit-exact against the oracle under the same rule (whose
difference from the default rule is pinned in test_oracle_pins.py), on the 100-heliostat
field (packed and window accumulation), on slices of the 5,000- and 25,000-heliostat fields,
and through a batch; and the rule stops a share of the rays at their own mirror.
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


def _records_equal(g, o, label):
    for k in ("shade", "block", "pixel"):
        bad = np.nonzero(g[k] != o[k])[0]
        assert bad.size == 0, f"{label}: {k} differs at {bad.size} rays, first {bad[:5]}"


def _maps_equal(maps, o, label):
    import torch
    torch.cuda.synchronize()
    assert np.array_equal(maps.counts.cpu().numpy().view(np.uint32), o["counts"]), label
    assert np.array_equal(maps.energy.cpu().numpy().view(np.uint64), o["energy"]), label


@pytest.mark.parametrize("acc", ["packed", "window"])
def test_small_field_literal_records_and_maps(om, acc):
    d, p = F.config_small()
    p = dict(p, literal_self=1, accumulate=om.ACC_PACKED if acc == "packed" else om.ACC_WINDOW)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    want = o.trace_f32(records=True)
    assert want["stats"]["self_hits"] > 0
    _records_equal(g.trace_records(), want, f"small literal {acc}")
    _maps_equal(g.trace_plain(), want, f"small literal {acc} maps")
    # the default rule gives other maps on this field
    default = oracle.OracleScene(d, dict(p, literal_self=0)).trace_f32()
    assert int(default["counts"].sum()) > int(want["counts"].sum())


@pytest.mark.parametrize("name,h0,h1", [("mid", 2000, 2012), ("large", 12000, 12006)])
def test_field_slice_literal_records(om, name, h0, h1):
    d, p = (F.config_mid if name == "mid" else F.config_large)()
    p = dict(p, literal_self=1)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    want = o.trace_f32(h0, h1, records=True)
    assert want["stats"]["self_hits"] > 0
    _records_equal(g.trace_records(h0, h1), want, f"{name} literal [{h0}, {h1})")
    maps = g.new_maps(smooth=False)
    g.trace_heliostats(h0, h1, maps)
    _maps_equal(maps, want, f"{name} literal maps [{h0}, {h1})")


def test_batch_literal_equals_oracle(om):
    import torch
    d, p = F.config_small()
    scenes, want = [], []
    for i, (el, az) in enumerate(((45.0, 180.0), (30.0, 130.0))):
        dd = dict(d, sun_dir=F.sun_direction(el, az))
        pp = dict(p, scene_index=i, literal_self=1)
        scenes.append(om.SolarScene(dd, pp))
        want.append(oracle.OracleScene(dd, pp).trace_f32())
    for wg in (256, om.BATCH_PER_SCENE):
        maps = om.Batch(scenes, workgroup=wg).run()
        torch.cuda.synchronize()
        for i, m in enumerate(maps):
            _maps_equal(m, want[i], f"batch {wg} scene {i}")


def test_differentiable_mode_refuses_literal(om):
    import torch
    d, p = F.config_small()
    g = om.SolarScene(d, dict(p, literal_self=1))
    from onemirror_amd import diff
    theta = torch.zeros((g.info["n_helio"], 4), dtype=torch.float32, device="cuda")
    with pytest.raises(om.OMError, match="literal_self"):
        diff.forward(g, theta)
