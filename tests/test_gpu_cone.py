"""The truncated-cone receiver on the GPU, bit-exact against the oracle: maps and flux,
per-ray records, a batch holding cone, cylinder and rectangle scenes, and a fork."""
import numpy as np
import pytest

import oracle
from onemirror_amd import fields as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def om():
    import onemirror_amd
    return onemirror_amd


def _maps(scene):
    import torch
    m = scene.trace_plain()
    torch.cuda.synchronize()
    return {k: getattr(m, k).cpu().numpy().copy() for k in ("counts", "energy", "flux", "flux_smooth")}


@pytest.mark.parametrize("lattice", [False, True])
@pytest.mark.parametrize("size", [(1.5, 2.5, 4.0), (2.5, 1.0, 4.0), (0.0, 2.5, 5.0)])
def test_cone_scene_bitexact(om, lattice, size):
    d, p = F.config_small_cone()
    d = dict(d, recv_size=np.float32(size))
    if lattice:
        p = F.with_lattice(p)
    print(f"cone {size} lattice={lattice}", flush=True)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    got = _maps(g)
    orc = o.trace_f32(records=True)
    assert np.array_equal(got["counts"].view(np.uint32), orc["counts"])
    assert np.array_equal(got["energy"].view(np.uint64), orc["energy"])
    assert int(orc["counts"].sum()) > 10_000
    scale, total = o.flux_scale()
    raw = oracle.energy_to_float(orc["energy"])
    assert np.array_equal(got["flux"].view(np.uint32), (raw * np.float32(scale)).astype(np.float32).view(np.uint32))
    gr = g.trace_records()
    for k in ("shade", "block", "pixel"):
        assert np.array_equal(gr[k], orc[k]), k


def test_cone_in_a_batch_and_as_a_fork(om):
    import torch
    d, p = F.config_small_cone()
    cone = om.SolarScene(d, dict(p, scene_index=0))
    # a fork keeps the root's microsamples: with the root's scene index it is the fresh scene
    same = cone.fork(F.sun_direction(35.0, 150.0), 900.0, scene_index=0)
    fresh = om.SolarScene(dict(d, sun_dir=F.sun_direction(35.0, 150.0), dni=np.float32(900.0)), dict(p, scene_index=0))
    f, g = _maps(fresh), _maps(same)
    for k in f:
        assert np.array_equal(g[k], f[k]), k
    fork = cone.fork(F.sun_direction(35.0, 150.0), 900.0, scene_index=1)
    dc, pc = F.config_small_cylinder()
    ds, ps = F.config_small()
    scenes = [cone, fork, om.SolarScene(dc, dict(pc, scene_index=2)), om.SolarScene(ds, dict(ps, scene_index=3))]
    want = [_maps(s) for s in scenes]
    b = om.Batch(scenes, workgroup=256, rays_per_thread=2)
    maps = b.run()
    torch.cuda.synchronize()
    for i, m in enumerate(maps):
        for k in want[i]:
            got = getattr(m, k).cpu().numpy()
            assert np.array_equal(got, want[i][k]), (i, k)
