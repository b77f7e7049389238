"""The allocation-free trace (om_trace_plain_noalloc) on the GPU, through the C ABI.

Pin: its maps (counts, energy, flux, smoothed flux) equal om_trace_plain's bit for bit, on
calls in a row that share the static per-GPU workspace: a 256 x 256 map in window mode,
then a smaller 128 x 128 map in packed mode (the first call's contents must not leak into
the second), then the first scene again.  A map whose workspace does not fit the static
array is refused before anything is launched: its output maps keep their contents.
"""
import numpy as np
import pytest

from onemirror_amd import fields as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def om():
    import onemirror_amd
    return onemirror_amd


def _host(maps):
    return {k: getattr(maps, k).cpu().numpy().copy() for k in ("counts", "energy", "flux", "flux_smooth")}


def _same(got, want, label):
    for k, v in want.items():
        g = got[k]
        assert np.array_equal(g.view(np.uint32) if g.dtype == np.float32 else g,
                              v.view(np.uint32) if v.dtype == np.float32 else v), (label, k)


def test_noalloc_equals_plain_on_calls_in_a_row(om):
    import torch
    dm, pm = F.config_mid()
    big = om.SolarScene(F.subset(dm, slice(0, 300)), dict(pm, accumulate=om.ACC_WINDOW))
    ds, ps = F.config_small()
    small = om.SolarScene(ds, dict(ps, accumulate=om.ACC_PACKED))
    assert small.info["res_x"] * small.info["res_y"] < big.info["res_x"] * big.info["res_y"]
    assert small.info["total_rays"] < big.info["total_rays"]
    want_big = _host(big.trace_plain())
    want_small = _host(small.trace_plain())
    torch.cuda.synchronize()
    assert want_big["counts"].sum() > 0 and want_small["counts"].sum() > 0
    for label, scene, want in (("256x256 window", big, want_big), ("128x128 packed", small, want_small),
                               ("256x256 window again", big, want_big)):
        print(f"noalloc call: {label}", flush=True)
        maps = scene.new_maps()
        scene.trace_plain_noalloc(maps)
        torch.cuda.synchronize()
        _same(_host(maps), want, label)


def test_noalloc_refuses_a_map_beyond_the_scratch(om):
    import torch
    d, p = F.config_small()
    d = dict(d, recv_pixel_length=np.float32(8.0 / 1100))
    s = om.SolarScene(d, dict(p, accumulate=om.ACC_PACKED))
    assert s.info["res_x"] * s.info["res_y"] >= 1024 * 1024
    maps = s.new_maps()
    maps.counts.fill_(7)
    torch.cuda.synchronize()
    with pytest.raises(om.OMError, match="workspace"):
        s.trace_plain_noalloc(maps)
    torch.cuda.synchronize()
    assert int((maps.counts != 7).sum()) == 0
