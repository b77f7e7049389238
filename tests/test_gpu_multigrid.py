"""Multi-grid fields on the GPU against the oracle's grids (bit-exact).

A two-grid field (two sub-fields, each in a grid of its own) is traced with every grid
(OM_GRIDS_ALL: each heliostat's rays walk the grid that holds it) and with one grid
selected (the reference's trace is grid 0's heliostats over grid 0, Gtest.cu:12,21-23),
in every accumulation mode, from arrays and from a .scn file the library wrote, in a batch,
as annual forks, and in the differentiable mode.  The oracle's grids are pinned against its
single-grid path on the CPU (tests/test_multigrid.py).
"""
import numpy as np
import pytest

import oracle
from onemirror_amd import fields as F

pytestmark = pytest.mark.gpu

N_EACH = 50


@pytest.fixture(scope="module")
def om():
    import onemirror_amd
    return onemirror_amd


def _maps(scene):
    import torch
    m = scene.trace_plain()
    torch.cuda.synchronize()
    return {k: getattr(m, k).cpu().numpy().copy() for k in ("counts", "energy", "flux", "flux_smooth")}


def _bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _same_maps(a, b, label):
    for k in a:
        assert np.array_equal(_bits(a[k]), _bits(b[k])), (label, k)


def _oracle_equal(got, orc, label):
    assert np.array_equal(got["counts"].view(np.uint32), orc["counts"]), label
    assert np.array_equal(got["energy"].view(np.uint64), orc["energy"]), label


@pytest.fixture(scope="module")
def two():
    return F.config_two_grids(n_each=N_EACH)


@pytest.mark.parametrize("lattice", [False, True])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_all_grids_bitexact(om, two, lattice, mode):
    d, p = two
    p = dict(F.with_lattice(p) if lattice else p, grid_select=om.GRIDS_ALL, accumulate=mode)
    print(f"two grids lattice={lattice} accumulate={mode}", flush=True)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    assert g.info["n_grids"] == 2 and g.info["total_rays"] == o.total_rays
    got = _maps(g)
    orc = o.trace_f32()
    _oracle_equal(got, orc, "all grids")
    assert int(orc["counts"].sum()) > 100_000
    scale, total = o.flux_scale()
    flux = (oracle.energy_to_float(orc["energy"]) * np.float32(scale)).astype(np.float32)
    assert np.array_equal(got["flux"].view(np.uint32), flux.view(np.uint32))


def test_all_grids_records_bitexact(om, two):
    d, p = two
    p = dict(p, grid_select=om.GRIDS_ALL)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    for h0, h1 in ((0, 2 * N_EACH), (N_EACH - 3, N_EACH + 4), (N_EACH, 2 * N_EACH)):
        print(f"records [{h0}, {h1})", flush=True)
        gr, orc = g.trace_records(h0, h1), o.trace_f32(h0, h1, records=True)
        for k in ("shade", "block", "pixel"):
            assert np.array_equal(gr[k], orc[k]), (h0, h1, k)
    gr = g.trace_records()
    assert (gr["shade"] >= 0).sum() > 0 and (gr["block"] >= 0).sum() > 0
    # blockers of each sub-field's rays are its own heliostats
    off = g.ray_offset()
    for k in ("shade", "block"):
        first, second = gr[k][: off[N_EACH]], gr[k][off[N_EACH]:]
        assert first.max() < N_EACH and second[second >= 0].min() >= N_EACH


@pytest.mark.parametrize("k", [0, 1])
def test_selected_grid_is_its_sub_field(om, two, k):
    d, p = two
    sub = dict(F.subset(d, slice(k * N_EACH, (k + 1) * N_EACH)))
    g0 = d["grids"][k]
    sub.update(grid_pos=g0["pos"], grid_size=g0["size"], grid_interval=g0["interval"])
    sub.pop("grids")
    g, o = om.SolarScene(d, dict(p, grid_select=k)), oracle.OracleScene(sub, p)
    assert g.info["n_helio"] == N_EACH
    _oracle_equal(_maps(g), o.trace_f32(), f"grid {k}")


def test_scn_file_traces_as_the_arrays(om, two, tmp_path):
    """The field written as .scn and read back traces bit for bit as the arrays it was
    written from (the scene file is the reference's input format)."""
    d, p = two
    path = str(tmp_path / "two.scn")
    om.scn_write(path, d)
    p = dict(p, grid_select=om.GRIDS_ALL)
    a = om.SolarScene(d, p)
    # the .scn format carries no receiver pixel length (the reference's is the global
    # solarenergy::receiver_pixel_length): take the description's
    desc = om.scn_parse(path)
    desc.update(sun_dir=d["sun_dir"], dni=d["dni"], recv_pixel_length=d["recv_pixel_length"])
    c = om.SolarScene(desc, p)
    _same_maps(_maps(a), _maps(c), "scn")
    b = om.SolarScene.from_scn(path, d["sun_dir"], float(d["dni"]), p)
    assert b.info["n_grids"] == 2 and b.info["total_rays"] == a.info["total_rays"]


def test_batch_with_a_multigrid_scene_equals_plain(om, two):
    import torch
    d, p = two
    scenes = [om.SolarScene(d, dict(p, grid_select=om.GRIDS_ALL, scene_index=0))]
    scenes.append(scenes[0].fork(F.sun_direction(30.0, 140.0), 900.0, scene_index=1))
    ds, ps = F.config_small()
    scenes.append(om.SolarScene(ds, dict(ps, scene_index=2)))
    scenes.append(om.SolarScene(d, dict(p, grid_select=1, scene_index=3)))
    want = [_maps(s) for s in scenes]
    b = om.Batch(scenes, workgroup=256, rays_per_thread=2)
    for rep in range(2):
        maps = b.run()
        torch.cuda.synchronize()
        for i, m in enumerate(maps):
            _same_maps({k: getattr(m, k).cpu().numpy() for k in want[i]}, want[i], f"scene {i} replay {rep}")


def test_fork_of_a_multigrid_scene_equals_fresh(om, two):
    d, p = two
    p = dict(p, grid_select=om.GRIDS_ALL)
    root = om.SolarScene(d, p)
    sun = F.sun_direction(35.0, 205.0)
    f = root.fork(sun, 800.0, scene_index=0)
    fresh = om.SolarScene(dict(d, sun_dir=sun, dni=np.float32(800.0)), p)
    _same_maps(_maps(f), _maps(fresh), "fork")


def test_two_grids_with_sector_windows(om, two):
    """Both grids around a cylinder whose map is larger than one LDS window (320 x 160): each
    grid's window trace gives its queues that grid's azimuth sectors.  The maps equal the
    oracle's bit for bit, for the root scene and for a fork at another sun (a fork uses its
    root's sector order)."""
    d, p = two
    d = dict(d, recv_type=1, recv_pos=np.array([0.0, 30.0, 0.0], np.float32),
             recv_size=np.array([2.0, 4.0, 0.0], np.float32), recv_res_x=320, recv_res_y=160)
    p = dict(p, grid_select=om.GRIDS_ALL, accumulate=om.ACC_WINDOW)
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    assert g.info["n_grids"] == 2
    orc = o.trace_f32()
    _oracle_equal(_maps(g), orc, "two grids, sector windows")
    assert int(orc["counts"].sum()) > 100_000
    sun = F.sun_direction(35.0, 205.0)
    f = g.fork(sun, 800.0, scene_index=0)
    d2 = dict(d, sun_dir=sun, dni=np.float32(800.0))
    _oracle_equal(_maps(f), oracle.OracleScene(d2, p).trace_f32(), "fork, sector windows")


def test_diff_forward_and_gradients_over_two_grids(om, two):
    """The differentiable mode walks each heliostat's grid for its visibility ballots.  At
    zero parameters its rays are the plain trace's up to the rounding of one more
    normalisation of the aimed normal, so its ballots agree with the oracle's records: a
    visible ray is neither shaded nor blocked, and a ray that reaches the receiver unshaded
    and unblocked is visible (all but a few rays on a boundary).  The splat matches the FP64
    model and the adjoint's gradients match FP64 central differences for heliostats of both
    grids."""
    import torch
    import onemirror_amd.diff as D
    from test_gpu_diff import _fd_check, _theta
    d, p = two
    p = F.with_lattice(dict(p, grid_select=om.GRIDS_ALL))
    g, o = om.SolarScene(d, p), oracle.OracleScene(d, p)
    H = 2 * N_EACH
    _, _, mask = D.forward(g, torch.zeros((H, 4), dtype=torch.float32, device="cuda"))
    torch.cuda.synchronize()
    vis = D.visibility(g, mask).astype(bool)
    rec = o.trace_f32(records=True)
    clear = (rec["shade"] < 0) & (rec["block"] < 0)
    n = vis.size
    assert vis.sum() > 100_000
    assert (vis & ~clear).sum() <= 1e-4 * n, (vis & ~clear).sum()
    assert ((rec["pixel"] >= 0) & clear & ~vis).sum() <= 1e-4 * n
    assert clear.sum() < n  # the field shades and blocks
    th = _theta(H, 3)
    S, E, mask = D.forward(g, torch.tensor(th, device="cuda"))
    torch.cuda.synchronize()
    vis = D.visibility(g, mask)
    S64, E64 = o.diff_splat64(th.astype(np.float64), vis)
    assert abs(float(E) - E64) < 1e-5 * E64
    assert np.abs(S.cpu().numpy() - S64).max() < 5e-4 * S64.max()
    _fd_check(o, D, g, th, [3, N_EACH - 1, N_EACH, H - 5])
