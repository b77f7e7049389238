"""Annual time steps on the GPU: a fork of a field under another sun traces exactly like a
scene built from scratch for that sun (maps bit-exact, Poisson and lattice microsamples,
rectangle and cylinder receivers), forks of forks share the root, and the annual driver's
batched steps equal plain traces of the same forks, with the intercept ratios of
results_with_given_normal."""
import numpy as np
import pytest

from onemirror_amd import fields as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def om():
    import onemirror_amd
    return onemirror_amd


def _maps(m):
    import torch
    torch.cuda.synchronize()
    return (m.counts.cpu().numpy().view(np.uint32).copy(), m.energy.cpu().numpy().view(np.uint64).copy(),
            m.flux.cpu().numpy().view(np.uint32).copy())


SUNS = [((0.3, -0.8, 0.52), 850.0), ((-0.55, -0.6, -0.3), 620.0)]


@pytest.mark.parametrize("cfg,lattice", [("small", False), ("small", True), ("mid_slice", False),
                                         ("cylinder", False)])
def test_fork_equals_fresh_scene(om, cfg, lattice):
    if cfg == "small":
        d, p = F.config_small()
    elif cfg == "cylinder":
        d, p = F.config_small_cylinder()
    else:
        d, p = F.config_mid()
        d = F.subset(d, slice(1500, 1700))
    if lattice:
        p = F.with_lattice(p)
    root = om.SolarScene(d, p)
    for sun, dni in SUNS:
        f = root.fork(sun, dni)
        fresh = om.SolarScene(dict(d, sun_dir=np.float32(sun), dni=np.float32(dni)), p)
        assert f.info["total_rays"] == fresh.info["total_rays"]
        assert f.info["total_energy"] == fresh.info["total_energy"]
        a, b = _maps(f.trace_plain()), _maps(fresh.trace_plain())
        for x, y in zip(a, b):
            assert np.array_equal(x, y), (cfg, sun)
        assert a[0].sum() > 0
        ff = f.fork(SUNS[0][0], SUNS[0][1])   # a fork of a fork shares the root
        assert ff.root is root


def test_fork_scene_index_changes_rays_not_microsamples(om):
    d, p = F.config_small()
    root = om.SolarScene(d, p)
    f0, f1 = root.fork(SUNS[0][0], SUNS[0][1], 0), root.fork(SUNS[0][0], SUNS[0][1], 7)
    assert np.array_equal(f0.export("micro_count", 4 * f0.info["n_helio"], np.int32),
                          f1.export("micro_count", 4 * f1.info["n_helio"], np.int32))
    a, b = _maps(f0.trace_plain()), _maps(f1.trace_plain())
    assert not np.array_equal(a[1], b[1])
    # the totals of independent streams agree statistically
    assert abs(int(a[0].sum()) - int(b[0].sum())) < 0.02 * a[0].sum()


@pytest.mark.parametrize("mode", ["batch", "plain"])
def test_annual_run_equals_plain_traces(om, mode):
    from onemirror_amd import annual as A
    d, p = F.config_small()
    root = om.SolarScene(d, p)
    sched = [(3, 1, 8), (3, 1, 10), (3, 1, 12), (6, 16, 14), (12, 4, 16), (12, 4, 18)]
    steps = A.annual_steps(sched)
    assert 3 <= len(steps) <= 6
    run = A.AnnualRun(root, steps, batch_size=4, mode=mode)
    res = run.run(keep_maps=True)
    assert len(res["steps"]) == len(steps)
    L = om.lib()
    for i, (st, f) in enumerate(zip(steps, run.forks)):
        m = f.trace_plain()
        a = _maps(m)
        b = _maps(run.maps()[i])
        for x, y in zip(a, b):
            assert np.array_equal(x, y), i
        flux = m.flux.cpu().numpy().astype(np.float32).reshape(f.res_y, f.res_x)
        want = om.intercept_ratio(flux, f.info["total_energy"])
        assert res["steps"][i]["ratio"] == pytest.approx(want, rel=0, abs=0)
        assert 0.0 < res["steps"][i]["ratio"] <= 1.0
    assert 0.0 < res["annual_ratio"] <= 1.0


@pytest.mark.parametrize("cfg", ["small", "mid_slice"])
def test_fork_with_own_stream_equals_oracle(om, cfg):
    """A fork under another sun, on a Philox stream of its own (scene_index != 0), against
    the oracle built for that sun with the root's microsamples and the fork's ray stream
    (main.cpp:56-116 builds each time step's scene on the same field): per-ray records and
    integer maps bit-exact."""
    import oracle
    import torch
    if cfg == "small":
        d, p = F.config_small()
    else:
        d, p = F.config_mid()
        d = F.subset(d, slice(1500, 1650))
    root = om.SolarScene(d, p)
    for k, (sun, dni) in enumerate(SUNS):
        idx = 5 + k
        f = root.fork(sun, dni, idx)
        o = oracle.OracleScene(dict(d, sun_dir=np.float32(sun), dni=np.float32(dni)), p)
        o.set_ray_stream(idx)
        want = o.trace_f32(records=True)
        rec = f.trace_records()
        for key in ("shade", "block", "pixel"):
            assert np.array_equal(rec[key], want[key]), (cfg, k, key)
        m = f.trace_plain()
        torch.cuda.synchronize()
        assert np.array_equal(m.counts.cpu().numpy().view(np.uint32).ravel(), want["counts"].ravel()), (cfg, k)
        assert np.array_equal(m.energy.cpu().numpy().view(np.uint64).ravel(), want["energy"].ravel()), (cfg, k)
        assert want["counts"].sum() > 0


def test_annual_batch_steps_equal_oracle(om):
    """Every time step of a batched annual run (forks i = 0.. with stream i) against the
    oracle of that step's sun and DNI: integer maps bit-exact, and the flux scale (DNI) of
    the step."""
    import oracle
    from onemirror_amd import annual as A
    d, p = F.config_small()
    root = om.SolarScene(d, p)
    steps = A.annual_steps([(3, 1, 8), (6, 16, 12), (9, 10, 16), (12, 4, 10)])
    run = A.AnnualRun(root, steps, batch_size=4, mode="batch")
    run.run(keep_maps=True)
    for i, st in enumerate(steps):
        o = oracle.OracleScene(dict(d, sun_dir=np.float32(st["sun_dir"]), dni=np.float32(st["dni"])), p)
        o.set_ray_stream(i)
        want = o.trace_f32()
        c, e, _ = _maps(run.maps()[i])
        assert np.array_equal(c.ravel(), want["counts"].ravel()), i
        assert np.array_equal(e.ravel(), want["energy"].ravel()), i
        assert run.forks[i].info["flux_scale"] == o.flux_scale()[0], i
