"""Statistical agreement with the reference's own sampling scheme.

The reference draws rays from two pools (groups x lights directions, rebuilt per
heliostat) with a random pool start per microsample and takes the perturbation of
microsample m-1 (scene_instance_process.cu:25-226, rectgrid_raytracing.cu:171-185).  Its
cuRAND streams are closed and time-seeded, so the comparison is statistical: total
intercepted power from independent Philox per-ray draws and from the pooled scheme (driven
by an independent Philox key) must agree within 3 sigma, sigma from repeated runs."""
import numpy as np

import oracle
from onemirror_amd import fields as F


def test_total_power_per_ray_vs_pooled_within_3_sigma():
    d, p = F.config_small()
    p = F.with_lattice(p)
    d = dict(d)
    d["recv_size"] = np.float32([3.0, 3.0, 0.0])   # a receiver that clips: power depends on the sampling
    d["recv_pixel_length"] = np.float32(3.0 / 48)
    per_ray, pooled = [], []
    for seed in range(6):
        pp = dict(p)
        pp["seed"] = 1000 + seed
        s = oracle.OracleScene(d, pp)
        per_ray.append(float(s.trace_f32()["energy"].sum()) * 2.0 ** -31)
        pooled.append(float(s.trace_pooled(groups=128, lights=p["rays_per_sample"],
                                           pool_seed=77 + seed).sum()))
    a, b = np.array(per_ray), np.array(pooled)
    sigma = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    assert sigma > 0
    assert abs(a.mean() - b.mean()) < 3 * sigma + 1e-9 * a.mean(), (a.mean(), b.mean(), sigma)
    # and the clipping receiver actually makes the result sampling-dependent
    assert 0.2 < a.mean() / (s.total_rays) < 0.99
