"""Differentiable mode on the GPU against the oracle's FP64 model.

* forward: the splat map and energy of the GPU (FP32 rays, fixed-point sums) against the
  FP64 model of the same rays with the GPU's visibility;
* backward: the adjoint kernel's dL/dtheta against FP64 central finite differences of the
  same loss (aim offsets stepped by 1e-4 m, tracking deltas by 1e-4 mrad), visibility held
  fixed as the model defines it.
"""
import numpy as np
import pytest

import oracle
from onemirror_amd import fields as F

pytestmark = pytest.mark.gpu

STEP = 1e-4
LAM = 1.0


@pytest.fixture(scope="module")
def D():
    import onemirror_amd.diff as D
    return D


def _theta(H, seed):
    rng = np.random.default_rng(seed)
    th = np.zeros((H, 4), np.float32)
    th[:, :2] = rng.normal(0.0, 0.3, (H, 2))   # metres
    th[:, 2:] = rng.normal(0.0, 0.3, (H, 4 - 2))  # milliradians
    return th


def _fd_check(o, D, scene, th, heliostats):
    """GPU gradient (autograd through the adjoint kernel) against FP64 central differences."""
    import torch
    H = th.shape[0]
    p = torch.tensor(th, device="cuda", requires_grad=True)
    S, E = D.splat(scene, p)
    tau = float(0.05 * S.detach().max())
    L = D.spill_peak_loss(S, E, tau, LAM)
    L.backward()
    g = p.grad.cpu().numpy().astype(np.float64)
    assert np.isfinite(g).all(), np.where(~np.isfinite(g).all(1))[0][:8]
    _, _, mask = D.forward(scene, p)
    vis = D.visibility(scene, mask)
    off = o.micro()["ray_offset"]
    S64, E64 = o.diff_splat64(th.astype(np.float64), vis)
    # the GPU loss and the FP64 loss at theta agree
    L64 = D.spill_peak_loss_np(S64, E64, tau, LAM)
    assert abs(float(L) - L64) < 1e-5 * abs(L64) + 1e-9
    fd = np.zeros((len(heliostats), 4))
    for a, j in enumerate(heliostats):
        vj = vis[off[j] - off[0]: off[j + 1] - off[0]]
        Sj, Ej = o.diff_splat64(th.astype(np.float64), vj, j, j + 1)
        for k in range(4):
            vals = []
            for sgn in (1.0, -1.0):
                t2 = th.astype(np.float64)
                t2[j, k] += sgn * STEP
                S2, E2 = o.diff_splat64(t2, vj, j, j + 1)
                vals.append(D.spill_peak_loss_np(S64 - Sj + S2, E64 - Ej + E2, tau, LAM))
            fd[a, k] = (vals[0] - vals[1]) / (2 * STEP)
    got = g[heliostats]
    # The GPU evaluates each ray in FP32 (positions within ~3e-4 px of the FP64 rays) and a
    # heliostat's gradient is a sum over ~40,000 rays with cancellation, so it carries an
    # error of a few 1e-3 of the largest gradient of that parameter.
    for k in range(4):
        scale = np.abs(fd[:, k]).max()
        assert scale > 0, k
        err = np.abs(got[:, k] - fd[:, k])
        assert err.max() < 5e-3 * scale, (k, err.max(), scale, got[:, k][:4], fd[:, k][:4])
        assert np.median(err) < 1e-3 * scale, (k, np.median(err), scale)
    return g, fd


@pytest.fixture(scope="module")
def sub(om_lib):
    import onemirror_amd as om
    d, p = F.config_mid()
    d = F.subset(d, slice(1000, 1300))
    return om.SolarScene(d, p), oracle.OracleScene(d, p)


@pytest.fixture(scope="module")
def om_lib():
    import onemirror_amd
    return onemirror_amd


def test_forward_matches_fp64_model(D, sub):
    import torch
    scene, o = sub
    th = _theta(o.n_helio, 1)
    S, E, mask = D.forward(scene, torch.tensor(th, device="cuda"))
    torch.cuda.synchronize()
    vis = D.visibility(scene, mask)
    assert 0.3 < vis.mean() < 1.0
    S64, E64 = o.diff_splat64(th.astype(np.float64), vis)
    S = S.cpu().numpy()
    assert abs(float(E) - E64) < 1e-5 * E64
    # FP32 ray geometry: positions off by ~3e-4 px at 500 m move single-ray splats slightly
    assert np.abs(S - S64).max() < 5e-4 * S64.max()
    assert S.sum() > 0.5 * E64


def test_forward_is_deterministic_and_backward_repeats(D, sub):
    import torch
    scene, o = sub
    p = torch.tensor(_theta(o.n_helio, 2), device="cuda")
    S1, E1, m1 = D.forward(scene, p)
    S2, E2, m2 = D.forward(scene, p)
    assert torch.equal(S1, S2) and torch.equal(E1, E2) and torch.equal(m1, m2)
    gS = torch.randn_like(S1)
    gE = torch.tensor([0.3], dtype=torch.float64, device="cuda")
    g1 = D.backward(scene, p, m1, gS, gE)
    g2 = D.backward(scene, p, m1, gS, gE)
    assert torch.equal(g1, g2)


def test_gradient_matches_fp64_central_differences(D, sub):
    scene, o = sub
    th = _theta(o.n_helio, 3)
    rng = np.random.default_rng(4)
    _fd_check(o, D, scene, th, sorted(rng.choice(o.n_helio, 16, replace=False).tolist()))


def test_mid_field_gradient_on_32_heliostats(D, om_lib):
    """The differentiable workload: the 5,000-heliostat field, 4 parameters per heliostat,
    FP64 central differences on 32 random heliostats (seed 4)."""
    d, p = F.config_mid()
    scene, o = om_lib.SolarScene(d, p), oracle.OracleScene(d, p)
    th = np.zeros((o.n_helio, 4), np.float32)
    rng = np.random.default_rng(4)
    _fd_check(o, D, scene, th, sorted(rng.choice(o.n_helio, 32, replace=False).tolist()))
