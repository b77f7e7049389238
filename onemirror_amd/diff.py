"""Differentiable mode: the Gaussian-kernel flux splat of a field as a function of
per-heliostat parameters, behind a torch.autograd.Function whose backward is the library's
hand-written adjoint kernel (om_diff_forward / om_diff_backward).

Parameters, one row per heliostat: (a_u, a_v, t_u, t_v) = aim-point offset in metres along
the receiver's two edges and tracking-angle deltas in milliradians about the heliostat's
frame axes; all zero is the nominal aiming.  The forward pass returns the splat map S
(energy units, float64) and the energy E of the visible rays; visibility (shading, blocking,
plane hit) is decided in the forward pass and held fixed in the backward pass, which
recomputes the rays from their Philox counters.  No reference counterpart: the reference
only traces fixed aims.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _check, _stream_ptr, _vp, lib


def mask_words(scene) -> int:
    n = ctypes.c_int64()
    _check(lib().om_diff_mask_words(scene._h, ctypes.byref(n)), "om_diff_mask_words")
    return n.value


def forward(scene, params: torch.Tensor):
    """(S, E, mask) without autograd: S [res_y, res_x] float64, E [1] float64, mask the
    visibility ballots (int64 words) the backward pass needs."""
    params = params.detach().to(device="cuda", dtype=torch.float32).contiguous()
    H = scene.info["n_helio"]
    if params.shape != (H, 4):
        raise ValueError(f"params must be [{H}, 4]")
    mask = torch.zeros(max(1, mask_words(scene)), dtype=torch.int64, device="cuda")
    S = torch.empty((scene.res_y, scene.res_x), dtype=torch.float64, device="cuda")
    E = torch.empty(1, dtype=torch.float64, device="cuda")
    _check(lib().om_diff_forward(scene._h, _vp(params.data_ptr()), _vp(mask.data_ptr()),
                                 _vp(S.data_ptr()), _vp(E.data_ptr()), _stream_ptr()),
           "om_diff_forward")
    return S, E, mask


def backward(scene, params: torch.Tensor, mask: torch.Tensor, gS: torch.Tensor, gE: torch.Tensor):
    """dL/dparams [H, 4] float32 from dL/dS and dL/dE (the adjoint kernel)."""
    params = params.detach().to(device="cuda", dtype=torch.float32).contiguous()
    gS = gS.to(device="cuda", dtype=torch.float64).contiguous()
    gE = gE.to(device="cuda", dtype=torch.float64).reshape(1).contiguous()
    grad = torch.empty_like(params)
    _check(lib().om_diff_backward(scene._h, _vp(params.data_ptr()), _vp(mask.data_ptr()),
                                  _vp(gS.data_ptr()), _vp(gE.data_ptr()), _vp(grad.data_ptr()),
                                  _stream_ptr()), "om_diff_backward")
    return grad


def visibility(scene, mask: torch.Tensor, h0: int = 0, h1=None) -> np.ndarray:
    """Visibility of every ray of heliostats [h0, h1) from the ballots, in ray order."""
    h1 = scene.info["n_helio"] if h1 is None else h1
    off = scene.ray_offset()
    n = int(off[h1] - off[h0])
    out = torch.zeros(max(1, n), dtype=torch.uint8, device="cuda")
    _check(lib().om_diff_visibility(scene._h, _vp(mask.data_ptr()), int(h0), int(h1),
                                    _vp(out.data_ptr()), n, _stream_ptr()), "om_diff_visibility")
    torch.cuda.synchronize()
    return out[:n].cpu().numpy()


class Splat(torch.autograd.Function):
    """(S, E) = splat(params); backward through the adjoint kernel."""

    @staticmethod
    def forward(ctx, params, scene):
        S, E, mask = forward(scene, params)
        ctx.scene = scene
        ctx.save_for_backward(params.detach(), mask)
        return S, E

    @staticmethod
    def backward(ctx, gS, gE):
        params, mask = ctx.saved_tensors
        scene = ctx.scene
        if gS is None:
            gS = torch.zeros((scene.res_y, scene.res_x), dtype=torch.float64, device="cuda")
        if gE is None:
            gE = torch.zeros(1, dtype=torch.float64, device="cuda")
        g = backward(scene, params, mask, gS, gE)
        return g.to(params.dtype), None


def splat(scene, params: torch.Tensor):
    return Splat.apply(params, scene)


def spill_peak_loss(S: torch.Tensor, E: torch.Tensor, tau: float, lam: float = 1.0) -> torch.Tensor:
    """Spillage plus a softmax peak-flux penalty, both as fractions of the reflected energy:
    L = (1 - sum(S) / E) + lam * tau * logsumexp(S / tau) / E."""
    e = E.reshape(())
    spill = 1.0 - S.sum() / e
    peak = tau * torch.logsumexp(S.reshape(-1) / tau, 0) / e
    return spill + lam * peak


def spill_peak_loss_np(S: np.ndarray, E: float, tau: float, lam: float = 1.0) -> float:
    """The same loss in numpy float64 (for checks against the FP64 model)."""
    z = S.reshape(-1) / tau
    mx = z.max()
    lse = mx + np.log(np.exp(z - mx).sum())
    return float((1.0 - S.sum() / E) + lam * tau * lse / E)
