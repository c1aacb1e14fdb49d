"""torch.autograd boundary of the SEM solver.

``SEMLikelihoodAdjointFn`` keeps the reference's signature and behaviour
(Bspline_Coupling/RQS_Coupling/RQS_Coupling.py:295-351): forward returns the
scalar log-likelihood of the control-point offsets z, backward multiplies the
upstream gradient into the adjoint gradient saved in forward.

Both drivers pass the solver sqrt(noise_std^2 + EPS), not noise_std: EPS =
1e-12 in Bspline_Coupling/RQS_Coupling/RQS_Coupling.py:67-70 (used at
:323-333), 1e-30 by default in CO2_plume/one_stage/single_cubic/RQS_NF.py:161
(used at :1009-1013).  The Functions here take the same EPS.

``MultiSourceLikelihoodFn`` is the CO2 driver's multi-source Function
(RQS_NF.py:987-1098) over a whole batch of samples: z [S, 2K] in, float32
log-likelihoods [S] out (a 1-D z gives the reference's 0-d result), every
(sample, source) solve in one batched call (``likelihood.MultiSourceLikelihood``).

``BatchedLikelihoodFn`` takes a whole batch of tensor B-spline node values
and all shots at once and returns one log-likelihood per sample; its backward
is the batched adjoint gradient.  Both replace the reference drivers'
per-sample loop and per-source thread pool (RQS_NF.py:1035-1072, 1488-1492).
"""
from __future__ import annotations

import json

import numpy as np
import torch

from .batch import BatchSolver
from .likelihood import MultiSourceLikelihood
from .simulation import SEMSimulation

_SIM_CACHE = {}
_MULTI_CACHE = {}

# EPS added to noise_std^2 by the drivers before calling the solver
EPS_BSPLINE_DRIVER = 1e-12   # Bspline_Coupling/RQS_Coupling/RQS_Coupling.py:67-70
EPS_CO2_DRIVER = 1e-30       # CO2_plume/one_stage/single_cubic/RQS_NF.py:161-162


def effective_noise_std(noise_std, eps: float) -> float:
    """sqrt(noise_std^2 + eps), the noise_std the drivers hand the solver
    (RQS_Coupling.py:323-333, RQS_NF.py:1009-1013)."""
    ns = float(noise_std.detach().cpu().item()) if isinstance(noise_std, torch.Tensor) else float(noise_std)
    return float(np.sqrt(ns ** 2 + float(eps)))


def _cfg_key(cfg: dict) -> str:
    def enc(o):
        if isinstance(o, np.ndarray):
            return o.tolist()
        if isinstance(o, torch.Tensor):
            return o.detach().cpu().tolist()
        return str(o)
    return json.dumps(cfg, sort_keys=True, default=enc)


def get_or_make_sim(cfg: dict, device=None) -> SEMSimulation:
    """One cached simulation per distinct config (as get_or_make_sim,
    RQS_Coupling.py:28-66); the cache key covers the whole config."""
    key = (_cfg_key(cfg), str(device))
    sim = _SIM_CACHE.get(key)
    if sim is None:
        sim = SEMSimulation(cfg, device=device)
        _SIM_CACHE[key] = sim
    return sim


class SEMLikelihoodAdjointFn(torch.autograd.Function):
    """loglik(z) with z the (12,) offsets of the 6 closed-curve control points
    from cfg['velocity']['control_points']; the solver sees
    sqrt(noise_std^2 + eps), eps = 1e-12 as in RQS_Coupling.py:319-333."""

    @staticmethod
    def forward(ctx, z, y_obs, noise_std, sem_cfg_dict, obs_dt, obs_nt, eps=EPS_BSPLINE_DRIVER):
        if z.dim() != 1:
            raise ValueError("SEMLikelihoodAdjointFn expects z shape (12,), got {}".format(tuple(z.shape)))
        sim = get_or_make_sim(sem_cfg_dict)
        base = np.asarray(sem_cfg_dict["velocity"]["control_points"], dtype=np.float64).reshape(6, 2)
        abs_ctrl = base + z.detach().cpu().numpy().reshape(-1, 2)
        ns = effective_noise_std(noise_std, eps)
        out = sim.run_forward_and_adjoint({
            "bspline_ctrl": abs_ctrl,
            "y_obs": y_obs.detach().cpu().numpy() if isinstance(y_obs, torch.Tensor) else np.asarray(y_obs),
            "obs_dt": float(obs_dt), "obs_nt": int(obs_nt), "noise_std": ns})
        g = torch.as_tensor(np.asarray(out["grad_wrt_ctrl"], dtype=np.float64).reshape(-1), device=z.device)
        ctx.save_for_backward(g)
        return torch.tensor(out["loglik"], dtype=torch.float64, device=z.device)

    @staticmethod
    def backward(ctx, grad_output):
        (g,) = ctx.saved_tensors
        return grad_output * g, None, None, None, None, None, None


def get_or_make_multi(cfg: dict, device=None) -> MultiSourceLikelihood:
    """One cached multi-source likelihood per distinct configuration."""
    key = (_cfg_key(cfg), str(device))
    ms = _MULTI_CACHE.get(key)
    if ms is None:
        ms = MultiSourceLikelihood(cfg, device=device)
        _MULTI_CACHE[key] = ms
    return ms


class MultiSourceLikelihoodFn(torch.autograd.Function):
    """The CO2 driver's SEMLikelihoodAdjointFn (RQS_NF.py:987-1098) for a
    batch: z [S, 2K] offsets from the configuration's anomaly control points
    (or [2K] for one sample), y_obs (rows, nsrc * nrec) with the sources'
    receiver blocks side by side; returns the float32 log-likelihood summed
    over sources per sample ([S], or 0-d for a 1-D z).  Backward multiplies the
    upstream gradient into the float32 control-point gradient summed over
    sources, as the reference's backward does."""

    @staticmethod
    def forward(ctx, z, y_obs, noise_std, sem_cfg_dict, obs_dt, obs_nt, eps=EPS_CO2_DRIVER):
        if z.dim() not in (1, 2):
            raise ValueError(f"MultiSourceLikelihoodFn expects z of shape (2K,) or (S, 2K), got {tuple(z.shape)}")
        ms = get_or_make_multi(sem_cfg_dict)
        ll, grad, _ = ms.evaluate(z.detach(), y_obs, effective_noise_std(noise_std, eps), obs_dt, obs_nt)
        g = grad.to(device=z.device, dtype=torch.float32)
        out = torch.as_tensor(ll, dtype=torch.float32).to(z.device)
        if z.dim() == 1:
            g, out = g[0], out[0]
        ctx.save_for_backward(g)
        return out

    @staticmethod
    def backward(ctx, grad_output):
        (g,) = ctx.saved_tensors
        go = grad_output.to(g.dtype)
        return (go[..., None] * g if g.dim() == 2 else go * g), None, None, None, None, None, None


class BatchedLikelihoodFn(torch.autograd.Function):
    """loglik[s] for a batch of B-spline node values coef [S, nbz, nbx]
    (summed over the solver's shots); backward gives d loglik[s] / d coef[s]."""

    @staticmethod
    def forward(ctx, coef, solver: BatchSolver, obs, noise_std: float):
        ll, grad, _ = solver.loglik_grad(coef.detach(), obs, noise_std, want_grad=True)
        ctx.save_for_backward(grad.to(coef.device))
        return ll.to(device=coef.device, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_output):
        (g,) = ctx.saved_tensors
        return grad_output.reshape(-1, 1, 1).to(g.dtype) * g, None, None, None


def batched_loglik(coef: torch.Tensor, solver: BatchSolver, obs, noise_std: float) -> torch.Tensor:
    return BatchedLikelihoodFn.apply(coef, solver, obs, noise_std)
