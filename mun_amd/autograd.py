"""torch.autograd boundary of the SEM solver.

``SEMLikelihoodAdjointFn`` keeps the reference's signature and behaviour
(Bspline_Coupling/RQS_Coupling/RQS_Coupling.py:295-351): forward returns the
scalar log-likelihood of the control-point offsets z, backward multiplies the
upstream gradient into the adjoint gradient saved in forward.

``BatchedLikelihoodFn`` takes a whole batch of parameter vectors (ELBO
samples, MCMC chains) and all shots at once and returns one log-likelihood
per sample; its backward is the batched adjoint gradient.  It replaces the
reference drivers' per-sample loop and per-source thread pool
(CO2_plume/one_stage/single_cubic/RQS_NF.py:987-1098, 1488-1492).
"""
from __future__ import annotations

import json

import numpy as np
import torch

from .batch import BatchSolver
from .simulation import SEMSimulation

_SIM_CACHE = {}


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
    from cfg['velocity']['control_points']."""

    @staticmethod
    def forward(ctx, z, y_obs, noise_std, sem_cfg_dict, obs_dt, obs_nt):
        if z.dim() != 1:
            raise ValueError("SEMLikelihoodAdjointFn expects z shape (12,), got {}".format(tuple(z.shape)))
        sim = get_or_make_sim(sem_cfg_dict)
        base = np.asarray(sem_cfg_dict["velocity"]["control_points"], dtype=np.float64).reshape(6, 2)
        abs_ctrl = base + z.detach().cpu().numpy().reshape(-1, 2)
        ns = float(noise_std.detach().cpu().item()) if isinstance(noise_std, torch.Tensor) else float(noise_std)
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
        return grad_output * g, None, None, None, None, None


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
