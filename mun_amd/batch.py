"""Batched forward+adjoint solves of many velocity models and shots.

``BatchSolver`` binds a plan to one acquisition (shots + receivers +
wavelet) and a velocity parameterisation (tensor cubic B-spline nodes on the
domain, optionally over a fixed background).  One call advances every
(model, shot) pair together in the same kernel launches, replacing the
reference drivers' per-sample loop and per-source thread pool
(Bspline_Coupling/RQS_Coupling/RQS_Coupling.py:795-798,
CO2_plume/one_stage/single_cubic/RQS_NF.py:1035-1072, 1488-1492).
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .solver import Plan


class BatchSolver:
    def __init__(self, plan: Plan, sources: np.ndarray, receivers: np.ndarray, f0: float,
                 nbz: int, nbx: int, base: Optional[np.ndarray] = None, amp: float = 1.0):
        self.plan = plan
        self.dev = plan.device
        self.sources = np.asarray(sources, dtype=np.float64).reshape(-1, 2)
        self.nshot = self.sources.shape[0]
        self.rec_set = plan.add_receivers(receivers)
        self.nrec = plan.rec_sizes[self.rec_set]
        self.src_nodes = [plan.source_node(x, z) for x, z in self.sources]
        self.wavelet = plan.wavelet(f0, amp)
        self.nbz, self.nbx = nbz, nbx
        self.base = None if base is None else torch.as_tensor(base, dtype=torch.float64, device=self.dev)
        self.tuning = None

    @classmethod
    def from_workload(cls, wl, tile_rows="auto", tune=None, **plan_kw):
        """Solver for a workload; tile_rows='auto' asks the autotuner for the
        fastest one-step tile height at this mesh and batch (models x shots).
        The plan picks its schedule per call from free device memory
        (Plan.schedule_auto) unless plan_kw sets schedule='whole'.  tune=True
        (default: with schedule='auto') also tunes the two-step tile shape,
        the full-history groups and lanes, and the correlation chunk for this
        batch (autotune.tune_batch), which schedule_auto then uses."""
        plan_kw.setdefault("schedule", "auto")
        batch = wl.nmodel * wl.nshot
        if tile_rows == "auto":
            from .autotune import best_tile_rows
            tile_rows = best_tile_rows(wl.mesh.nelem_x, wl.mesh.nelem_z, batch,
                                       plan_kw.get("precision", "fp32"), plan_kw.get("device"))
        plan = Plan.from_mesh(wl.mesh, tile_rows=int(tile_rows), **plan_kw)
        nbz, nbx = wl.coef.shape[1:]
        solver = cls(plan, wl.sources, wl.receivers, wl.f0, nbz, nbx, wl.base)
        if tune is None:
            tune = plan_kw["schedule"] == "auto"
        if tune:
            from .autotune import tune_batch
            solver.tuning = tune_batch(plan, batch, solver.nrec)
        return solver

    def velocity(self, coef):
        coef = torch.as_tensor(coef, dtype=torch.float64, device=self.dev)
        if coef.dim() == 2:
            coef = coef[None]
        return self.plan.bspline_map(coef, self.base)

    def _expand(self, nmodel):
        src = np.tile(np.asarray(self.src_nodes, dtype=np.int32), nmodel)
        model_of = np.repeat(np.arange(nmodel, dtype=np.int32), self.nshot)
        return src, model_of

    def forward(self, coef) -> torch.Tensor:
        """Seismograms [nmodel, nshot, nt, nrec] for B-spline node values
        coef [nmodel, nbz, nbx]."""
        _, c2 = self.velocity(coef)
        M = c2.shape[0]
        src, model_of = self._expand(M)
        seis = self.plan.forward(c2, self.wavelet, src, [self.rec_set] * len(src), model_of=model_of)
        return seis.view(M, self.nshot, self.plan.nt, -1)

    def loglik_grad(self, coef, obs, noise_std: float, want_grad: bool = True):
        """log-likelihood summed over shots and its gradient w.r.t. the
        B-spline node values.

        obs: [nshot, nt, nrec] (shared by all models) or [nmodel, nshot, nt, nrec].
        Returns (loglik [nmodel], grad [nmodel, nbz, nbx] or None, seis)."""
        c, c2 = self.velocity(coef)
        M = c2.shape[0]
        src, model_of = self._expand(M)
        obs = torch.as_tensor(obs, device=self.dev)
        if obs.dim() == 3:
            obs = obs[None].expand(M, -1, -1, -1)
        obs = obs.reshape(M * self.nshot, self.plan.nt, -1)
        out = self.plan.forward_adjoint(c2, self.wavelet, src, obs, noise_std, [self.rec_set] * len(src),
                                        model_of=model_of, want_grad_c=want_grad)
        ll = out["loglik"].view(M, self.nshot).sum(1)
        grad = None
        if want_grad:
            g_c = out["grad_c"].view(M, self.nshot, -1).sum(1)
            grad = self.plan.bspline_project(g_c, self.nbz, self.nbx)
        return ll, grad, out["seis"].view(M, self.nshot, self.plan.nt, -1)
