"""Batched log-likelihood of the reference drivers' curve parameterisation.

The CO2 driver evaluates its likelihood one ELBO sample at a time
(CO2_plume/one_stage/single_cubic/RQS_NF.py:1488-1492), and every sample one
source at a time: a ``SEMSimulation`` per source, run by a thread pool of
three (RQS_NF.py:1035-1072), with each source's float64 log-likelihood added
to a Python float and its control-point gradient, cast to float32, added to a
float32 accumulator in source order (RQS_NF.py:1030-1091).

``MultiSourceLikelihood`` runs the same arithmetic for S samples x nsrc
sources in one batched call: the S closed-curve velocity models in one
launch, every (sample, source) solve in the same step launches (shots share
their model through ``model_of``), and the projections per source over all
samples.  Each solve's bits do not depend on the batch it runs in, so the
float32 gradient sums and the float64 log-likelihood sums equal the
reference's per-source sequence exactly.
"""
from __future__ import annotations

import copy
from typing import Optional

import numpy as np
import torch

from .simulation import SEMSimulation, _np


def source_positions(cfg: dict) -> np.ndarray:
    """The sources of a configuration: ``source.positions`` (the CO2 drivers'
    multi-source schema, RQS_NF.py:887) or the single ``source.position``."""
    src = cfg["source"]
    pos = src.get("positions")
    if pos is None or len(pos) == 0:
        if "position" not in src:
            raise ValueError("source.positions must be provided and non-empty")
        pos = [src["position"]]
    return np.asarray(pos, dtype=np.float64).reshape(-1, 2)


class MultiSourceLikelihood:
    """One inversion configuration (CO2 schema: gridded background, periodic
    closed-curve anomaly, misfit window; or the clamped 6-point curve) with
    several sources sharing one receiver set."""

    def __init__(self, cfg: dict, device=None, ctrl_base: Optional[np.ndarray] = None):
        self.positions = source_positions(cfg)
        cfg0 = copy.deepcopy(cfg)
        cfg0["source"] = dict(cfg0["source"])
        cfg0["source"]["position"] = [float(q) for q in self.positions[0]]
        self.sim = SEMSimulation(cfg0, device=device)
        if self.sim.grid is not None:
            raise ValueError("MultiSourceLikelihood takes closed-curve configurations")
        base = self.sim.ctrl_pts_original if ctrl_base is None else np.asarray(ctrl_base, dtype=np.float64)
        if base is None:
            raise ValueError("the configuration has no anomaly control points")
        self.ctrl_base = base.reshape(-1, 2)
        self.nsrc = self.positions.shape[0]
        self.nrec = self.sim.num_receivers

    def evaluate(self, z, y_obs, noise_std: float, obs_dt: float, obs_nt: int, want_grad: bool = True):
        """z: control-point offsets [S, 2K] (added to ``ctrl_base``); y_obs:
        (obs rows, nsrc * nrec), the sources' receiver blocks side by side
        (RQS_NF.py:1019-1031).  noise_std is passed to each solve as is (the
        reference's sqrt(noise_std^2 + EPS)).

        Returns (loglik float64 [S]: the per-source values summed in source
        order; grad float32 [S, 2K] or None: the per-source float32 gradients
        summed in source order; per-source loglik [S, nsrc])."""
        sim = self.sim
        z = np.asarray(_np(z), dtype=np.float64)
        z = z.reshape(1, -1) if z.ndim == 1 else z
        S, K2 = z.shape
        K = self.ctrl_base.shape[0]
        if K2 != 2 * K:
            raise ValueError(f"z must have {2 * K} offsets per sample, got {K2}")
        y = np.asarray(_np(y_obs), dtype=np.float64)
        if y.ndim != 2 or y.shape[1] != self.nsrc * self.nrec:
            raise ValueError(f"Multi-source likelihood: expected {self.nsrc * self.nrec} traces "
                             f"({self.nsrc} sources * {self.nrec} receivers) but got "
                             f"{y.shape[1] if y.ndim == 2 else y.shape}.")
        # one cached simulation per source, called sample after sample, in
        # the reference: its dt shrinks as a running minimum (quirk Q1)
        dt_prev, nt_prev = sim.dt, sim.nt
        M, c2, groups, aux = sim._batch_models(z, ctrl_base=self.ctrl_base, sequential_cfl=True)
        dt_after, nt_after = sim.dt, sim.nt
        sim.dt, sim.nt = dt_prev, nt_prev
        ll_src = np.zeros((S, self.nsrc))
        grad = torch.zeros((S, 2 * K), dtype=torch.float32, device=sim.device) if want_grad else None
        try:
            for (dt, nt), idx in groups.items():
                sim.dt, sim.nt = dt, nt
                window = sim._window()
                plan = sim._plan("post", window)
                obs = torch.stack([torch.as_tensor(
                    sim._observations(plan, y[:, k * self.nrec:(k + 1) * self.nrec], obs_dt, obs_nt, window),
                    dtype=torch.float64, device=sim.device) for k in range(self.nsrc)])
                Mi = len(idx)
                nodes = [plan.source_node(x, zz) for x, zz in self.positions]
                src = np.tile(np.asarray(nodes, dtype=np.int32), Mi)
                model_of = np.repeat(np.arange(Mi, dtype=np.int32), self.nsrc)
                obs_b = obs[None].expand(Mi, -1, -1, -1).reshape(Mi * self.nsrc, nt, self.nrec)
                sel = torch.as_tensor(np.asarray(idx, dtype=np.int64), device=sim.device)
                out = plan.forward_adjoint(c2.index_select(0, sel), sim._wavelet(plan), src, obs_b,
                                           float(noise_std), [0] * (Mi * self.nsrc), model_of=model_of,
                                           want_grad_c=want_grad)
                ll_src[idx] = out["loglik"].view(Mi, self.nsrc).cpu().numpy()
                if want_grad:
                    g = out["grad_c"].view(Mi, self.nsrc, -1)
                    aux_i = aux.index_select(0, sel)
                    a = sim._curve_args()
                    acc = grad.index_select(0, sel)
                    for k in range(self.nsrc):
                        # the reference hands dL/dv = -dJ_dv to the projection (core.py:1278-1287)
                        gk = plan.curve_project(-g[:, k], aux_i, K, a["kind"], a["v_in"], a["v_out"], a["tau"],
                                                a["background"], a["blend"])
                        acc = acc + gk.to(torch.float32)
                    grad.index_copy_(0, sel, acc)
        finally:
            sim.dt, sim.nt = dt_after, nt_after
        total = np.zeros(S)
        for k in range(self.nsrc):
            total = total + ll_src[:, k]
        return total, grad, ll_src
