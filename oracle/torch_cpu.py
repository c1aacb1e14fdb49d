"""FP64 torch-CPU form of the oracle's forward + adjoint, for the timed CPU baseline.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): ``bench.py``'s
``cpu_baseline`` leg times it on the host cores; ``tests/`` check it against
``sem_oracle.forward_adjoint``.

It runs the recurrences of ``sem_oracle.forward`` / ``sem_oracle.adjoint``
the way the reference's torch path does (Bspline_Coupling/SEM/sem_waveform/
core.py:1197-1275): the assembled stiffness as a torch sparse CSR tensor, one
sparse mat-vec per step, the sponge, source and rotation as torch vector
operations, and the gradient correlation as dense per-element products of
the 25 x 25 element operators (core.py:49-85), on as many threads as
``torch.set_num_threads`` allows.  The forward keeps its FP64 history in
memory, so ``nsteps`` bounds the steps run (a timing sample).
"""
from __future__ import annotations

import time
from typing import Optional

import numpy as np
import torch

from . import sem_oracle as O


def _csr(K) -> torch.Tensor:
    return torch.sparse_csr_tensor(torch.as_tensor(K.indptr, dtype=torch.int64),
                                   torch.as_tensor(K.indices, dtype=torch.int64),
                                   torch.as_tensor(K.data, dtype=torch.float64), size=K.shape)


def forward_adjoint_torch(disc: O.Discretization, c: np.ndarray, y_obs: np.ndarray, noise_std: float,
                          mode: str = "reference", nsteps: Optional[int] = None, threads: Optional[int] = None):
    """Forward (post-update sampling, full history) + adjoint + gradient of
    the first ``nsteps`` steps (default: all), as ``sem_oracle.forward_adjoint``
    computes them without a misfit window.  Returns dict(loglik, dloglik_dc,
    seis, times={'assembly', 'forward', 'adjoint'} in seconds, steps)."""
    if threads is not None:
        torch.set_num_threads(int(threads))
    p = disc.prob
    nt = p.nt if nsteps is None else min(int(nsteps), p.nt)
    N = disc.mesh.N
    times = {}
    t0 = time.perf_counter()
    K = _csr(O.stiffness_csr(c, disc.mesh, disc.ops))
    times["assembly"] = time.perf_counter() - t0
    dt2 = p.dt * p.dt
    D = torch.as_tensor(disc.decay)
    Minv_f = torch.as_tensor(1.0 / disc.M)
    Minv_a = torch.as_tensor(1.0 / (disc.M + 1e-30))
    R = _csr(disc.R)
    RT = _csr(disc.R.T.tocsr())
    s = torch.as_tensor(disc.wavelet)
    src = disc.src_node
    Msrc = float(disc.M[src])
    hist = torch.empty((nt, N), dtype=torch.float64)
    seis = torch.empty((nt, R.shape[0]), dtype=torch.float64)
    # forward (sem_oracle.forward; S/core.py:1197-1215)
    t0 = time.perf_counter()
    u_prev = torch.zeros(N, dtype=torch.float64)
    u_curr = torch.zeros(N, dtype=torch.float64)
    for it in range(nt):
        rhs = -(K @ u_curr)
        rhs[src] += float(s[it]) * Msrc
        u_next = 2.0 * u_curr - u_prev + dt2 * (Minv_f * rhs)
        u_prev, u_curr = D * u_curr, D * u_next
        seis[it] = R @ u_curr
        hist[it] = u_curr
    times["forward"] = time.perf_counter() - t0
    # misfit (S/core.py:1224-1230)
    sigma2 = float(noise_std) ** 2 + 1e-24
    r = seis - torch.as_tensor(np.asarray(y_obs, dtype=np.float64)[:nt])
    loglik = -0.5 * float(torch.sum(r * r)) / sigma2
    rs = r / sigma2
    # adjoint and correlation (sem_oracle.adjoint; S/core.py:1247-1275, 49-85)
    Gx, Gz, S, w2D = (torch.as_tensor(a) for a in disc.ops)
    conn = torch.as_tensor(disc.mesh.conn, dtype=torch.int64)
    flat = conn.reshape(-1)
    t0 = time.perf_counter()
    lam_curr = torch.zeros(N, dtype=torch.float64)
    lam_next = torch.zeros(N, dtype=torch.float64)
    g = torch.zeros(N, dtype=torch.float64)
    exact = mode == "exact"
    for it in range(nt - 1, -1, -1):
        rhs = (RT @ rs[it]) - K @ lam_curr
        if exact:
            lam_prev = D * (2.0 * lam_curr - D * lam_next + dt2 * (Minv_a * rhs))
        else:
            lam_prev = 2.0 * lam_curr - lam_next + dt2 * (Minv_a * rhs)
        if it < nt - 1:
            ul = hist[it][conn]
            ll = lam_curr[conn]
            gq = w2D[None, :] * ((ul @ Gx.T) * (ll @ Gx.T) + (ul @ Gz.T) * (ll @ Gz.T))
            g.index_add_(0, flat, (gq @ S).reshape(-1))
        lam_next, lam_curr = lam_curr, lam_prev
    times["adjoint"] = time.perf_counter() - t0
    return {"loglik": loglik, "dloglik_dc": (2.0 * torch.as_tensor(c) * g).numpy(), "seis": seis.numpy(),
            "times": times, "steps": nt}
