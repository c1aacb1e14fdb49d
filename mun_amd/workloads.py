"""Seeded synthetic workloads for the SEM solver.

Five workloads, each a concrete mesh / velocity / source / receiver layout:

  cpu40    40x40 deg-4 elements on 2 km x 2 km, 8x8 cubic B-spline velocity
           nodes (linear depth gradient 1500-3000 m/s + Gaussian anomaly,
           amp 300 m/s, sigma 300 m), 10 Hz Ricker at 50 m depth, 50 surface
           receivers, dt 0.5 ms, 2000 steps, sponge on left/right/bottom.
  single200  200x200 deg-4 elements on 4 km x 4 km, 5 horizontal layers
           1600-3500 m/s + 16x16 B-spline perturbation N(0, 100 m/s),
           15 Hz Ricker, 200 receivers, dt 0.25 ms, 6000 steps,
           checkpoints every 100 steps.
  batch200 256 velocity models on the single200 mesh, 16x16 perturbations
           N(0, 150 m/s) around the layered model.
  elbo120  64 samples x 8 shots, 120x120 deg-4 elements on 3 km x 3 km,
           10 Hz, 100 receivers per shot, 4000 steps at dt 0.4 ms, 12x12 nodes.
  served   Poisson stream of requests on {20,40,80,160}^2 element meshes with
           1000-5000 steps, 8x8..16x16 B-spline draws in 1500-3500 m/s and
           20-200 receivers.

Coordinates: x to the right, z = depth (0 at the free surface, increasing
downward).  The free surface (natural Neumann boundary) is the z = zmin side;
the sponge damps xmin, xmax and zmax.  Density is accepted and unused: the
reference's equation u_tt = div(c^2 grad u) has no density
(Bspline_Coupling/SEM/sem_waveform/operators.py:97-107).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import bspline

SIDES_NO_SURFACE = ("xmin", "xmax", "zmax")


@dataclass
class MeshSpec:
    nelem_x: int
    nelem_z: int
    xmin: float
    xmax: float
    zmin: float
    zmax: float
    order: int = 4
    dt: float = 5e-4
    nt: int = 2000
    sponge_width: float = 200.0
    sponge_cmax: float = 3000.0
    sponge_sides: tuple = SIDES_NO_SURFACE
    checkpoint_every: int = 100

    @property
    def NX(self):
        return self.nelem_x * self.order + 1

    @property
    def NZ(self):
        return self.nelem_z * self.order + 1

    @property
    def N(self):
        return self.NX * self.NZ


@dataclass
class Workload:
    name: str
    mesh: MeshSpec
    f0: float
    sources: np.ndarray            # (nshot, 2)
    receivers: np.ndarray          # (nrec, 2)
    coef: np.ndarray               # (nmodel, nbz, nbx) B-spline node values
    base: Optional[np.ndarray]     # (NZ*NX,) background added to the spline, or None
    true_coef: np.ndarray          # (nbz, nbx) model that generates the observations
    noise_std: float = 1e-9
    meta: dict = field(default_factory=dict)

    @property
    def nmodel(self):
        return self.coef.shape[0]

    @property
    def nshot(self):
        return self.sources.shape[0]


def surface_receivers(xmin, xmax, n, margin, z=0.0):
    return np.stack([np.linspace(xmin + margin, xmax - margin, n), np.full(n, z)], axis=1)


def gradient_plus_gaussian_coef(nb, L, vtop=1500.0, vbot=3000.0, amp=300.0, sigma=300.0,
                                center=None):
    """8x8-style control values: linear depth gradient at the Greville
    abscissae plus a Gaussian anomaly (the spline reproduces the linear part
    exactly)."""
    g = bspline.greville(0.0, L, nb)
    if center is None:
        center = (0.5 * L, 0.5 * L)
    Z, X = np.meshgrid(g, g, indexing="ij")
    lin = vtop + (vbot - vtop) * Z / L
    gau = amp * np.exp(-((X - center[0]) ** 2 + (Z - center[1]) ** 2) / (2 * sigma ** 2))
    return lin + gau


def layered_profile(z, L, nlayer=5, vtop=1600.0, vbot=3500.0):
    """Horizontal layers, equal thickness, velocities linear from vtop to vbot,
    layer index by searchsorted(side='right') as
    CO2_plume/layered_model/inversion/sem_waveform/velocity.py:217."""
    interfaces = np.linspace(0.0, L, nlayer + 1)[1:-1]
    vel = np.linspace(vtop, vbot, nlayer)
    return vel[np.searchsorted(interfaces, z, side="right")]


def mesh_axes(m: MeshSpec):
    from .geometry import gll_axis
    return gll_axis(m.xmin, m.xmax, m.nelem_x, m.order), gll_axis(m.zmin, m.zmax, m.nelem_z, m.order)


def cpu40(seed: int = 0) -> Workload:
    L = 2000.0
    m = MeshSpec(40, 40, 0.0, L, 0.0, L, 4, dt=5e-4, nt=2000, sponge_width=200.0, sponge_cmax=3000.0)
    coef = gradient_plus_gaussian_coef(8, L)
    rng = np.random.default_rng(seed)
    true = coef + rng.normal(0.0, 50.0, coef.shape)
    return Workload("cpu40", m, 10.0, np.array([[1000.0, 50.0]]),
                    surface_receivers(0.0, L, 50, 200.0), coef[None], None, true,
                    meta={"density": 1000.0})


def _layered_workload(name, nel, L, nb, sigma_v, nmodel, seed, f0, nrec, dt, nt, width,
                      nshot=1, vtop=1600.0, vbot=3500.0):
    m = MeshSpec(nel, nel, 0.0, L, 0.0, L, 4, dt=dt, nt=nt, sponge_width=width,
                 sponge_cmax=vbot, checkpoint_every=100)
    x1d, z1d = mesh_axes(m)
    base = np.repeat(layered_profile(z1d, L, 5, vtop, vbot)[:, None], m.NX, axis=1).reshape(-1)
    rng = np.random.default_rng(seed)
    coef = rng.normal(0.0, sigma_v, (nmodel, nb, nb))
    true = rng.normal(0.0, sigma_v, (nb, nb))
    if nshot == 1:
        src = np.array([[0.5 * L, 50.0]])
    else:
        src = np.stack([np.linspace(width, L - width, nshot), np.full(nshot, 50.0)], axis=1)
    return Workload(name, m, f0, src, surface_receivers(0.0, L, nrec, width), coef, base, true,
                    meta={"density": 1000.0})


def single200(seed: int = 1) -> Workload:
    return _layered_workload("single200", 200, 4000.0, 16, 100.0, 1, seed, 15.0, 200, 2.5e-4, 6000, 300.0)


def batch200(nmodel: int = 256, seed: int = 2, nt: int = 6000) -> Workload:
    return _layered_workload("batch200", 200, 4000.0, 16, 150.0, nmodel, seed, 15.0, 200, 2.5e-4, nt, 300.0)


def elbo120(nsample: int = 64, nshot: int = 8, seed: int = 3, nt: int = 4000) -> Workload:
    return _layered_workload("elbo120", 120, 3000.0, 12, 100.0, nsample, seed, 10.0, 100, 4e-4, nt, 250.0,
                             nshot=nshot)


@dataclass
class Request:
    rid: int
    arrival: float
    workload: Workload


SERVED_SIZES = (20, 40, 80, 160)


def served_request(rid: int, rng: np.random.Generator, arrival: float, nt_scale: float = 1.0) -> Request:
    """One request of the served stream: mesh size uniform over SERVED_SIZES,
    element size 25 m, steps 1000..5000 growing with the mesh, B-spline grid
    8..16 nodes per axis with values 1500..3500 m/s, 20..200 receivers."""
    k = int(rng.integers(len(SERVED_SIZES)))
    nel = SERVED_SIZES[k]
    L = 25.0 * nel
    nt = int(round((1000 + 4000 * k / (len(SERVED_SIZES) - 1)) * nt_scale))
    dt = 4e-4
    width = min(0.15 * L, 300.0)
    m = MeshSpec(nel, nel, 0.0, L, 0.0, L, 4, dt=dt, nt=nt, sponge_width=width, sponge_cmax=3500.0,
                 checkpoint_every=100)
    nb = int(rng.integers(8, 17))
    coef = rng.uniform(1500.0, 3500.0, (1, nb, nb))
    true = np.clip(coef[0] + rng.normal(0.0, 100.0, (nb, nb)), 1500.0, 3500.0)
    nrec = int(rng.integers(20, 201))
    wl = Workload(f"served{nel}", m, 10.0, np.array([[0.5 * L, 0.1 * L]]),
                  surface_receivers(0.0, L, nrec, width), coef, None, true)
    return Request(rid, arrival, wl)


def served_stream(nreq: int, rate_hz: float, seed: int = 5, nt_scale: float = 1.0) -> List[Request]:
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, nreq))
    return [served_request(i, rng, float(t[i]), nt_scale) for i in range(nreq)]


WORKLOADS = {"cpu40": cpu40, "single200": single200, "batch200": batch200, "elbo120": elbo120}


def served_solve_requests(nreq: int, rate_hz: float, seed: int = 5, nt_scale: float = 1.0, device="cuda",
                          noise_std: float = 1e-9):
    """The served stream as solver requests, with observations generated by a
    forward solve of each request's perturbed ('true') model."""
    import torch
    from .serving import SolveRequest
    from .solver import Plan, source_wavelet
    reqs = []
    plans = {}
    for r in served_stream(nreq, rate_hz, seed, nt_scale):
        wl, m = r.workload, r.workload.mesh
        key = (m.nelem_x, m.nt)
        if key not in plans:
            plans[key] = Plan.from_mesh(m, precision="fp32", device=device)
        p = plans[key]
        rs = p.add_receivers(wl.receivers)
        _, c2 = p.bspline_map(torch.as_tensor(wl.true_coef)[None])
        w = torch.as_tensor(source_wavelet(m.nt, m.dt, wl.f0), dtype=torch.float32, device=device)
        obs = p.forward(c2, w, [p.source_node(*wl.sources[0])], [rs])[0].cpu().numpy()
        reqs.append(SolveRequest(r.rid, m.nelem_x, m.nelem_z, m.xmax, m.zmax, m.dt, m.nt, wl.coef[0],
                                 tuple(wl.sources[0]), wl.f0, wl.receivers, obs, noise_std, m.sponge_width,
                                 m.sponge_cmax, arrival=r.arrival))
    return reqs
