"""Shared set-up for GPU parity tests: one small problem built twice, for the
oracle (CPU, FP64) and for the HIP path (through the C ABI)."""
import numpy as np
import torch

import oracle as O


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def small_problem(nex=20, nez=17, nt=400, dt=1e-3, sides=("xmin", "xmax", "zmax"), seed=0, nrec=12,
                  src=(260.0, 40.0), width=120.0, f0=12.0):
    """Mesh spanning several 15x15-element tiles (partial tiles included)."""
    L_x, L_z = 25.0 * nex, 25.0 * nez
    rec = np.stack([np.linspace(30.0, L_x - 30.0, nrec), np.full(nrec, 5.0)], axis=1)
    prob = O.Problem(0.0, L_x, 0.0, L_z, nex, nez, 4, dt, nt, src, f0, rec,
                     sponge_width=width, sponge_cmax=3000.0, sponge_sides=tuple(sides))
    disc = O.discretize(prob)
    rng = np.random.default_rng(seed)
    coef = 1800.0 + 900.0 * rng.random((6, 7))
    c, _, _ = O.tensor_bspline_velocity(disc.mesh, coef)
    return prob, disc, coef, c


def make_plan(prob, precision="fp64", adjoint_mode="reference", record="post", checkpoint_every=None,
              tile_rows=0):
    from mun_amd.solver import Plan
    return Plan(prob.nelem_x, prob.nelem_z, prob.xmin, prob.xmax, prob.zmin, prob.zmax, prob.dt, prob.nt, 4,
                prob.sponge_width, prob.sponge_cmax, prob.sponge_R, prob.sponge_sides, precision=precision,
                adjoint_mode=adjoint_mode, record=record, checkpoint_every=checkpoint_every, tile_rows=tile_rows)


def gpu_setup(prob, disc, c, plan):
    rs = plan.add_receivers(prob.rec_xz)
    src = plan.source_node(*prob.src_xz)
    assert src == disc.src_node
    w = torch.as_tensor(disc.wavelet, dtype=plan.dtype, device=plan.device)
    c2 = torch.as_tensor(c * c, dtype=plan.dtype, device=plan.device)[None]
    return rs, src, w, c2
