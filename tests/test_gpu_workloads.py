"""The benchmark workloads at their real sizes.

* cpu40 (40x40 deg-4, 2000 steps): FP32 GPU vs FP64 oracle, seismograms,
  log-likelihood and the gradient on the 64 B-spline nodes.
* single200 (200x200 deg-4, 6000 steps, checkpoints every 100): FP32 vs an
  FP64 build of the same kernels (FP64 itself pinned to the oracle at small
  sizes in test_gpu_parity.py): seismogram and gradient relative L2.
"""
import numpy as np
import pytest
import torch

import oracle as O
from tests.gpu_helpers import rel_l2

pytestmark = pytest.mark.gpu


def test_cpu40_fp32_vs_oracle():
    from mun_amd import workloads as W
    from mun_amd.batch import BatchSolver
    print("[cpu40] oracle ...", flush=True)
    wl = W.cpu40()
    m = wl.mesh
    prob = O.Problem(m.xmin, m.xmax, m.zmin, m.zmax, m.nelem_x, m.nelem_z, 4, m.dt, m.nt, tuple(wl.sources[0]),
                     wl.f0, wl.receivers, sponge_width=m.sponge_width, sponge_cmax=m.sponge_cmax,
                     sponge_sides=m.sponge_sides)
    disc = O.discretize(prob)
    ct, _, _ = O.tensor_bspline_velocity(disc.mesh, wl.true_coef)
    yobs, _ = O.forward(disc, ct, history=None)
    c, Bz, Bx = O.tensor_bspline_velocity(disc.mesh, wl.coef[0])
    ref = O.forward_adjoint(disc, c, yobs, wl.noise_std, mode="reference", checkpoint_every=100)
    gref = Bz.T @ ref["dloglik_dc"].reshape(disc.mesh.NZ, disc.mesh.NX) @ Bx
    print("[cpu40] gpu ...", flush=True)
    s = BatchSolver.from_workload(wl, precision="fp32", adjoint_mode="reference")
    ll, g, seis = s.loglik_grad(torch.as_tensor(wl.coef, device="cuda"), torch.as_tensor(yobs[None], device="cuda"),
                                wl.noise_std)
    es = rel_l2(seis[0, 0].double().cpu().numpy(), ref["seis"])
    el = abs(ll.item() - ref["loglik"]) / abs(ref["loglik"])
    eg = rel_l2(g[0].cpu().numpy(), gref)
    print(f"  seis {es:.2e} loglik {el:.2e} grad(64 nodes) {eg:.2e}")
    assert es < 1e-5 and el < 1e-5 and eg < 1e-5


def test_single200_fp32_vs_fp64_kernels():
    from mun_amd import workloads as W
    from mun_amd.batch import BatchSolver
    wl = W.single200()
    out = {}
    for prec in ("fp64", "fp32"):
        print(f"[single200] {prec}", flush=True)
        s = BatchSolver.from_workload(wl, precision=prec, adjoint_mode="reference")
        obs = s.forward(torch.as_tensor(wl.true_coef[None], device="cuda"))[0].double()
        if prec == "fp64":
            obs64 = obs
        ll, g, seis = s.loglik_grad(torch.as_tensor(wl.coef, device="cuda"), obs64, wl.noise_std)
        out[prec] = (ll.item(), g[0].cpu().numpy(), seis[0, 0].double().cpu().numpy())
        del s
        torch.cuda.empty_cache()
    es = rel_l2(out["fp32"][2], out["fp64"][2])
    eg = rel_l2(out["fp32"][1], out["fp64"][1])
    el = abs(out["fp32"][0] - out["fp64"][0]) / abs(out["fp64"][0])
    print(f"  seis {es:.2e} loglik {el:.2e} grad(256 nodes) {eg:.2e}")
    assert es < 1e-5
    assert eg < 1e-5
