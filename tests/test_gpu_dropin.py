"""Drop-in surface on the GPU: SEMSimulation / autograd / batched solves
compared with the FP64 oracle (reference-style configs)."""
import numpy as np
import pytest
import torch

import oracle as O
from tests.gpu_helpers import rel_l2

pytestmark = pytest.mark.gpu

# the reference's control points (Bspline_Coupling/SEM/sem_forward.py:42-43) moved off the mesh
# lattice: a node lying exactly on the curve has distance ~1e-14 and its outward normal
# r/(|r| + 1e-15) is rounding noise in any implementation, the reference's included
CTRL = np.array([[-100, -100], [100, -100], [150, -50], [200, 0], [0, 200], [-200, 100]], dtype=np.float64) \
    + np.array([3.7, -2.3])


def reference_config(precision="fp64", mode="reference", nelem=24):
    """A config in the reference's schema (Bspline_Coupling/SEM/sem_forward.py:15-58),
    p = 4, dt below the reference CFL limit so no shrink happens."""
    return {
        "domain": {"xmin": -600, "xmax": 600, "zmin": -600, "zmax": 600, "nelem_x": nelem, "nelem_z": nelem},
        "time": {"total_time": 0.25, "dt": 2.0e-4},
        "source": {"position": (0.0, 0.0), "frequency": 15.0, "amplitude": 1.0},
        "receivers": {"num_receivers": 16, "radius": 400.0},
        "method": {"polynomial_order": 4, "pml_thickness": 150.0, "precision": precision, "adjoint_mode": mode,
                   "checkpoint_every": 50},
        "velocity": {"inside_velocity": 2000.0, "outside_velocity": 3000.0, "control_points": CTRL,
                     "perturbations": np.zeros((6, 2)), "tau": 10.0, "spline_samples": 800},
        "output": {"save_wavefield": False, "save_seismograms": True, "visualize": False},
    }


def oracle_for(cfg, sim):
    d = cfg["domain"]
    prob = O.Problem(d["xmin"], d["xmax"], d["zmin"], d["zmax"], d["nelem_x"], d["nelem_z"], 4, sim.dt, sim.nt,
                     cfg["source"]["position"], cfg["source"]["frequency"],
                     np.asarray(sim.receiver_coords_sem), sponge_width=cfg["method"]["pml_thickness"],
                     sponge_cmax=cfg["velocity"]["outside_velocity"])
    return O.discretize(prob)


def test_curve_velocity_matches_oracle():
    from mun_amd.simulation import SEMSimulation
    print("[curve velocity]", flush=True)
    cfg = reference_config()
    sim = SEMSimulation(cfg)
    disc = oracle_for(cfg, sim)
    v, sd, us, no = O.closed_curve_velocity(disc.mesh.coords, CTRL, 2000.0, 3000.0, 10.0, 800, 7)
    c, c2, aux = sim._curve_velocity(CTRL)
    got = c[0].cpu().numpy()
    np.testing.assert_allclose(got, v, rtol=0, atol=1e-9)
    A = aux[0].cpu().numpy()
    np.testing.assert_allclose(A[0], sd, atol=1e-8)
    np.testing.assert_allclose(A[1], us, atol=1e-9)
    np.testing.assert_allclose(A[2:].T, no, atol=1e-9)
    # the stored model uses the reference node numbering
    perm = O.reference_node_order(24, 24, 4)
    np.testing.assert_allclose(sim.global_coords, disc.mesh.coords[perm], atol=1e-9)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_run_matches_oracle(precision, tol):
    from mun_amd.simulation import SEMSimulation
    print(f"[run] {precision}", flush=True)
    cfg = reference_config(precision)
    sim = SEMSimulation(cfg)
    out = sim.run()
    disc = oracle_for(cfg, sim)
    v, _, _, _ = O.closed_curve_velocity(disc.mesh.coords, CTRL, 2000.0, 3000.0, 10.0, 800, 7)
    ref, _ = O.forward(disc, v, record="pre", history=None)
    assert out["receiver_data"].shape == (sim.nt, 16) and out["nt"] == sim.nt
    assert rel_l2(out["receiver_data"], ref) < tol


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_run_forward_and_adjoint_matches_oracle(mode):
    from mun_amd.simulation import SEMSimulation
    print(f"[rfa] {mode}", flush=True)
    cfg = reference_config("fp64", mode)
    sim = SEMSimulation(cfg)
    disc = oracle_for(cfg, sim)
    rng = np.random.default_rng(0)
    true_ctrl = CTRL + rng.normal(0, 20.0, CTRL.shape)
    vt, _, _, _ = O.closed_curve_velocity(disc.mesh.coords, true_ctrl, 2000.0, 3000.0, 10.0, 800, 7)
    yobs, _ = O.forward(disc, vt, record="post", history=None)
    out = sim.run_forward_and_adjoint({"bspline_ctrl": CTRL.reshape(-1), "y_obs": yobs, "obs_dt": sim.dt,
                                       "obs_nt": sim.nt, "noise_std": 1e-7})
    v, sd, us, no = O.closed_curve_velocity(disc.mesh.coords, CTRL, 2000.0, 3000.0, 10.0, 800, 7)
    ref = O.forward_adjoint(disc, v, yobs, 1e-7, mode=mode)
    g12 = O.closed_curve_projection(-ref["dloglik_dc"], sd, us, no, 2000.0, 3000.0, 10.0)
    assert abs(out["loglik"] - ref["loglik"]) < 1e-8 * abs(ref["loglik"])
    assert rel_l2(out["clean_data"], ref["seis"]) < 1e-9
    assert rel_l2(out["grad_wrt_ctrl"], g12) < 1e-6
    assert out["grad_wrt_ctrl"].shape == (12,)


def test_autograd_function_reference_signature():
    """SEMLikelihoodAdjointFn hands the solver sqrt(noise_std^2 + 1e-12)
    (RQS_Coupling.py:319-333).  At the driver's obs_noise_std = 5e-9
    (RQS_Coupling.py:82) the log-likelihood and gradient are therefore those
    of noise_std = 5e-9 scaled by (2.5e-17 + 1e-24) / (2.5e-17 + 1e-12 +
    1e-24) (the core adds 1e-24, core.py:1086), and equal the oracle's at
    sigma^2 = 2.5e-17 + 1e-12 + 1e-24."""
    from mun_amd.autograd import SEMLikelihoodAdjointFn, get_or_make_sim
    print("[autograd]", flush=True)
    cfg = reference_config("fp64")
    sim = get_or_make_sim(cfg)
    yobs = sim.run()["receiver_data"] * 1.05
    ns = 5e-9
    z = torch.zeros(12, dtype=torch.float64, requires_grad=True)
    ll = SEMLikelihoodAdjointFn.apply(z, torch.as_tensor(yobs), ns, cfg, sim.dt, sim.nt)
    ll.backward()
    raw = sim.run_forward_and_adjoint({"bspline_ctrl": CTRL, "y_obs": yobs, "obs_dt": sim.dt, "obs_nt": sim.nt,
                                       "noise_std": ns})
    scale = (ns ** 2 + 1e-24) / (ns ** 2 + 1e-12 + 1e-24)
    print(f"  scale {scale:.6e}: loglik {ll.item():.9e} vs raw x scale {raw['loglik'] * scale:.9e}")
    assert abs(ll.item() - raw["loglik"] * scale) <= 1e-12 * abs(raw["loglik"] * scale)
    np.testing.assert_allclose(z.grad.numpy(), raw["grad_wrt_ctrl"] * scale, rtol=1e-11)
    # against the oracle at the reference's variance
    disc = oracle_for(cfg, sim)
    v, sd, us, no = O.closed_curve_velocity(disc.mesh.coords, CTRL, 2000.0, 3000.0, 10.0, 800, 7)
    ref = O.forward_adjoint(disc, v, yobs, np.sqrt(ns ** 2 + 1e-12), mode="reference")
    g12 = O.closed_curve_projection(-ref["dloglik_dc"], sd, us, no, 2000.0, 3000.0, 10.0)
    el = abs(ll.item() - ref["loglik"]) / abs(ref["loglik"])
    eg = rel_l2(z.grad.numpy(), g12)
    print(f"  vs oracle: loglik {el:.2e} grad {eg:.2e}")
    assert el < 1e-9 and eg < 1e-6


def _grid_solver(precision, mode, nshot=1, nt=300, seed=0):
    from mun_amd.batch import BatchSolver
    from mun_amd.solver import Plan
    L = 800.0
    plan = Plan(32, 32, 0.0, L, 0.0, L, 5e-4, nt, 4, 150.0, 3500.0, 1e-8, ("xmin", "xmax", "zmax"),
                precision=precision, adjoint_mode=mode, checkpoint_every=64)
    src = np.stack([np.linspace(200.0, 600.0, nshot), np.full(nshot, 30.0)], axis=1)
    rec = np.stack([np.linspace(50.0, 750.0, 20), np.zeros(20)], axis=1)
    return BatchSolver(plan, src, rec, 15.0, 8, 8, None), src, rec


def test_batched_gradient_matches_central_fd_fp64():
    """d loglik / d node (h = 1 m/s central differences on 8 seeded nodes) vs
    the exact-mode adjoint, both in FP64 on the GPU."""
    from mun_amd.autograd import batched_loglik
    print("[fd]", flush=True)
    s, _, _ = _grid_solver("fp64", "exact")
    rng = np.random.default_rng(5)
    coef = torch.as_tensor(2200.0 + 300.0 * rng.random((1, 8, 8)), device="cuda")
    obs = s.forward(coef + torch.as_tensor(rng.normal(0, 40, (1, 8, 8)), device="cuda"))[0]
    c = coef.clone().requires_grad_(True)
    ll = batched_loglik(c, s, obs, 1e-8)
    ll.sum().backward()
    g = c.grad[0].cpu().numpy()
    idx = rng.choice(64, 8, replace=False)

    def fd(k, h):
        e = torch.zeros_like(coef)
        e.view(-1)[k] = h
        lp = s.loglik_grad(coef + e, obs, 1e-8, want_grad=False)[0].item()
        lm = s.loglik_grad(coef - e, obs, 1e-8, want_grad=False)[0].item()
        return (lp - lm) / (2 * h)

    f1 = np.array([fd(k, 1.0) for k in idx])           # h = 1 m/s as the workload spec states
    f2 = np.array([fd(k, 0.5) for k in idx])
    rich = (4.0 * f2 - f1) / 3.0                        # O(h^4) Richardson extrapolation
    ga = g.reshape(-1)[idx]
    scale = np.abs(g).max()
    e1 = np.abs(ga - f1) / scale
    er = np.abs(ga - rich) / scale
    print("  h=1 FD err / max|g| =", e1.max(), " Richardson err =", er.max())
    assert e1.max() < 1e-3          # FD truncation at h = 1 m/s
    assert er.max() < 1e-6          # the adjoint itself


def test_batched_gradient_matches_oracle_and_shots_share_models():
    """Two models x three shots in one launch sequence == oracle per shot."""
    print("[multishot]", flush=True)
    s, src, rec = _grid_solver("fp32", "reference", nshot=3, nt=240)
    rng = np.random.default_rng(6)
    coef = 2200.0 + 300.0 * rng.random((2, 8, 8))
    prob = [O.Problem(0.0, 800.0, 0.0, 800.0, 32, 32, 4, 5e-4, 240, tuple(src[k]), 15.0, rec, sponge_width=150.0,
                      sponge_cmax=3500.0, sponge_sides=("xmin", "xmax", "zmax")) for k in range(3)]
    discs = [O.discretize(p) for p in prob]
    ctrue, _, _ = O.tensor_bspline_velocity(discs[0].mesh, coef[0] + 30.0)
    obs = np.stack([O.forward(d, ctrue, history=None)[0] for d in discs])
    ll, grad, seis = s.loglik_grad(torch.as_tensor(coef, device="cuda"), torch.as_tensor(obs, device="cuda"), 1e-8)
    for m in range(2):
        c, Bz, Bx = O.tensor_bspline_velocity(discs[0].mesh, coef[m])
        tot_ll, tot_g = 0.0, 0.0
        for k in range(3):
            r = O.forward_adjoint(discs[k], c, obs[k], 1e-8, mode="reference")
            tot_ll += r["loglik"]
            tot_g = tot_g + r["dloglik_dc"]
            assert rel_l2(seis[m, k, :, :20].double().cpu().numpy(), r["seis"]) < 1e-5
        gref = Bz.T @ tot_g.reshape(discs[0].mesh.NZ, discs[0].mesh.NX) @ Bx
        el = abs(ll[m].item() - tot_ll) / abs(tot_ll)
        eg = rel_l2(grad[m].cpu().numpy(), gref)
        print(f"  model {m}: loglik {el:.2e} grad {eg:.2e}")
        assert el < 1e-5 and eg < 1e-5
