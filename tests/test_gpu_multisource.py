"""The CO2 driver's multi-source likelihood over a batch of samples
(mun_amd.likelihood.MultiSourceLikelihood, autograd.MultiSourceLikelihoodFn)
against its per-source sequential form (CO2_plume/one_stage/single_cubic/
RQS_NF.py:987-1098: one cached SEMSimulation per source, float64 log-likelihoods
summed in source order, float32 gradients summed in source order) and
against the FP64 oracle."""
import copy

import numpy as np
import pytest
import torch

import oracle as O
from tests.gpu_helpers import rel_l2
from tests.test_gpu_extensions import co2_config

pytestmark = pytest.mark.gpu

POSITIONS = [(-300.0, -400.0), (0.0, -400.0), (250.0, -420.0)]


def multi_config(tmp_path, precision):
    cfg, ctrl, grid = co2_config(tmp_path, precision=precision, mode="reference", window=True, K=8)
    cfg["output"]["save_wavefield"] = False
    cfg["source"] = {"positions": [list(p) for p in POSITIONS], "frequency": 12.0, "amplitude": 1.0}
    return cfg, ctrl, grid


def observations(ms, ctrl_true):
    """Windowed, decimated data of a perturbed curve, the sources' receiver
    blocks side by side (the driver's y_obs layout)."""
    sim = ms.sim
    win = sim._window()
    plan = sim._plan("post")
    a = sim._curve_args()
    _, c2t, _ = plan.curve_velocity(ctrl_true[None], a["kind"], a["v_in"], a["v_out"], a["tau"], 800, 7,
                                    a["background"], a["blend"])
    blocks = []
    for x, z in POSITIONS:
        seis = plan.forward(c2t, sim._wavelet(plan), [plan.source_node(x, z)], [0])[0].double().cpu().numpy()
        blocks.append(seis[win[0]:win[1]:win[2], :ms.nrec])
    return np.concatenate(blocks, axis=1), sim.dt * win[2]


def sequential(cfg, base, z, y, noise_std, obs_dt, obs_nt):
    """The reference's per-sample, per-source loop (RQS_NF.py:1030-1091)."""
    from mun_amd.autograd import get_or_make_sim
    nrec = len(cfg["receivers"]["positions"])
    lls, grads = [], []
    for zi in z:
        total, acc = 0.0, np.zeros(zi.size, dtype=np.float32)
        for k, pos in enumerate(cfg["source"]["positions"]):
            cfg_k = copy.deepcopy(cfg)
            cfg_k["source"]["position"] = [float(pos[0]), float(pos[1])]
            sim_k = get_or_make_sim(cfg_k)
            out = sim_k.run_forward_and_adjoint({"bspline_ctrl": base + zi.reshape(-1, 2),
                                                 "y_obs": y[:, k * nrec:(k + 1) * nrec], "obs_dt": obs_dt,
                                                 "obs_nt": obs_nt, "noise_std": noise_std})
            total += float(out["loglik"])
            acc += np.asarray(out["grad_wrt_ctrl"], dtype=np.float32).reshape(-1)
        lls.append(total)
        grads.append(acc)
    return np.asarray(lls), np.stack(grads)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batched_multisource_equals_per_source_sequence(tmp_path, precision):
    from mun_amd.autograd import MultiSourceLikelihoodFn, effective_noise_std, EPS_CO2_DRIVER
    from mun_amd.likelihood import MultiSourceLikelihood
    print(f"[multisource] {precision}", flush=True)
    cfg, ctrl, _ = multi_config(tmp_path, precision)
    ms = MultiSourceLikelihood(cfg)
    rng = np.random.default_rng(4)
    y, obs_dt = observations(ms, ctrl + rng.normal(0.0, 6.0, ctrl.shape))
    z = rng.normal(0.0, 5.0, (3, 2 * ctrl.shape[0]))
    ns = effective_noise_std(1e-9, EPS_CO2_DRIVER)
    ll, grad, per = ms.evaluate(z, y, ns, obs_dt, y.shape[0])
    ll_seq, g_seq = sequential(cfg, ctrl, z, y, ns, obs_dt, y.shape[0])
    assert per.shape == (3, len(POSITIONS))
    np.testing.assert_array_equal(ll, ll_seq)
    np.testing.assert_array_equal(grad.cpu().numpy(), g_seq)
    # the Function: float32 log-likelihoods, backward = upstream * float32 gradient
    zt = torch.as_tensor(z, dtype=torch.float64).requires_grad_(True)
    out = MultiSourceLikelihoodFn.apply(zt, torch.as_tensor(y), 1e-9, cfg, obs_dt, y.shape[0])
    assert out.dtype == torch.float32 and out.shape == (3,)
    np.testing.assert_array_equal(out.detach().numpy(), ll_seq.astype(np.float32))
    (out * torch.tensor([1.0, 2.0, -0.5])).sum().backward()
    want = g_seq * np.array([1.0, 2.0, -0.5], dtype=np.float32)[:, None]
    np.testing.assert_array_equal(zt.grad.numpy(), want.astype(np.float64))
    one = MultiSourceLikelihoodFn.apply(torch.as_tensor(z[1]), torch.as_tensor(y), 1e-9, cfg, obs_dt, y.shape[0])
    assert one.dim() == 0 and one.item() == np.float32(ll_seq[1])


def test_batched_multisource_matches_oracle(tmp_path):
    from mun_amd.likelihood import MultiSourceLikelihood
    print("[multisource oracle]", flush=True)
    cfg, ctrl, (xg, zg, vp) = multi_config(tmp_path, "fp64")
    ms = MultiSourceLikelihood(cfg)
    sim = ms.sim
    rng = np.random.default_rng(12)
    y, obs_dt = observations(ms, ctrl + rng.normal(0.0, 6.0, ctrl.shape))
    z = rng.normal(0.0, 5.0, (2, 2 * ctrl.shape[0]))
    ll, grad, _ = ms.evaluate(z, y, 1e-9, obs_dt, y.shape[0])
    d = cfg["domain"]
    rec = np.asarray(cfg["receivers"]["positions"])
    win = sim._window()
    x2, z2, v2 = O.increasing_grid(xg, zg, vp)
    tot_ll, tot_g = 0.0, 0.0
    nrec = rec.shape[0]
    for k, pos in enumerate(POSITIONS):
        prob = O.Problem(d["xmin"], d["xmax"], d["zmin"], d["zmax"], 20, 20, 4, sim.dt, sim.nt, pos, 12.0, rec,
                         sponge_width=150.0, sponge_cmax=sim.c_max)
        disc = O.discretize(prob)
        bg = O.grid_background(disc.mesh.coords, x2, z2, v2, "linear")
        v, sd, us, no = O.periodic_curve_velocity(disc.mesh.coords, ctrl + z[0].reshape(-1, 2), 1900.0, bg, 12.0,
                                                  800, 7)
        yk = np.zeros((sim.nt, nrec))
        yk[win[0]:win[1]:win[2]] = y[:, k * nrec:(k + 1) * nrec]
        ref = O.forward_adjoint(disc, v, yk, 1e-9, mode="reference", window=win)
        tot_ll += ref["loglik"]
        tot_g = tot_g + O.periodic_curve_projection(-ref["dloglik_dc"], sd, us, no, 1900.0, bg, 12.0,
                                                    ctrl.shape[0])
    el = abs(ll[0] - tot_ll) / abs(tot_ll)
    eg = rel_l2(grad[0].cpu().numpy(), tot_g)
    print(f"  sample 0 vs oracle: loglik {el:.2e} grad (float32 sums) {eg:.2e}")
    assert el < 1e-9 and eg < 1e-6
