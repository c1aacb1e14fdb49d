"""GPU parity of the HIP path against the FP64 CPU oracle (through the C ABI).

Tolerances: FP64 plans agree with the oracle to <= 1e-9 relative L2 (the
kernel uses exact GLL collocation, the oracle the reference's 1e-12-perturbed
Lagrange factors); FP32 plans to <= 1e-5 relative L2 for seismograms and
gradients (the solver's stated FP32 tolerance).
"""
import numpy as np
import pytest
import torch

import oracle as O
from tests.gpu_helpers import gpu_setup, make_plan, rel_l2, small_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob_small():
    prob, disc, coef, c = small_problem()
    seis_post, _ = O.forward(disc, c, record="post", history=None)
    seis_pre, _ = O.forward(disc, c, record="pre", history=None)
    return prob, disc, coef, c, seis_post, seis_pre


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
@pytest.mark.parametrize("record", ["post", "pre"])
def test_forward_seismograms(prob_small, precision, tol, record):
    prob, disc, coef, c, sp, sq = prob_small
    print(f"[forward] precision={precision} record={record}", flush=True)
    plan = make_plan(prob, precision, record=record)
    rs, src, w, c2 = gpu_setup(prob, disc, c, plan)
    seis = plan.forward(c2, w, [src], [rs])
    torch.cuda.synchronize()
    got = seis[0].double().cpu().numpy()
    ref = sp if record == "post" else sq
    err = rel_l2(got, ref)
    print(f"  rel-L2 = {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("mode", ["reference", "exact"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_forward_adjoint_gradient(prob_small, mode, precision, tol):
    prob, disc, coef, c, sp, _ = prob_small
    print(f"[fwd+adj] mode={mode} precision={precision}", flush=True)
    rng = np.random.default_rng(7)
    cobs, _, _ = O.tensor_bspline_velocity(disc.mesh, coef + 40.0 * rng.standard_normal(coef.shape))
    yobs, _ = O.forward(disc, cobs, history=None)
    noise = 1e-9
    ref = O.forward_adjoint(disc, c, yobs, noise, mode=mode)
    plan = make_plan(prob, precision, adjoint_mode=mode, checkpoint_every=64)
    rs, src, w, c2 = gpu_setup(prob, disc, c, plan)
    obs = torch.as_tensor(yobs, dtype=torch.float64, device=plan.device)[None]
    out = plan.forward_adjoint(c2, w, [src], obs, noise, [rs])
    torch.cuda.synchronize()
    e_seis = rel_l2(out["seis"][0].double().cpu().numpy(), ref["seis"])
    e_ll = abs(float(out["loglik"][0]) - ref["loglik"]) / abs(ref["loglik"])
    e_g = rel_l2(out["grad_c"][0].cpu().numpy(), ref["dloglik_dc"])
    print(f"  seis {e_seis:.3e}  loglik {e_ll:.3e}  grad {e_g:.3e}")
    assert e_seis < tol
    assert e_ll < tol
    assert e_g < tol


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_checkpoint_recompute_bit_identical(prob_small, precision):
    """Checkpoint/recompute replays the forward exactly: the gradient equals
    the one from the stored full history bit for bit."""
    prob, disc, coef, c, sp, _ = prob_small
    print(f"[checkpoint] precision={precision}", flush=True)
    outs = []
    for ck in (None, 37, 100):
        plan = make_plan(prob, precision, adjoint_mode="exact", checkpoint_every=ck)
        rs, src, w, c2 = gpu_setup(prob, disc, c, plan)
        obs = torch.as_tensor(sp * 0.9, dtype=torch.float64, device=plan.device)[None]
        o = plan.forward_adjoint(c2, w, [src], obs, 1e-9, [rs])
        outs.append((o["grad_c"][0].cpu(), o["misfit"][0].item()))
    for g, m in outs[1:]:
        assert torch.equal(g, outs[0][0])
        assert m == outs[0][1]


def test_batch_independence(prob_small):
    """A solve gives the same bits alone and inside a mixed batch."""
    prob, disc, coef, c, sp, _ = prob_small
    print("[batch independence]", flush=True)
    plan = make_plan(prob, "fp32", adjoint_mode="exact", checkpoint_every=50)
    rs, src, w, c2 = gpu_setup(prob, disc, c, plan)
    rng = np.random.default_rng(3)
    models = [c]
    for k in range(4):
        ck, _, _ = O.tensor_bspline_velocity(disc.mesh, coef + 60.0 * rng.standard_normal(coef.shape))
        models.append(ck)
    C2 = torch.as_tensor(np.stack(models) ** 2, dtype=plan.dtype, device=plan.device)
    obs = torch.as_tensor(sp * 1.1, dtype=torch.float64, device=plan.device)[None]
    alone = plan.forward_adjoint(C2[:1], w, [src], obs, 1e-9, [rs])
    srcs = [src, disc.mesh.find_closest_node(100.0, 60.0), src, src, src]
    mix = plan.forward_adjoint(C2[[1, 2, 0, 3, 4]], w, srcs, obs.expand(5, -1, -1), 1e-9, [rs] * 5)
    assert torch.equal(alone["grad_c"][0], mix["grad_c"][2])
    assert torch.equal(alone["seis"][0], mix["seis"][2])
    assert alone["misfit"][0].item() == mix["misfit"][2].item()


def test_bspline_map_and_project(prob_small):
    prob, disc, coef, c, sp, _ = prob_small
    print("[bspline]", flush=True)
    plan = make_plan(prob, "fp64")
    cg, c2g = plan.bspline_map(torch.as_tensor(coef)[None])
    torch.cuda.synchronize()
    assert rel_l2(cg[0].cpu().numpy(), c) < 1e-14
    assert rel_l2(c2g[0].cpu().numpy(), c * c) < 1e-14
    rng = np.random.default_rng(1)
    g = rng.standard_normal(disc.mesh.N)
    _, Bz, Bx = O.tensor_bspline_velocity(disc.mesh, coef)
    ref = Bz.T @ g.reshape(disc.mesh.NZ, disc.mesh.NX) @ Bx
    got = plan.bspline_project(torch.as_tensor(g)[None], *coef.shape)[0].cpu().numpy()
    assert rel_l2(got, ref) < 1e-13


@pytest.mark.parametrize("tile_rows", [15, 7, 3])
def test_every_tile_height_matches_oracle(prob_small, tile_rows):
    """Each step-kernel instance (tile heights 15, 7, 3) reproduces the
    oracle's seismograms and exact-mode gradient in FP64."""
    prob, disc, coef, c, sp, _ = prob_small
    print(f"[tile_rows={tile_rows}]", flush=True)
    ref = O.forward_adjoint(disc, c, sp * 0.8, 1e-9, mode="exact")
    plan = make_plan(prob, "fp64", adjoint_mode="exact", checkpoint_every=45, tile_rows=tile_rows)
    rs, src, w, c2 = gpu_setup(prob, disc, c, plan)
    obs = torch.as_tensor(sp * 0.8, dtype=torch.float64, device=plan.device)[None]
    out = plan.forward_adjoint(c2, w, [src], obs, 1e-9, [rs])
    torch.cuda.synchronize()
    assert rel_l2(out["seis"][0].cpu().numpy(), ref["seis"]) < 1e-9
    assert rel_l2(out["grad_c"][0].cpu().numpy(), ref["dloglik_dc"]) < 1e-9


def test_autotuner_picks_and_caches(tmp_path, monkeypatch):
    from mun_amd import autotune
    monkeypatch.setenv("MUN_AMD_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    autotune._MEM.clear()
    rep = {}
    best = autotune.best_tile_rows(20, 20, 8, "fp32", steps=10, report=rep)
    assert best in autotune.TILE_ROWS
    assert set(rep["step_us"]) == {"15", "7", "3"}
    print("  20x20 batch 8:", rep)
    autotune._MEM.clear()
    assert autotune.best_tile_rows(20, 20, 8, "fp32", steps=10) == best  # from the file cache


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_residual_formed_in_double_near_a_fit(prob_small, precision):
    """Near a fit (|r| / |d| = 1e-3) the residual d_syn - y_obs is formed in
    double from the double observations, as core.py:1224-1227 forms it: the
    misfit equals 0.5 sum ((double) d - y)^2 to rounding.  Casting y to FP32
    first would lose the residual's low digits (about 6e-8 |d| per sample,
    1e-4 of |r|)."""
    prob = prob_small[0]
    print(f"[residual near fit] {precision}", flush=True)
    plan = make_plan(prob, precision)
    rng = np.random.default_rng(5)
    B, nt, pitch, nrec = 2, 300, 16, 13
    d = torch.as_tensor(rng.standard_normal((B, nt, pitch)), dtype=plan.dtype, device=plan.device)
    y = d.double() * (1.0 + 1e-3 * torch.as_tensor(rng.standard_normal((B, nt, pitch)), device=plan.device))
    got = plan.misfit(d, y, nrec).cpu().numpy()
    r = (d.double() - y)[:, :, :nrec].cpu().numpy()
    want = 0.5 * np.sum(r * r, axis=(1, 2))
    r32 = (d - y.to(plan.dtype)).double()[:, :, :nrec].cpu().numpy()
    lossy = 0.5 * np.sum(r32 * r32, axis=(1, 2))
    err = np.abs(got / want - 1).max()
    print(f"  misfit rel err {err:.2e} (FP32-cast observations: {np.abs(lossy / want - 1).max():.2e})")
    assert err < 1e-12
    if precision == "fp32":
        assert np.abs(lossy / want - 1).max() > 1e-7
