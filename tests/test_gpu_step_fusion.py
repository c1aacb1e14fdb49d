"""Two time steps per step-kernel launch (sem_plan_set_step_fusion(plan, 2))
give the same bits as one step per launch: seismograms, misfit and gradient,
on a mesh of several ragged two-step tiles with the source and receivers on
tile boundaries, an odd step count, both precisions, both adjoint modes,
both recording conventions, and the full-history, lane and checkpoint
schedules."""
import numpy as np
import pytest
import torch

import oracle as O
from tests.gpu_helpers import make_plan, rel_l2

pytestmark = pytest.mark.gpu


def _problem(nt=263):
    # 37 x 50 elements: 3 x 3 two-step tiles of 16 x 23 owned elements (the
    # last column and row partial); tile corners at x = 400 m, z = 575 m
    nex, nez = 37, 50
    L_x, L_z = 25.0 * nex, 25.0 * nez
    xs = np.linspace(30.0, L_x - 30.0, 9)
    zs = np.linspace(30.0, L_z - 30.0, 7)
    rec = np.concatenate([np.stack([xs, np.full(xs.size, 575.0)], axis=1),     # along a tile row boundary
                          np.stack([np.full(zs.size, 400.0), zs], axis=1),     # along a tile column boundary
                          np.array([[399.0, 574.0], [800.0, 1150.0], [L_x, L_z], [0.0, 0.0]])])
    prob = O.Problem(0.0, L_x, 0.0, L_z, nex, nez, 4, 1e-3, nt, (400.0, 575.0), 12.0, rec,
                     sponge_width=120.0, sponge_cmax=3000.0, sponge_sides=("xmin", "xmax", "zmax"))
    disc = O.discretize(prob)
    rng = np.random.default_rng(3)
    coef = 1800.0 + 900.0 * rng.random((6, 7))
    c, _, _ = O.tensor_bspline_velocity(disc.mesh, coef)
    return prob, disc, coef, c


@pytest.fixture(scope="module")
def fusion_case():
    prob, disc, coef, c = _problem()
    seis, _ = O.forward(disc, c, record="post", history=None)
    return prob, disc, coef, c, seis


def _inputs(prob, disc, coef, c, plan, nsolve=5):
    rs = plan.add_receivers(prob.rec_xz)
    src = plan.source_node(*prob.src_xz)
    assert src == disc.src_node
    w = torch.as_tensor(disc.wavelet, dtype=plan.dtype, device=plan.device)
    rng = np.random.default_rng(9)
    models = [c] + [O.tensor_bspline_velocity(disc.mesh, coef + 80.0 * rng.standard_normal(coef.shape))[0]
                    for _ in range(2)]
    C2 = torch.as_tensor(np.stack(models) ** 2, dtype=plan.dtype, device=plan.device)
    src2 = disc.mesh.find_closest_node(130.0, 990.0)
    srcs = [src2 if k % 2 else src for k in range(nsolve)]
    model_of = [k % 3 for k in range(nsolve)]
    return C2, w, srcs, model_of, rs


@pytest.mark.parametrize("precision,mode,record", [("fp32", "reference", "post"), ("fp32", "exact", "pre"),
                                                   ("fp64", "exact", "post"), ("fp64", "reference", "pre")])
def test_two_steps_per_launch_give_the_same_bits(fusion_case, precision, mode, record):
    prob, disc, coef, c, seis = fusion_case
    print(f"[fusion] {precision} {mode} {record}", flush=True)
    plan = make_plan(prob, precision, adjoint_mode=mode, record=record, checkpoint_every=40)
    C2, w, srcs, model_of, rs = _inputs(prob, disc, coef, c, plan)
    B = len(srcs)
    obs = torch.as_tensor(seis * 0.9, dtype=plan.dtype, device=plan.device)[None].expand(B, -1, -1)
    assert plan.get_step_fusion() == 2  # the default
    for sch in [(0, 0, 1), (0, 2, 3), (40, 0, 1), (33, 2, 2)]:
        plan.set_schedule(*sch)
        outs = []
        for steps in (1, 2):
            plan.set_step_fusion(steps)
            o = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
            torch.cuda.synchronize()
            outs.append({k: o[k].cpu() for k in ("seis", "misfit", "grad_c")})
        for k in ("seis", "misfit", "grad_c"):
            assert torch.equal(outs[0][k], outs[1][k]), (sch, k)
        assert torch.isfinite(outs[1]["grad_c"]).all() and outs[1]["grad_c"].abs().max() > 0


def test_two_steps_with_nontemporal_streams_give_the_same_bits(fusion_case):
    """The non-temporal instances (NTS, chosen when a launch's working set
    exceeds the last-level cache; forced here) under one- and two-step
    launches, full-history and checkpointed schedules."""
    prob, disc, coef, c, seis = fusion_case
    print("[fusion] nontemporal", flush=True)
    plan = make_plan(prob, "fp32", adjoint_mode="exact", record="post", checkpoint_every=40)
    C2, w, srcs, model_of, rs = _inputs(prob, disc, coef, c, plan)
    B = len(srcs)
    obs = torch.as_tensor(seis * 0.9, dtype=plan.dtype, device=plan.device)[None].expand(B, -1, -1)
    plan.set_streaming("nontemporal")
    for sch in [(0, 0, 1), (40, 2, 2)]:
        plan.set_schedule(*sch)
        outs = []
        for steps in (1, 2):
            plan.set_step_fusion(steps)
            o = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
            torch.cuda.synchronize()
            outs.append({k: o[k].cpu() for k in ("seis", "misfit", "grad_c")})
        for k in ("seis", "misfit", "grad_c"):
            assert torch.equal(outs[0][k], outs[1][k]), (sch, k)
    plan.set_streaming("auto")


@pytest.mark.parametrize("precision,record", [("fp32", "post"), ("fp64", "pre")])
def test_forward_only_two_steps_give_the_same_bits(fusion_case, precision, record):
    """The forward-only pass (ping-pong states) under one- and two-step
    launches, the odd last step included."""
    prob, disc, coef, c, seis = fusion_case
    print(f"[fusion] forward {precision} {record}", flush=True)
    plan = make_plan(prob, precision, record=record)
    C2, w, srcs, model_of, rs = _inputs(prob, disc, coef, c, plan)
    outs = []
    for steps in (1, 2):
        plan.set_step_fusion(steps)
        outs.append(plan.forward(C2, w, srcs, [rs] * len(srcs), model_of=model_of).cpu())
    assert torch.equal(outs[0], outs[1])
    if precision == "fp64" and record == "pre":
        ref, _ = O.forward(disc, c, record="pre", history=None)
        assert rel_l2(outs[1][0, :, :ref.shape[1]].numpy(), ref) < 1e-9


@pytest.mark.parametrize("every", [1, 2, 3])
def test_snapshot_runs_with_two_steps_give_the_same_bits(fusion_case, every):
    """Snapshot runs (a state every `every` steps, which a step pair may end
    on but not hold in its middle) under one- and two-step launches."""
    prob, disc, coef, c, seis = fusion_case
    print(f"[fusion] snapshots every {every}", flush=True)
    plan = make_plan(prob, "fp32", record="pre")
    C2, w, srcs, model_of, rs = _inputs(prob, disc, coef, c, plan, nsolve=3)
    box = (10, 120, 30, 170)
    outs = []
    for steps in (1, 2):
        plan.set_step_fusion(steps)
        sm, sn = plan.forward_snapshots(C2, w, srcs, every, box, [rs] * len(srcs), model_of=model_of)
        outs.append((sm.cpu(), sn.cpu()))
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    assert outs[1][1].abs().max() > 0


def test_two_steps_per_launch_match_the_oracle(fusion_case):
    """Solve 0 (the fixture's model and source) under two-step launches
    against the FP64 oracle's exact-mode gradient and post-update seismogram."""
    prob, disc, coef, c, seis = fusion_case
    print("[fusion] oracle", flush=True)
    plan = make_plan(prob, "fp64", adjoint_mode="exact", record="post", checkpoint_every=40)
    C2, w, srcs, model_of, rs = _inputs(prob, disc, coef, c, plan, nsolve=3)
    B = len(srcs)
    obs = torch.as_tensor(seis * 0.9, dtype=plan.dtype, device=plan.device)[None].expand(B, -1, -1)
    plan.set_schedule(0, 0, 1)
    plan.set_step_fusion(2)
    o = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
    torch.cuda.synchronize()
    ref = O.forward_adjoint(disc, c, seis * 0.9, 1e-9, mode="exact")
    assert rel_l2(o["seis"][0].cpu().numpy(), seis) < 1e-9
    assert rel_l2(o["grad_c"][0].cpu().numpy(), ref["dloglik_dc"]) < 1e-9


def test_step_fusion_rejects_other_counts():
    from mun_amd._lib import SemError
    prob, disc, coef, c = _problem(nt=20)
    plan = make_plan(prob, "fp32")
    for bad in (0, 3, -1):
        with pytest.raises(SemError):
            plan.set_step_fusion(bad)
    assert plan.get_step_fusion() == 2


def test_two_steps_at_the_benchmark_mesh_give_the_same_bits():
    """The 200 x 200-element benchmark mesh (117 two-step tiles, 200 surface
    receivers across tile boundaries, 5-layer model with B-spline
    perturbations), 4 solves in 2 lanes of full-history groups and the
    checkpointed whole batch, 301 steps: one- and two-step launches give the
    same log-likelihood, gradient and seismograms."""
    from mun_amd import workloads as W
    from mun_amd.batch import BatchSolver
    print("[fusion] batch200 mesh", flush=True)
    wl = W.batch200(nmodel=4, nt=301)
    s = BatchSolver.from_workload(wl, tile_rows=15, precision="fp32", adjoint_mode="exact")
    coef = torch.as_tensor(wl.coef, device="cuda")
    obs = s.forward(torch.as_tensor(wl.true_coef[None], device="cuda"))[0]
    for sch in [(0, 2, 2), (50, 0, 1)]:
        s.plan.set_schedule(*sch)
        outs = []
        for steps in (1, 2):
            s.plan.set_step_fusion(steps)
            ll, g, seis = s.loglik_grad(coef, obs, wl.noise_std)
            torch.cuda.synchronize()
            outs.append((ll.cpu(), g.cpu(), seis.cpu()))
        for k in range(3):
            assert torch.equal(outs[0][k], outs[1][k]), (sch, k)
        assert torch.isfinite(outs[1][1]).all() and outs[1][1].abs().max() > 0
    s.plan.set_step_fusion(2)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_every_two_step_tile_shape_gives_the_same_bits(fusion_case, precision):
    """The two-step kernel's tile shapes (sem_plan_set_tile2; 0..4, shape 3
    FP32 only) give the bits of one step per launch, under the full-history
    lanes and the checkpointed schedule."""
    prob, disc, coef, c, seis = fusion_case
    print(f"[tile2] {precision}", flush=True)
    plan = make_plan(prob, precision, adjoint_mode="exact", record="post", checkpoint_every=40)
    C2, w, srcs, model_of, rs = _inputs(prob, disc, coef, c, plan)
    B = len(srcs)
    obs = torch.as_tensor(seis * 0.9, dtype=torch.float64, device=plan.device)[None].expand(B, -1, -1)
    shapes = (0, 1, 2, 3, 4) if precision == "fp32" else (0, 1, 2, 4)
    if precision == "fp64":
        with pytest.raises(Exception, match="FP64"):
            plan.set_tile2(3)
    for sch in [(0, 2, 3), (40, 0, 1)]:
        plan.set_schedule(*sch)
        plan.set_step_fusion(1)
        o = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
        ref = {k: o[k].cpu() for k in ("seis", "misfit", "grad_c")}
        plan.set_step_fusion(2)
        for sh in shapes:
            plan.set_tile2(sh)
            assert plan.get_tile2()[0] == sh
            o = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
            torch.cuda.synchronize()
            for k in ref:
                assert torch.equal(o[k].cpu(), ref[k]), (sch, sh, k)
        plan.set_tile2(0)


def test_every_correlation_chunk_gives_the_same_bits(fusion_case):
    prob, disc, coef, c, seis = fusion_case
    print("[corr chunk]", flush=True)
    plan = make_plan(prob, "fp32", adjoint_mode="reference", record="post", checkpoint_every=40)
    C2, w, srcs, model_of, rs = _inputs(prob, disc, coef, c, plan)
    B = len(srcs)
    obs = torch.as_tensor(seis * 0.9, dtype=torch.float64, device=plan.device)[None].expand(B, -1, -1)
    outs = []
    for ck in (64, 1, 2, 7, 16, 128):
        plan.set_corr_chunk(ck)
        assert plan.get_corr_chunk() == ck
        for sch in [(0, 2, 3), (40, 0, 1)]:
            plan.set_schedule(*sch)
            o = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
            torch.cuda.synchronize()
            outs.append({k: o[k].cpu() for k in ("seis", "misfit", "grad_c")})
    for o in outs[1:]:
        for k in o:
            assert torch.equal(o[k], outs[0][k]), k
