"""Schedules of sem_forward_adjoint: a batch run whole, with checkpoints or
the full history, or split into groups on several lanes (streams of the
plan's own) gives the same bits for every solve; the automatic planner picks
a schedule that fits the memory budget it is given."""
import numpy as np
import pytest
import torch

import oracle as O
from tests.gpu_helpers import gpu_setup, make_plan, rel_l2, small_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob_sched():
    prob, disc, coef, c = small_problem(nt=260)
    seis, _ = O.forward(disc, c, record="post", history=None)
    return prob, disc, coef, c, seis


def _batch(prob, disc, coef, c, plan, nsolve=7, shared_models=True):
    """nsolve solves over 4 models (model_of) or one model per solve, two
    source positions."""
    rs, src, w, c2 = gpu_setup(prob, disc, c, plan)
    rng = np.random.default_rng(5)
    nm = 4 if shared_models else nsolve
    models = [c] + [O.tensor_bspline_velocity(disc.mesh, coef + 70.0 * rng.standard_normal(coef.shape))[0]
                    for _ in range(nm - 1)]
    C2 = torch.as_tensor(np.stack(models) ** 2, dtype=plan.dtype, device=plan.device)
    src2 = disc.mesh.find_closest_node(110.0, 55.0)
    srcs = [src2 if k % 3 == 2 else src for k in range(nsolve)]  # solve 0: the fixture's model and source
    model_of = [k % nm for k in range(nsolve)] if shared_models else None
    return C2, w, srcs, model_of, rs


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("shared_models", [True, False])
def test_every_schedule_gives_the_same_bits(prob_sched, precision, shared_models):
    prob, disc, coef, c, seis = prob_sched
    print(f"[schedules] {precision} shared_models={shared_models}", flush=True)
    plan = make_plan(prob, precision, adjoint_mode="exact", checkpoint_every=50)
    C2, w, srcs, model_of, rs = _batch(prob, disc, coef, c, plan, shared_models=shared_models)
    B = len(srcs)
    obs = torch.as_tensor(seis * 0.9, dtype=torch.float64, device=plan.device)[None].expand(B, -1, -1)
    # (checkpoint_every, group, lanes): whole batch with checkpoints, whole with
    # the full history, groups on 1-3 lanes with either policy
    schedules = [(50, 0, 1), (0, 0, 1), (0, 2, 3), (0, 3, 1), (37, 3, 2), (0, 1, 3), (50, 2, 2)]
    outs = []
    for sch in schedules:
        plan.set_schedule(*sch)
        assert plan.get_schedule() == dict(zip(("checkpoint_every", "group", "lanes"), sch))
        o = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
        torch.cuda.synchronize()
        outs.append({k: o[k].cpu() for k in ("seis", "misfit", "grad_c")})
    for sch, o in zip(schedules[1:], outs[1:]):
        for k in ("seis", "misfit", "grad_c"):
            assert torch.equal(o[k], outs[0][k]), (sch, k)
    # and solve 0 (the fixture's model) still matches the oracle
    ref = O.forward_adjoint(disc, c, seis * 0.9, 1e-9, mode="exact")
    tol = 1e-9 if precision == "fp64" else 1e-5
    e_g = rel_l2(outs[0]["grad_c"][0].numpy(), ref["dloglik_dc"])
    print(f"  solve 0 vs oracle: grad {e_g:.2e}")
    assert e_g < tol


def test_schedule_auto_fits_the_budget(prob_sched):
    prob, disc, coef, c, seis = prob_sched
    print("[schedule auto]", flush=True)
    plan = make_plan(prob, "fp32", adjoint_mode="reference", checkpoint_every=40)
    C2, w, srcs, model_of, rs = _batch(prob, disc, coef, c, plan)
    B, pitch = len(srcs), prob.rec_xz.shape[0]
    obs = torch.as_tensor(seis, dtype=torch.float64, device=plan.device)[None].expand(B, -1, -1)
    ws_bytes = plan.workspace_bytes
    # a large budget: the whole batch with its full history, one group per lane
    assert plan.schedule_auto(B, pitch, 1 << 40) == {"checkpoint_every": 0, "group": -(-B // 3), "lanes": 3}
    full = ws_bytes(B, pitch, True)
    ref = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
    # a budget below the full history of one solve: checkpoints (the plan's 40 steps)
    plan.set_schedule(0, 1, 1)
    one_full = ws_bytes(1, pitch, True)
    sch = plan.schedule_auto(B, pitch, one_full - 1)
    assert sch["checkpoint_every"] == 40
    assert ws_bytes(B, pitch, True) <= one_full - 1
    got = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
    torch.cuda.synchronize()
    for k in ("seis", "misfit", "grad_c"):
        assert torch.equal(got[k], ref[k]), k
    # a budget between: groups with the full history, never above the budget
    budget = full // 2
    sch = plan.schedule_auto(B, pitch, budget)
    assert ws_bytes(B, pitch, True) <= budget
    assert 0 < sch["group"] < B or sch["checkpoint_every"] > 0
    # below one checkpointed solve: a clean error
    from mun_amd._lib import SemError
    with pytest.raises(SemError):
        plan.schedule_auto(B, pitch, 1024)


def test_lanes_at_the_headline_mesh():
    """The benchmark's mesh (200 x 200 elements, 801^2 nodes): lanes of
    full-history groups (as many tiles and solves in flight as the bench
    runs) give the same bits as the whole batch with checkpoints."""
    from mun_amd import workloads as W
    from mun_amd.batch import BatchSolver
    print("[schedules 801^2]", flush=True)
    wl = W.batch200(nmodel=8, nt=60)
    s = BatchSolver.from_workload(wl, precision="fp32", adjoint_mode="exact", tile_rows=15, schedule="whole")
    obs = s.forward(torch.as_tensor(wl.true_coef[None], device="cuda"))[0]
    coef = torch.as_tensor(wl.coef, device="cuda")
    outs = []
    for sch in ((20, 0, 1), (0, 3, 3), (0, 8, 1)):
        s.plan.set_schedule(*sch)
        ll, g, seis = s.loglik_grad(coef, obs, wl.noise_std)
        torch.cuda.synchronize()
        outs.append((ll.cpu(), g.cpu(), seis.cpu()))
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert torch.equal(a, b)
    assert bool(torch.isfinite(outs[0][1]).all()) and float(outs[0][1].abs().max()) > 0


def test_streaming_hints_give_the_same_bits(prob_sched):
    """Plain and non-temporal state streams change only cache behaviour."""
    prob, disc, coef, c, seis = prob_sched
    print("[streaming hints]", flush=True)
    plan = make_plan(prob, "fp32", adjoint_mode="exact", checkpoint_every=40)
    C2, w, srcs, model_of, rs = _batch(prob, disc, coef, c, plan)
    B = len(srcs)
    obs = torch.as_tensor(seis, dtype=torch.float64, device=plan.device)[None].expand(B, -1, -1)
    outs = []
    for mode in ("plain", "nontemporal", "auto"):
        plan.set_streaming(mode)
        o = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
        torch.cuda.synchronize()
        outs.append(o)
    for o in outs[1:]:
        for k in ("seis", "misfit", "grad_c"):
            assert torch.equal(o[k], outs[0][k]), k


def test_misfit_only_call_skips_the_adjoint(prob_sched):
    """Without gradient outputs the call returns the same misfit and
    seismograms as with them."""
    prob, disc, coef, c, seis = prob_sched
    print("[misfit only]", flush=True)
    plan = make_plan(prob, "fp32", adjoint_mode="reference", checkpoint_every=40)
    C2, w, srcs, model_of, rs = _batch(prob, disc, coef, c, plan)
    B = len(srcs)
    obs = torch.as_tensor(seis, dtype=torch.float64, device=plan.device)[None].expand(B, -1, -1)
    for sch in ((40, 0, 1), (0, 3, 3)):
        plan.set_schedule(*sch)
        full = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
        only = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of, want_grad_c=False)
        torch.cuda.synchronize()
        assert only["grad_c"] is None
        assert torch.equal(only["misfit"], full["misfit"])
        assert torch.equal(only["seis"], full["seis"])


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_schedules_with_a_misfit_window(prob_sched, mode):
    """A misfit window (the adjoint runs over the window only in the
    reference mode) gives the same bits under every schedule."""
    from mun_amd.solver import Plan
    prob, disc, coef, c, seis = prob_sched
    print(f"[schedules window] {mode}", flush=True)
    plan = Plan(prob.nelem_x, prob.nelem_z, prob.xmin, prob.xmax, prob.zmin, prob.zmax, prob.dt, prob.nt, 4,
                prob.sponge_width, prob.sponge_cmax, prob.sponge_R, prob.sponge_sides, precision="fp32",
                adjoint_mode=mode, checkpoint_every=30, misfit_window=(40, 200, 3))
    C2, w, srcs, model_of, rs = _batch(prob, disc, coef, c, plan)
    B = len(srcs)
    obs = torch.as_tensor(seis * 0.95, dtype=torch.float64, device=plan.device)[None].expand(B, -1, -1)
    outs = []
    for sch in ((30, 0, 1), (0, 2, 3), (45, 3, 2)):
        plan.set_schedule(*sch)
        o = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
        torch.cuda.synchronize()
        outs.append(o)
    assert float(outs[0]["grad_c"].abs().max()) > 0
    for o in outs[1:]:
        for k in ("seis", "misfit", "grad_c"):
            assert torch.equal(o[k], outs[0][k]), k


def test_schedules_with_per_solve_wavelets(prob_sched):
    """Per-solve source wavelets (rows of a [B, nt] tensor) follow their
    solves into the groups of every schedule."""
    prob, disc, coef, c, seis = prob_sched
    print("[schedules per-solve wavelets]", flush=True)
    plan = make_plan(prob, "fp32", adjoint_mode="reference", checkpoint_every=50)
    C2, w, srcs, model_of, rs = _batch(prob, disc, coef, c, plan)
    B = len(srcs)
    W2 = torch.stack([w * (1.0 + 0.1 * k) for k in range(B)]).contiguous()
    obs = torch.as_tensor(seis, dtype=torch.float64, device=plan.device)[None].expand(B, -1, -1)
    outs = []
    for sch in ((50, 0, 1), (0, 2, 3), (0, 3, 1)):
        plan.set_schedule(*sch)
        o = plan.forward_adjoint(C2, W2, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
        torch.cuda.synchronize()
        outs.append(o)
    s0 = outs[0]["seis"]
    for o in outs[1:]:
        for k in ("seis", "misfit", "grad_c"):
            assert torch.equal(o[k], outs[0][k]), k
    # solve k's wavelet is (1 + 0.1 k) w, and the seismograms are linear in it
    one = plan.forward_adjoint(C2, w, srcs, obs, 1e-9, [rs] * B, model_of=model_of)
    torch.cuda.synchronize()
    ratio = (s0[3].double().norm() / one["seis"][3].double().norm()).item()
    assert abs(ratio - 1.3) < 1e-5
