"""Continuous batching: each result is the same alone and in the mix.

A stream of uneven requests (three mesh sizes, different step counts,
sources, receiver counts and B-spline grids) is served with overlapping
admission; every result is compared with the same request served alone,
with the plain one-at-a-time batch path, and one request with the FP64
oracle."""
import numpy as np
import pytest
import torch

import oracle as O
from tests.gpu_helpers import rel_l2

pytestmark = pytest.mark.gpu


def make_requests(n=10, seed=0):
    from mun_amd.serving import SolveRequest
    from mun_amd.solver import Plan, source_wavelet
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        nel = int(rng.choice([8, 12, 20]))
        Lx = 25.0 * nel
        nt = int(rng.integers(150, 420))
        nb = int(rng.integers(5, 9))
        coef = rng.uniform(1800.0, 3000.0, (nb, nb))
        true = coef + rng.normal(0.0, 60.0, (nb, nb))
        nrec = int(rng.integers(5, 30))
        rec = np.stack([np.linspace(20.0, Lx - 20.0, nrec), np.full(nrec, 2.0)], axis=1)
        src = (float(rng.uniform(0.3, 0.7) * Lx), 30.0)
        # observations from the true model, forward path
        p = Plan(nel, nel, 0.0, Lx, 0.0, Lx, 4e-4, nt, 4, 60.0, 3500.0, 1e-8, ("xmin", "xmax", "zmax"),
                 precision="fp32", checkpoint_every=100)
        rs = p.add_receivers(rec)
        _, c2 = p.bspline_map(torch.as_tensor(true)[None])
        w = torch.as_tensor(source_wavelet(nt, 4e-4, 12.0), dtype=torch.float32, device="cuda")
        obs = p.forward(c2, w, [p.source_node(*src)], [rs])[0].cpu().numpy()
        reqs.append(SolveRequest(i, nel, nel, Lx, Lx, 4e-4, nt, coef, src, 12.0, rec, obs, 1e-8, 60.0, 3500.0,
                                 arrival=0.002 * i))
    return reqs


def test_served_results_equal_alone_in_mix_and_plain():
    from mun_amd.serving import SEMServer, latency_stats, run_plain
    print("[serve] building requests", flush=True)
    reqs = make_requests()
    print("[serve] mix", flush=True)
    srv = SEMServer(hbm_budget_bytes=8e9, segment=60, capacity={8: 3, 12: 2, 20: 2}, use_graph=True,
                    rec_pitch=32, nt_max=420)
    mix = srv.run(reqs)
    assert len(mix) == len(reqs)
    assert max(l.rounds for l in srv.lanes.values()) > 1
    # the stream really overlapped: some request was admitted while another was resident
    adm = sorted((r.admitted, r.finished) for r in mix.values())
    assert any(adm[i + 1][0] < adm[i][1] for i in range(len(adm) - 1))
    print("[serve] alone", flush=True)
    for r in reqs[:4]:
        solo = SEMServer(hbm_budget_bytes=8e9, segment=60, use_graph=False, rec_pitch=32, nt_max=420).run([r])
        a, b = solo[r.rid], mix[r.rid]
        assert a.loglik == b.loglik
        np.testing.assert_array_equal(a.grad, b.grad)
        np.testing.assert_array_equal(a.seis, b.seis)
    print("[serve] plain", flush=True)
    plain = run_plain(reqs, segment=60)
    for rid, p in plain.items():
        m = mix[rid]
        assert abs(p.loglik - m.loglik) <= 1e-6 * abs(p.loglik)
        assert rel_l2(m.grad, p.grad) < 1e-6
    st = latency_stats(mix)
    print("  served", st, "\n  plain", latency_stats(plain))


def test_served_request_matches_oracle():
    from mun_amd.serving import SEMServer
    r = make_requests(1, seed=3)[0]
    print(f"[serve-oracle] nel={r.nelem_x} nt={r.nt}", flush=True)
    res = SEMServer(hbm_budget_bytes=4e9, segment=64, use_graph=True, rec_pitch=32, nt_max=500).run([r])[r.rid]
    prob = O.Problem(0.0, r.xmax, 0.0, r.zmax, r.nelem_x, r.nelem_z, 4, r.dt, r.nt, r.src_xz, r.f0, r.rec_xz,
                     sponge_width=r.sponge_width, sponge_cmax=r.sponge_cmax, sponge_sides=("xmin", "xmax", "zmax"))
    d = O.discretize(prob)
    c, Bz, Bx = O.tensor_bspline_velocity(d.mesh, r.coef)
    ref = O.forward_adjoint(d, c, r.obs, r.noise_std, mode="reference")
    gref = Bz.T @ ref["dloglik_dc"].reshape(d.mesh.NZ, d.mesh.NX) @ Bx
    es, eg = rel_l2(res.seis, ref["seis"]), rel_l2(res.grad, gref)
    el = abs(res.loglik - ref["loglik"]) / abs(ref["loglik"])
    print(f"  seis {es:.2e} grad {eg:.2e} loglik {el:.2e}")
    assert es < 1e-5 and eg < 1e-5 and el < 1e-5


def test_history_and_checkpoint_slots_give_the_same_bits():
    """A request served with its full forward history (no recompute) gives the
    same bits as with checkpoints, alone and when the budget admits some
    requests one way and some the other."""
    from mun_amd.serving import SEMServer, request_footprint
    reqs = make_requests(6, seed=7)
    print("[serve-history] checkpoints", flush=True)
    kw = dict(segment=60, capacity={8: 3, 12: 3, 20: 3}, use_graph=True, rec_pitch=32, nt_max=420)
    ck = SEMServer(hbm_budget_bytes=8e9, history="checkpoints", **kw).run(reqs)
    print("[serve-history] full history", flush=True)
    srv = SEMServer(hbm_budget_bytes=8e9, history="auto", **kw)
    full = srv.run(reqs)
    assert all(a is None for l in srv.lanes.values() for a in l.slots)
    # a budget that holds one full-history request of the largest mesh plus
    # checkpointed ones: admission mixes both kinds
    big = max(reqs, key=lambda r: (r.nelem_x, r.nt))
    lb = srv.lanes[big.lane_key]
    fp_full = request_footprint(lb.plan.NX, lb.plan.NZ, lb.E, big.nt, lb.rec_pitch, 60, grad_elems=lb.GE,
                                full_history=True, two_step=srv.two_step)
    print(f"[serve-history] mixed budget {1.5 * fp_full / 1e6:.1f} MB", flush=True)
    mixed = SEMServer(hbm_budget_bytes=1.5 * fp_full, history="auto", **kw).run(reqs)
    for r in reqs:
        for other in (full, mixed):
            a, b = ck[r.rid], other[r.rid]
            assert a.loglik == b.loglik
            np.testing.assert_array_equal(a.grad, b.grad)
            np.testing.assert_array_equal(a.seis, b.seis)


@pytest.mark.parametrize("history", ["auto", "checkpoints"])
def test_two_step_rounds_give_the_same_bits(history):
    """Serving rounds that advance every slot two steps per launch
    (sem_slots_step2; odd segment tails take one step, the increment buffers
    swap accordingly, checkpointed forward slots keep u by step pair) give the
    same bits as one step per launch, with captured rounds, for history and
    checkpointed requests."""
    from mun_amd.serving import SEMServer
    reqs = make_requests(6, seed=11)
    kw = dict(segment=60, capacity={8: 3, 12: 3, 20: 3}, use_graph=True, rec_pitch=32, nt_max=420,
              history=history)
    print(f"[serve-two-step] {history}: one step per launch", flush=True)
    one = SEMServer(hbm_budget_bytes=8e9, two_step=False, **kw).run(reqs)
    assert any(r.nt % 2 == 1 for r in reqs)  # odd segment tails occur
    for tile2 in (None, 1, 2):  # the lanes' tuned tile shape, then two fixed ones
        print(f"[serve-two-step] {history}: two steps per launch, tile2={tile2}", flush=True)
        srv = SEMServer(hbm_budget_bytes=8e9, two_step=True, tile2=tile2, **kw)
        assert srv.two_step
        two = srv.run(reqs)
        if tile2 is not None:
            assert all(lane.plan.get_tile2()[0] == tile2 for lane in srv.lanes.values())
        for r in reqs:
            a, b = one[r.rid], two[r.rid]
            assert a.loglik == b.loglik
            np.testing.assert_array_equal(a.grad, b.grad)
            np.testing.assert_array_equal(a.seis, b.seis)


def test_servers_closed_one_after_another_give_the_same_bits():
    """Servers created and closed one after another in one process (each
    lane on a new stream, destroyed by close()) give the same bits; closing
    twice is harmless, and a server deleted without close() keeps its
    streams without disturbing the next."""
    import gc

    from mun_amd.serving import SEMServer
    reqs = make_requests(3, seed=17)
    kw = dict(segment=60, capacity={8: 2, 12: 2, 20: 2}, use_graph=True, rec_pitch=32, nt_max=420)
    first = None
    for k, loop in enumerate(("device", "host", "device", "host", "device")):
        print(f"[serve-close] server {k} ({loop})", flush=True)
        srv = SEMServer(hbm_budget_bytes=4e9, loop=loop, **kw)
        res = srv.run(reqs)
        if k < 4:
            srv.close()
            srv.close()
        del srv
        gc.collect()
        if first is None:
            first = res
        for r in reqs:
            assert res[r.rid].loglik == first[r.rid].loglik
            np.testing.assert_array_equal(res[r.rid].grad, first[r.rid].grad)
            np.testing.assert_array_equal(res[r.rid].seis, first[r.rid].seis)


@pytest.mark.parametrize("loop", ["device", "host"])
def test_receiver_tables_growing_mid_stream_give_the_same_bits(loop):
    """Lanes that start with room for one receiver set grow their tables
    (and re-capture their rounds) as requests arrive; every request gets the
    bits of a server that never grows them."""
    from mun_amd.serving import SEMServer
    reqs = make_requests(6, seed=13)
    kw = dict(segment=60, capacity={8: 2, 12: 2, 20: 2}, use_graph=True, rec_pitch=32, nt_max=420, loop=loop)
    print(f"[serve-grow] {loop}: roomy tables", flush=True)
    base = SEMServer(hbm_budget_bytes=8e9, **kw).run(reqs)
    print(f"[serve-grow] {loop}: one set to start", flush=True)
    srv = SEMServer(hbm_budget_bytes=8e9, max_sets=1, **kw)
    grown = srv.run(reqs)
    assert any(l.max_sets > 1 for l in srv.lanes.values())
    for r in reqs:
        a, b = base[r.rid], grown[r.rid]
        assert a.loglik == b.loglik
        np.testing.assert_array_equal(a.grad, b.grad)
        np.testing.assert_array_equal(a.seis, b.seis)


@pytest.mark.parametrize("loop", ["device", "host"])
def test_requests_beyond_the_lane_limits_are_refused(loop):
    """A request with more receivers than the lane's pitch, or more time
    steps than the server's nt_max, raises ValueError before it takes arena
    memory; the server then serves valid requests as before."""
    import dataclasses

    from mun_amd.serving import SEMServer
    r = make_requests(1, seed=2)[0]
    srv = SEMServer(hbm_budget_bytes=2e9, segment=60, use_graph=False, rec_pitch=32, nt_max=r.nt, loop=loop)
    many = np.stack([np.linspace(20.0, r.xmax - 20.0, 40), np.full(40, 2.0)], axis=1)
    for bad, what in ((dataclasses.replace(r, rid=101, rec_xz=many, obs=np.zeros((r.nt, 40))), "receivers"),
                      (dataclasses.replace(r, rid=102, nt=r.nt + 1), "time steps")):
        with pytest.raises(ValueError, match=what):
            srv.run([bad])
        assert srv.in_use == 0
        for lane in srv.lanes.values():
            lane.queue.clear()
    ok = srv.run([r])[r.rid]
    assert np.isfinite(ok.loglik) and np.abs(ok.grad).max() > 0


@pytest.mark.parametrize("history", ["auto", "checkpoints"])
def test_exact_adjoint_served_matches_one_step_and_plain(history):
    """The exact adjoint (the damped forward's discrete adjoint) through the
    server: two steps per launch give the bits of one step per launch, and
    the results match the plain path in the same mode (its own launch
    sequence at batch 1)."""
    from mun_amd.serving import SEMServer, run_plain
    reqs = make_requests(5, seed=9)
    kw = dict(segment=60, capacity={8: 2, 12: 2, 20: 2}, use_graph=True, rec_pitch=32, nt_max=420,
              adjoint_mode="exact", history=history)
    print(f"[serve-exact] {history}: two steps per launch", flush=True)
    two = SEMServer(hbm_budget_bytes=8e9, two_step=True, **kw).run(reqs)
    print(f"[serve-exact] {history}: one step per launch", flush=True)
    one = SEMServer(hbm_budget_bytes=8e9, two_step=False, **kw).run(reqs)
    print(f"[serve-exact] {history}: plain path", flush=True)
    plain = run_plain(reqs, adjoint_mode="exact", segment=60)
    ref = SEMServer(hbm_budget_bytes=8e9, two_step=True, **dict(kw, adjoint_mode="reference")).run(reqs[:2])
    for r in reqs:
        a, b, p = two[r.rid], one[r.rid], plain[r.rid]
        assert a.loglik == b.loglik
        np.testing.assert_array_equal(a.grad, b.grad)
        np.testing.assert_array_equal(a.seis, b.seis)
        e = rel_l2(a.grad, p.grad)
        print(f"  request {r.rid} ({r.nelem_x}^2, {r.nt} steps): grad vs plain {e:.2e}", flush=True)
        assert abs(a.loglik - p.loglik) <= 1e-6 * abs(p.loglik)
        assert e < 1e-6
    # the mode reaches the adjoint: reference-mode gradients differ where the
    # wave energy enters the sponge
    assert any(rel_l2(ref[r.rid].grad, two[r.rid].grad) > 1e-4 for r in reqs[:2])


@pytest.mark.parametrize("loop", ["device", "host"])
def test_size_class_split_into_two_lanes_gives_the_same_bits(loop):
    """A size class split into two lanes (two streams sharing its capacity,
    each request routed to the less loaded) gives every request the bits of
    one lane per class; the classes with several requests use both lanes."""
    from mun_amd.serving import SEMServer
    reqs = make_requests(8, seed=3)
    kw = dict(segment=60, capacity={8: 4, 12: 4, 20: 2}, use_graph=True, rec_pitch=32, nt_max=420, loop=loop)
    print(f"[serve-split] {loop}: one lane per class", flush=True)
    one = SEMServer(hbm_budget_bytes=8e9, lanes_per_class=1, **kw).run(reqs)
    print(f"[serve-split] {loop}: two lanes per class", flush=True)
    srv = SEMServer(hbm_budget_bytes=8e9, lanes_per_class=2, **kw)
    two = srv.run(reqs)
    assert len(srv.lanes) == 2 * len(srv.class_of)
    assert all(sum(l.C for l in ls) == kw["capacity"][ls[0].key[0]] for ls in srv.class_of.values())
    counts = {}
    for r in reqs:
        counts[r.lane_key] = counts.get(r.lane_key, 0) + 1
    busy = [ls for k, ls in srv.class_of.items() if counts[k] >= 2]
    assert busy and all(l.rounds > 0 for ls in busy for l in ls)
    for r in reqs:
        a, b = one[r.rid], two[r.rid]
        assert a.loglik == b.loglik
        np.testing.assert_array_equal(a.grad, b.grad)
        np.testing.assert_array_equal(a.seis, b.seis)


@pytest.mark.parametrize("two_step", [False, True])
def test_device_loop_gives_the_host_loops_bits(two_step):
    """The device-resident request loop (queue, slot states and round
    scheduling on the device, one captured round replayed every round) gives
    every request the bits of the host-scheduled loop.  Lanes of one or two
    slots make requests wait in the device queue; the same requests pushed to
    a warm server again, without graphs, give the same bits too."""
    from mun_amd.serving import LoopLane, SEMServer
    reqs = make_requests(8, seed=5)
    kw = dict(segment=60, capacity={8: 2, 12: 2, 20: 1}, rec_pitch=32, nt_max=420, two_step=two_step)
    print(f"[serve-loop] two_step={two_step}: host loop", flush=True)
    host = SEMServer(hbm_budget_bytes=8e9, loop="host", use_graph=True, **kw).run(reqs)
    print(f"[serve-loop] two_step={two_step}: device loop", flush=True)
    srv = SEMServer(hbm_budget_bytes=8e9, loop="device", use_graph=True, **kw)
    dev = srv.run(reqs)
    assert all(isinstance(l, LoopLane) for l in srv.lanes.values())
    for lane in srv.lanes.values():
        c = lane.counts()
        print(f"  lane {lane.key[0]}^2: {c}")
        assert c["pushed"] == c["completed"] and c["active"] == 0
        # every round the host launched ran on the device (a captured round is
        # enqueued through the library once and replayed after that)
        assert c["rounds_done"] == lane.rounds and c["rounds_enqueued"] == 1
    # some request of a one-slot lane waited on the device queue for the slot
    big = [r for r in reqs if r.nelem_x == 20]
    assert len(big) >= 2
    srv.use_graph = False
    sets = {k: len(l.plan.rec_sizes) for k, l in srv.lanes.items()}
    print(f"[serve-loop] two_step={two_step}: device loop again, warm, direct launches", flush=True)
    again = srv.run(reqs)
    # the same receiver layouts reuse the lanes' sets
    assert {k: len(l.plan.rec_sizes) for k, l in srv.lanes.items()} == sets
    for r in reqs:
        for other in (dev, again):
            a, b = host[r.rid], other[r.rid]
            assert a.loglik == b.loglik
            np.testing.assert_array_equal(a.grad, b.grad)
            np.testing.assert_array_equal(a.seis, b.seis)


def test_served_stream_at_its_workload_sizes_alone_equals_mix():
    """The served workload itself (meshes 20^2 .. 160^2 deg-4, 1000-5000 steps,
    31-186 receivers, Poisson arrivals at 20/s), a reduced count of 8
    requests covering all four mesh sizes: every result in the mix has the
    bits of the same request served alone."""
    from mun_amd import workloads as W
    from mun_amd.serving import SEMServer
    print("[serve workload] requests", flush=True)
    reqs = W.served_solve_requests(8, 20.0, seed=7, nt_scale=1.0)
    assert {r.nelem_x for r in reqs} == {20, 40, 80, 160} and max(r.nt for r in reqs) == 5000
    srv = SEMServer(hbm_budget_bytes=120e9, segment=100, use_graph=True, nt_max=5000)
    print("[serve workload] mix", flush=True)
    mix = srv.run(reqs)
    assert len(mix) == len(reqs)
    adm = sorted((r.admitted, r.finished) for r in mix.values())
    assert any(adm[i + 1][0] < adm[i][1] for i in range(len(adm) - 1))
    for r in reqs:
        print(f"[serve workload] alone: request {r.rid} ({r.nelem_x}^2, {r.nt} steps)", flush=True)
        a = srv.run([r])[r.rid]
        b = mix[r.rid]
        assert a.loglik == b.loglik
        np.testing.assert_array_equal(a.grad, b.grad)
        np.testing.assert_array_equal(a.seis, b.seis)
