"""Benchmark: batched forward+adjoint SEM solves per second on one MI355X.

Workload (one "step" = one batch): 256 velocity models on a 200 x 200
element, degree-4 mesh (801 x 801 GLL nodes, 4 km x 4 km), 5-layer background
+ 16 x 16 cubic B-spline perturbations N(0, 150 m/s), 15 Hz Ricker source,
200 surface receivers, dt 0.25 ms, 6000 steps; each solve = forward + misfit
+ adjoint + gradient projected onto the 256 B-spline nodes, FP32 state with
FP64 gradient accumulation.

The batch runs under the autotuned schedule (mun_amd.autotune.tune_batch,
cached per device and workload): on one MI355X, groups of 8 solves that keep
their whole FP32 forward history (no recompute), two groups in flight on
separate streams, two time steps per step-kernel launch on 19 x 13-element
tiles, gradient correlation per chunk of 64 adjoint steps.

Printed beside it: the plain path (one solve per launch sequence, batch 1),
the live roofline of the time-step kernel (HIP events on the launch stream,
launched on the solves in flight, and on 64 solves streaming from HBM), the
whole solve's algorithmic bandwidth over the timed region, the ELBO and 40 x
40 workloads, the served stream of uneven requests, and the CPU restatement
(FP64, structured like the reference: assembled CSR stiffness + sparse
mat-vec per step, torch on all host cores) timed on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--nt NT]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# hardware queues per process for HIP (its default is 4), read when the GPU is
# first used: the served leg runs one stream per mesh-size lane and a
# finishing stream beside the default stream and torch's stream pool; lanes
# whose streams land on shared queues run one after the other (measured: the
# saturated served stream at 83 instead of 108 solves/s with 4 queues, 108
# with 16)
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

HBM_PEAK_GBS = 8000.0
# algorithmic bytes of a time step per node and solve, the standard two-level
# leapfrog: u^n and u^{n-1} read, u^{n+1} written, c^2 read (mass and sponge
# are analytic, 0 B); a launch of S steps is credited S times this
STEP_BYTES_PER_NODE = 16
# the increment form's own minimum per launch, one or two steps: u^n, w^n,
# c^2 read, the new u and w written (u^{n+1} of a two-step launch stays in LDS)
INCREMENT_BYTES_PER_NODE_LAUNCH = 20
# one forward+adjoint solve step with the full forward history: forward and
# adjoint 20 B each with one step per launch; with two per launch each pair
# reads 12 B and writes 12 B (u^{n+1} and u^{n+2} into the history or the
# adjoint ring, w^{n+2}), 12 B per step; correlation 8 B (u^{k+1}, lambda_k
# read), FP64 accumulator (25 doubles per element of 16 nodes = 12.5 B per
# node) read and written once per chunk of 64 steps
SOLVE_BYTES_PER_NODE_STEP = {1: 20 + 20 + 8 + 25 / 64, 2: 12 + 12 + 8 + 25 / 64}
LLC_BYTES = 256 << 20


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads():
    """Host cores this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set (the share a box gives one GPU job)."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return (min(aff, int(env)) if env.isdigit() and int(env) > 0 else aff), aff


def cpu_baseline(log, steps200=200, steps200_1t=30):
    """The FP64 CPU restatement in the reference's torch form (assembled CSR
    stiffness, one torch sparse mat-vec per step, dense 25 x 25 element
    gradient correlation; oracle/torch_cpu.py) on the host cores: the 40x40
    workload in full (2000 steps), and the 200x200 benchmark mesh over
    `steps200` forward + adjoint steps extrapolated to 6000 (and over
    `steps200_1t` on one thread).  Observations are zero (the work per step
    does not depend on them)."""
    import torch
    import oracle as O
    from oracle.torch_cpu import forward_adjoint_torch
    from mun_amd import workloads as W
    cores, aff = host_threads()
    prev = torch.get_num_threads()

    def run(wl, nsteps, threads):
        m = wl.mesh
        prob = O.Problem(m.xmin, m.xmax, m.zmin, m.zmax, m.nelem_x, m.nelem_z, 4, m.dt, m.nt,
                         tuple(wl.sources[0]), wl.f0, wl.receivers, sponge_width=m.sponge_width,
                         sponge_cmax=m.sponge_cmax, sponge_sides=m.sponge_sides)
        disc = O.discretize(prob)
        c, _, _ = O.tensor_bspline_velocity(disc.mesh, wl.coef[0], wl.base)
        r = forward_adjoint_torch(disc, c, np.zeros((m.nt, wl.receivers.shape[0])), 1e-9, nsteps=nsteps,
                                  threads=threads)
        t = r["times"]
        per_step = (t["forward"] + t["adjoint"]) / r["steps"]
        return t["assembly"], per_step, t["assembly"] + per_step * m.nt, r["steps"]

    try:
        a1, ps1, t1, n1 = run(W.cpu40(), None, cores)
        log(f"[bench] cpu 40x40: {t1:.2f} s per solve ({n1} steps, {cores} threads)")
        a2, ps2, t2, n2 = run(W.batch200(nmodel=1), steps200, cores)
        log(f"[bench] cpu 200x200: {ps2 * 1e3:.1f} ms per step, {t2:.0f} s per solve ({cores} threads)")
        a3, ps3, t3, n3 = run(W.batch200(nmodel=1), steps200_1t, 1)
        log(f"[bench] cpu 200x200: {ps3 * 1e3:.1f} ms per step, {t3:.0f} s per solve (1 thread)")
    finally:
        torch.set_num_threads(prev)
    return {"value": round(1.0 / t2, 6), "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": (f"FP64 torch-CPU restatement of the reference path (CSR stiffness, one sparse mat-vec per "
                       f"step, dense 25x25 element correlation) on {cores} threads (affinity {aff}): 200x200 deg-4 "
                       f"mesh, assembly {a2:.2f} s + {n2} forward+adjoint steps at {ps2 * 1e3:.1f} ms per step, "
                       f"extrapolated to 6000 steps = {t2:.0f} s per solve (extrapolated, not run in full)"),
            "batch_seconds_sequential": round(256 * t2, 1),
            "one_thread": {"value": round(1.0 / t3, 6), "cores": 1, "ms_per_step": round(ps3 * 1e3, 2),
                           "sample": f"200x200, {n3} steps extrapolated to 6000 = {t3:.0f} s per solve"},
            "cpu40_full": {"value": round(1.0 / t1, 4), "cores": cores, "seconds": round(t1, 2),
                           "sample": f"40x40 deg-4, all {n1} steps forward+adjoint, assembly {a1:.2f} s"}}


def extra_workloads(log):
    """The other seeded workloads: 40x40 (batched 256 models vs the plain
    one-solve path) and the 64-sample x 8-shot ELBO batch."""
    import torch
    from mun_amd import workloads as W
    from mun_amd.batch import BatchSolver
    dev = torch.device("cuda:0")
    res = {}
    # cpu40: 40x40 deg-4, 2000 steps; 256 models in one batch vs one at a time
    wl = W.cpu40()
    nb = 256
    rng = np.random.default_rng(11)
    coefs = wl.coef[0][None] + rng.normal(0.0, 50.0, (nb,) + wl.coef.shape[1:])
    wl.coef = coefs
    s = BatchSolver.from_workload(wl, precision="fp32", device=dev)
    obs = s.forward(torch.as_tensor(wl.true_coef[None], device=dev))[0]
    c = torch.as_tensor(coefs, device=dev)
    s.loglik_grad(c, obs, wl.noise_std)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.loglik_grad(c, obs, wl.noise_std)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    wl1 = W.cpu40()
    s1 = BatchSolver.from_workload(wl1, precision="fp32", device=dev)
    s1.loglik_grad(c[:1], obs, wl.noise_std)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(4):
        s1.loglik_grad(c[k:k + 1], obs, wl.noise_std)
    torch.cuda.synchronize()
    tp = (time.perf_counter() - t0) / 4
    res["cpu40"] = {"workload": "40x40 deg-4 (161x161 nodes), 2000 steps, fwd+adjoint, 8x8 B-spline nodes",
                    "batched_solves_per_s": round(nb / tb, 2), "batch": nb,
                    "tile_rows": s.plan.tile_rows, "plain_solves_per_s": round(1.0 / tp, 2),
                    "batched_vs_plain": round(tp * nb / tb, 2)}
    log(f"[bench] cpu40: batched {nb / tb:.1f}/s, plain {1 / tp:.1f}/s, ratio {tp * nb / tb:.1f}")
    del s, s1
    torch.cuda.empty_cache()
    # elbo120: 64 samples x 8 shots, 120x120 deg-4, 4000 steps
    wl = W.elbo120()
    s = BatchSolver.from_workload(wl, precision="fp32", device=dev)
    obs = s.forward(torch.as_tensor(wl.true_coef[None], device=dev))[0]
    c = torch.as_tensor(wl.coef, device=dev)
    t0 = time.perf_counter()
    ll, g, _ = s.loglik_grad(c, obs, wl.noise_std)
    elbo_grad = g.mean(0)
    torch.cuda.synchronize()
    te = time.perf_counter() - t0
    res["elbo120"] = {"workload": "64 samples x 8 shots, 120x120 deg-4 (481x481 nodes), 4000 steps, "
                                  "mean gradient over 12x12 B-spline nodes",
                      "solves_per_s": round(wl.nmodel * wl.nshot / te, 2), "seconds": round(te, 2),
                      "tile_rows": s.plan.tile_rows, "mean_grad_norm": float(elbo_grad.norm())}
    log(f"[bench] elbo120: {te:.1f} s for {wl.nmodel * wl.nshot} solves")
    del s
    torch.cuda.empty_cache()
    res["elbo120"]["per_shot_rel_l2"] = per_shot_errors(wl, obs, (0, 17, 63))
    log(f"[bench] elbo120 per-shot FP32 vs FP64: {res['elbo120']['per_shot_rel_l2']}")
    return res


def per_shot_errors(wl, obs, samples):
    """FP32 against the FP64 build of the same kernels for every shot of a few
    samples of the ELBO batch (FP64 itself matches the oracle to ~1e-11):
    relative L2 of seismograms and of the B-spline node gradient, and the
    relative log-likelihood error, each the largest over the samples."""
    import torch
    from mun_amd.batch import BatchSolver
    out = {}
    for prec in ("fp64", "fp32"):
        s = BatchSolver.from_workload(wl, tile_rows=15, precision=prec, device="cuda:0")
        _, c2 = s.velocity(torch.as_tensor(wl.coef[list(samples)], device="cuda:0"))
        src, model_of = s._expand(len(samples))
        ob = obs.double()[None].expand(len(samples), -1, -1, -1).reshape(len(src), wl.mesh.nt, -1)
        o = s.plan.forward_adjoint(c2, s.wavelet, src, ob, wl.noise_std, [s.rec_set] * len(src), model_of=model_of)
        gn = s.plan.bspline_project(o["grad_c"], s.nbz, s.nbx)
        out[prec] = (o["seis"].double().cpu().numpy(), gn.cpu().numpy(), o["loglik"].cpu().numpy())
        del s, o
        torch.cuda.empty_cache()

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    ns = wl.nshot
    shots = {"seis": [], "grad": [], "loglik": []}
    for k in range(ns):
        rows = [i * ns + k for i in range(len(samples))]
        shots["seis"].append(max(rel(out["fp32"][0][b], out["fp64"][0][b]) for b in rows))
        shots["grad"].append(max(rel(out["fp32"][1][b], out["fp64"][1][b]) for b in rows))
        shots["loglik"].append(max(abs(out["fp32"][2][b] / out["fp64"][2][b] - 1.0) for b in rows))
    return {"samples": list(samples), **{k: [float(f"{v:.3g}") for v in vs] for k, vs in shots.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--nt", type=int, default=6000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-passes", type=int, default=4,
                    help="timed passes of the stream per engine; latencies pooled over them")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-nt-scale", type=float, default=1.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the cpu40 and elbo120 workloads")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from mun_amd import workloads as W
    from mun_amd.batch import BatchSolver

    dev = torch.device("cuda:0")
    wl = W.batch200(nmodel=args.batch, nt=args.nt)
    m = wl.mesh
    log(f"[bench] workload {wl.name}: batch {args.batch}, mesh {m.nelem_x}x{m.nelem_z} (N={m.N}), nt={m.nt}")
    t_tune = time.perf_counter()
    solver = BatchSolver.from_workload(wl, precision="fp32", adjoint_mode="reference", device=dev)
    plan = solver.plan
    if solver.tuning:
        log(f"[bench] tuning ({solver.tuning.get('source', 'timed')}, "
            f"{time.perf_counter() - t_tune:.1f} s): " +
            ", ".join(f"{k} {solver.tuning[k]}" for k in ("tile2", "group", "lanes", "chunk")))
    coef = torch.as_tensor(wl.coef, device=dev)
    # observations from the seeded true model (one forward solve)
    obs = solver.forward(torch.as_tensor(wl.true_coef[None], device=dev))[0]
    torch.cuda.synchronize()
    noise = 1e-9

    def step():
        return solver.loglik_grad(coef, obs, noise)

    for i in range(args.warmup):
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        log(f"[bench] warmup {i}: {time.perf_counter() - t0:.2f} s")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ll, grad, _ = step()
    torch.cuda.synchronize()
    dt_total = time.perf_counter() - t0
    ms_per_step = dt_total / args.steps * 1e3
    value = args.batch * args.steps / dt_total
    log(f"[bench] timed {args.steps} batch(es): {ms_per_step:.1f} ms/batch, {value:.3f} solves/s")
    assert bool(torch.isfinite(ll).all()) and bool(torch.isfinite(grad).all())
    sched = plan.get_schedule()
    nsolve = args.batch * wl.nshot
    if sched["group"] > 0:
        ngroups = -(-nsolve // sched["group"])
        ngroups = min(nsolve, -(-ngroups // sched["lanes"]) * sched["lanes"])
        group = -(-nsolve // ngroups)
        inflight = min(nsolve, group * sched["lanes"])
        sched_txt = (f"{ngroups} groups of <= {group} solves, {sched['lanes']} in flight on separate streams, "
                     + ("full forward history" if sched["checkpoint_every"] <= 0
                        else f"checkpoints every {sched['checkpoint_every']} steps"))
    else:
        inflight = nsolve
        sched_txt = ("whole batch at once, " + ("full forward history" if sched["checkpoint_every"] <= 0
                                                else f"checkpoints every {sched['checkpoint_every']} steps"))
    log(f"[bench] schedule: {sched_txt}")
    fusion = plan.get_step_fusion()
    solve_gbs = SOLVE_BYTES_PER_NODE_STEP[fusion] * m.N * m.nt * nsolve * args.steps / dt_total / 1e9

    # live roofline of the time-step kernel, launched on as many solves as the
    # schedule keeps in flight: a forward-only pass is nt / steps-per-launch
    # launches of the step kernel (plus one receiver pass), timed with events
    # on the stream the library launches on; measured for the plan's launch
    # form (two steps per launch by default), for one step per launch, and for
    # the two-step kernel on a launch too large for the last-level cache
    _, c2 = solver.velocity(coef)
    src, model_of = solver._expand(args.batch)
    st = torch.cuda.current_stream(dev)
    E = plan.layout()[1]
    nts = inflight * 4 * E * 4 > LLC_BYTES  # the library's choice of streams

    def time_forward(steps, n, streaming="auto"):
        plan.set_step_fusion(steps)
        plan.set_streaming(streaming)
        args_ = (c2[:n], solver.wavelet, src[:n], [solver.rec_set] * n)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        plan.forward(*args_, model_of=model_of[:n])
        e0.record(st)
        plan.forward(*args_, model_of=model_of[:n])
        e1.record(st)
        torch.cuda.synchronize()
        plan.set_step_fusion(fusion)
        plan.set_streaming("auto")
        return e0.elapsed_time(e1) / 1e3 / -(-m.nt // steps)

    t_single = time_forward(1, inflight)
    t_launch = time_forward(fusion, inflight)
    n_hbm = min(args.batch, 64)
    t_hbm = time_forward(fusion, n_hbm, "nontemporal")

    def gbs(bytes_, t):
        return round(bytes_ / t / 1e9, 1)

    bytes_launch = STEP_BYTES_PER_NODE * fusion * m.N * inflight
    achieved = bytes_launch / t_launch / 1e9
    _, t2x, t2z = plan.get_tile2()
    kernel = (f"step2_kernel<float,4,MODE_FWD,NTS={str(nts).lower()},{t2x},{t2z}>" if fusion == 2
              else f"step_kernel<float,4,MODE_FWD,{plan.tile_rows},NTS={str(nts).lower()}>")
    log(f"[bench] step kernel on {inflight} solves, {fusion} step(s) per launch: {t_launch * 1e6:.1f} us/launch "
        f"({t_launch / fusion * 1e6:.1f} us/step), {achieved:.0f} GB/s at 16 B per node-step; one step per launch: "
        f"{t_single * 1e6:.1f} us/step; {n_hbm} solves non-temporal: {t_hbm * 1e6:.1f} us/launch; whole solve "
        f"{solve_gbs:.0f} GB/s")
    # bytes per launch across L2 from the committed counter runs: exact when
    # they were taken at this launch size, else their measured ratio to the
    # increment form's algorithmic bytes applied to this launch (noted in traffic_basis)
    traffic, traffic_basis = None, None
    inc_bytes = INCREMENT_BYTES_PER_NODE_LAUNCH * m.N * inflight
    tp = os.path.join(ROOT, "profiles", "step_fwd_batch200_traffic.json")
    if os.path.exists(tp):
        with open(tp) as f:
            tj = json.load(f)
        fw = tj.get("forward", {})
        if fw.get("kernel", "").split("<")[0].split("::")[-1] != kernel.split("<")[0]:
            fw, tj = {}, {}
        if tj.get("batch") == inflight:
            traffic, traffic_basis = tj.get("bytes_per_launch"), f"counters at {inflight} solves per launch"
        elif fw.get("ratio"):
            traffic = int(fw["ratio"] * inc_bytes)
            traffic_basis = (f"counters at {tj.get('batch')} solves per launch: {fw['ratio']} x the increment "
                             f"form's bytes, applied to {inflight}")
    ws_mb = inflight * 4 * E * 4 / 2 ** 20

    out = {"metric": "fwd+adjoint SEM solves/s (1 GPU, batched vs plain); p50/p99 solve latency",
           "value": round(value, 4), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"batch200: {args.batch} models x (200x200 deg-4 elements, 801x801 nodes, "
                                  f"{m.nt} steps) forward+adjoint+gradient; {sched_txt}",
                      "batch": args.batch, "nt": m.nt, "nodes": m.N, "schedule": sched,
                      "tuned": ({k: solver.tuning[k] for k in ("tile2", "group", "lanes", "chunk", "source")}
                                if solver.tuning else None),
                      "tile2": list(plan.get_tile2()), "parallelism": "none (1 GPU)"},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                        "traffic_basis": traffic_basis, "kernel": kernel,
                        "basis": (f"{STEP_BYTES_PER_NODE} B per node-step (u^n, u^(n-1), c^2 read, u^(n+1) "
                                  f"written) x {fusion} steps per launch x {m.N} nodes x {inflight} solves "
                                  f"= {bytes_launch} B per launch"),
                        "steps_per_launch": fusion, "solves_per_launch": inflight,
                        "launch_us": round(t_launch * 1e6, 1), "us_per_step": round(t_launch / fusion * 1e6, 2),
                        "cache_resident": (f"the {inflight} solves' per-step working set (u^n, u^(n+1), w, c^2: "
                                           f"{ws_mb:.0f} MiB) fits the 256 MiB last-level cache, so much of "
                                           f"this launch's traffic is served there, not from HBM"),
                        "limiter": ("VALU work and memory latency, not bandwidth: the two-step kernel keeps "
                                    "u^(n+1) in LDS; its counters show the VALU carrying 71 % of the launch's "
                                    "cycles and waves waiting 43 % of their lifetime at four LDS-bound workgroups "
                                    "per compute unit (profiles/step_kernels_batch200_sq_ta_counters.md)"
                                    if fusion == 2 else "HBM / last-level cache bandwidth"),
                        "hbm_resident": {"solves_per_launch": n_hbm, "streams": "non-temporal",
                                         "launch_us": round(t_hbm * 1e6, 1),
                                         "achieved": gbs(STEP_BYTES_PER_NODE * fusion * m.N * n_hbm, t_hbm),
                                         "frac": round(STEP_BYTES_PER_NODE * fusion * m.N * n_hbm / t_hbm / 1e9
                                                       / HBM_PEAK_GBS, 4)},
                        "increment_form": {"bytes_per_node_launch": INCREMENT_BYTES_PER_NODE_LAUNCH,
                                           "achieved": gbs(inc_bytes, t_launch),
                                           "frac": round(inc_bytes / t_launch / 1e9 / HBM_PEAK_GBS, 4)},
                        "one_step_per_launch": {
                            "kernel": f"step_kernel<float,4,MODE_FWD,{plan.tile_rows},NTS={str(nts).lower()}>",
                            "launch_us": round(t_single * 1e6, 1),
                            "achieved": gbs(STEP_BYTES_PER_NODE * m.N * inflight, t_single),
                            "frac": round(STEP_BYTES_PER_NODE * m.N * inflight / t_single / 1e9 / HBM_PEAK_GBS, 4)},
                        "whole_solve": {"bytes_per_node_step": round(SOLVE_BYTES_PER_NODE_STEP[fusion], 3),
                                        "achieved": round(solve_gbs, 1),
                                        "frac": round(solve_gbs / HBM_PEAK_GBS, 4),
                                        "basis": "the schedule run: full forward history, forward and adjoint "
                                                 "two steps per launch, correlation per chunk of 64 steps"}},
           "latency_s": {"p50": round(ms_per_step / 1e3, 3), "p99": round(ms_per_step / 1e3, 3),
                         "note": "batched path: every solve completes with its batch"}}

    if not args.no_plain:
        plain = BatchSolver(plan, wl.sources, wl.receivers, wl.f0, solver.nbz, solver.nbx, wl.base)
        plain.loglik_grad(coef[:1], obs, noise)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plain.loglik_grad(coef[:1], obs, noise)
        torch.cuda.synchronize()
        tp1 = time.perf_counter() - t0
        out["plain_solves_per_s"] = round(1.0 / tp1, 4)
        out["batched_vs_plain"] = round(value * tp1, 3)
        log(f"[bench] plain path: {tp1:.2f} s per solve")
        # the same plain path with one launch per time step (the one-step kernel)
        fusion0 = plan.get_step_fusion()
        plan.set_step_fusion(1)
        plain.loglik_grad(coef[:1], obs, noise)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plain.loglik_grad(coef[:1], obs, noise)
        torch.cuda.synchronize()
        tp1s = time.perf_counter() - t0
        plan.set_step_fusion(fusion0)
        out["plain_one_step_per_launch"] = {"solves_per_s": round(1.0 / tp1s, 4),
                                            "batched_vs_plain": round(value * tp1s, 3)}
        log(f"[bench] plain path, one step per launch: {tp1s:.2f} s per solve")
    del solver, plan, coef, obs, c2, ll, grad
    if not args.no_plain:
        del plain
    torch.cuda.empty_cache()
    if not args.no_extra:
        out["workloads"] = extra_workloads(log)
    if not args.no_serve:
        from mun_amd.serving import SEMServer, latency_stats, run_plain
        log(f"[bench] served stream: {args.serve_requests} requests, Poisson {args.serve_rate}/s, "
            f"meshes 20..160^2, nt scale {args.serve_nt_scale}")
        import dataclasses
        import gc
        reqs = W.served_solve_requests(args.serve_requests, args.serve_rate, nt_scale=args.serve_nt_scale)
        at_once = [dataclasses.replace(r, arrival=0.0) for r in reqs]
        torch.cuda.synchronize()

        def serve(loop):
            # a warm server (lanes, captured rounds, tile choices, its arena),
            # as in service: the stream at its arrival times, then saturated
            # (every request present at t = 0: throughput, not latency)
            srv = SEMServer(hbm_budget_bytes=120e9, segment=100, use_graph=True, nt_max=5000, loop=loop)
            srv.run([dataclasses.replace(r, arrival=0.0) for r in reqs], realtime=False)
            # the stream at its arrival times, passes pooled (p99 over
            # passes x requests latencies; throughput the passes' mean)
            pooled, passes = {}, []
            for k in range(max(1, args.serve_passes)):
                res = srv.run([dataclasses.replace(r) for r in reqs], progress=log)
                passes.append(latency_stats(res))
                pooled.update({(k, rid): v for rid, v in res.items()})
            st = latency_stats(pooled)
            st["solves_per_s"] = float(np.mean([p_["solves_per_s"] for p_ in passes]))
            st["makespan_s"] = float(np.mean([p_["makespan_s"] for p_ in passes]))
            st["passes"] = len(passes)
            sat_ = latency_stats(srv.run(at_once))
            srv.close()
            del srv
            gc.collect()
            torch.cuda.synchronize()
            torch.cuda.empty_cache()
            return st, sat_

        served, sat = serve("device")
        served_host, sat_host = serve("host")
        plans = {}
        run_plain([dataclasses.replace(r, arrival=0.0) for r in reqs], plans=plans)
        plain = latency_stats(run_plain(reqs, progress=log, plans=plans))
        sat_plain = latency_stats(run_plain(at_once, plans=plans))
        out["served"] = {"workload": f"{args.serve_requests} Poisson({args.serve_rate}/s) forward+adjoint requests, "
                                     f"meshes {{20,40,80,160}}^2 deg-4, {int(1000 * args.serve_nt_scale)}-"
                                     f"{int(5000 * args.serve_nt_scale)} steps, 20-200 receivers",
                         "warm": "every path served the stream once before timing",
                         "latency_samples": (f"continuous: {args.serve_passes} passes of the stream pooled "
                                             f"({args.serve_passes * args.serve_requests} latencies); plain: one pass"),
                         "engine": "device-resident request loop (queue, slot states and round scheduling on "
                                   "the device; one captured round replayed)",
                         "continuous": served, "plain_one_at_a_time": plain,
                         "saturated_continuous": sat, "saturated_plain": sat_plain,
                         "saturated_throughput_ratio": round(sat["solves_per_s"] / sat_plain["solves_per_s"], 3),
                         "host_scheduled_loop": {"continuous": served_host, "saturated_continuous": sat_host}}
        out["latency_s"] = {"p50": round(served["p50_s"], 4), "p99": round(served["p99_s"], 4),
                            "note": "served stream (continuous batching), arrival to result"}
        log(f"[bench] served p50 {served['p50_s']:.3f}s p99 {served['p99_s']:.3f}s {served['solves_per_s']:.2f}/s; "
            f"plain p50 {plain['p50_s']:.3f}s p99 {plain['p99_s']:.3f}s {plain['solves_per_s']:.2f}/s")
    if not args.no_cpu:
        log("[bench] cpu baseline (bounded sample) ...")
        out["cpu_baseline"] = cpu_baseline(log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
