"""Self-written autotuner: the step kernels' tile shapes and the batched
path's schedule.

The step kernel has instances for tile heights of 15, 7 and 3 element rows
(256-, 128- and 64-thread workgroups; every tile is 16 owned elements wide so
rows stay 256-byte aligned).  Tall tiles waste less on the recomputed halo
row; short tiles give more workgroups, which small meshes and small batches
need to fill 256 CUs.  ``best_tile_rows`` times the forward step of every
instance on the actual mesh and batch size (HIP events around a run of
launches on the library's stream), keeps the fastest and caches the choice in
a JSON file keyed by device, mesh, batch and precision.  The tuned height
applies to one-step launches (checkpointed forward, recompute, one-step slot
launches, snapshots).

``tune_batch`` tunes what runs the batched forward+adjoint: the two-step
kernel's tile shape (four instances), the full-history group size and lane
count, and the gradient-correlation chunk, timed on the real mesh at the
in-flight batch and kept to configurations whose workspace fits memory.
``best_tile2`` times the two-step tile shapes for the serving lanes.
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, Optional

import numpy as np
import torch

TILE_ROWS = (15, 7, 3)
_MEM: Dict[str, dict] = {}


def cache_path() -> str:
    """MUN_AMD_AUTOTUNE_CACHE, else ~/.cache/mun_amd/autotune.json when the
    home directory is writable, else a file in the temporary directory."""
    env = os.environ.get("MUN_AMD_AUTOTUNE_CACHE")
    if env:
        return env
    home = os.path.join(os.path.expanduser("~"), ".cache", "mun_amd")
    try:
        os.makedirs(home, exist_ok=True)
        if os.access(home, os.W_OK):
            return os.path.join(home, "autotune.json")
    except OSError:
        pass
    import tempfile
    return os.path.join(tempfile.gettempdir(), f"mun_amd_autotune_{os.getuid()}.json")


def _load() -> dict:
    p = cache_path()
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


SHIPPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_mi355x.json")


def _shipped() -> dict:
    """Batch configurations measured on MI355X for the benchmark workloads
    (tune_batch's choice confirmed by full-size runs), used when the cache
    has no entry: a short probe can rank two configurations within 2 % of
    each other either way.  MUN_AMD_AUTOTUNE=retune ignores it and the cache."""
    try:
        with open(SHIPPED) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db: dict):
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + f".{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass


def device_tag(device) -> str:
    """Cache key of a device: its compute-unit count and memory size (the
    marketing name reads empty under some profilers)."""
    props = torch.cuda.get_device_properties(device)
    return f"cu{props.multi_processor_count}_mem{props.total_memory >> 30}G"


def _key(nelem_x, nelem_z, batch, precision, device) -> str:
    return f"{device_tag(device)}|{nelem_x}x{nelem_z}|b{batch}|{precision}"


def measure(nelem_x: int, nelem_z: int, batch: int, tile_rows: int, precision: str = "fp32", steps: int = 30,
            device=None) -> float:
    """Mean forward-step time (seconds) of one tile shape on a synthetic
    constant-velocity batch, one step per launch (the tile height is the
    one-step kernel's; two-step launches use a fixed tile)."""
    from .solver import Plan
    dev = torch.device(device if device is not None else "cuda")
    L = 25.0
    p = Plan(nelem_x, nelem_z, 0.0, L * nelem_x, 0.0, L * nelem_z, 2e-4, steps, 4, 4 * L, 3000.0, 1e-8,
             ("xmin", "xmax", "zmax"), precision=precision, checkpoint_every=None, device=dev, tile_rows=tile_rows)
    p.set_step_fusion(1)
    rs = p.add_receivers(np.array([[L, 0.0]]))
    c2 = torch.full((batch, p.N), 2000.0 ** 2, dtype=p.dtype, device=dev)
    w = torch.zeros(steps, dtype=p.dtype, device=dev)
    src = [p.source_node(L * nelem_x / 2, L)] * batch
    p.forward(c2, w, src, [rs] * batch)  # warm-up (workspace, receivers, code objects)
    st = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    p.forward(c2, w, src, [rs] * batch)
    e1.record(st)
    torch.cuda.synchronize(dev)
    return e0.elapsed_time(e1) / 1e3 / steps


def best_tile_rows(nelem_x: int, nelem_z: int, batch: int, precision: str = "fp32", device=None,
                   steps: int = 30, use_cache: bool = True, report: Optional[dict] = None) -> int:
    """Fastest tile height for this mesh and batch (timed once, then cached)."""
    dev = torch.device(device if device is not None else "cuda")
    key = _key(nelem_x, nelem_z, batch, precision, dev)
    if use_cache:
        if key in _MEM:
            hit = _MEM[key]
        else:
            hit = _load().get(key)
        if hit and int(hit.get("tile_rows", -1)) in TILE_ROWS:  # entries of other kernel builds are re-timed
            _MEM[key] = hit
            if report is not None:
                report.update(hit)
            return int(hit["tile_rows"])
    times = {}
    for ez in TILE_ROWS:
        times[ez] = measure(nelem_x, nelem_z, batch, ez, precision, steps, dev)
    best = min(times, key=times.get)
    rec = {"tile_rows": best, "step_us": {str(k): round(v * 1e6, 2) for k, v in times.items()},
           "tuned_at": time.strftime("%Y-%m-%dT%H:%M:%S")}
    _MEM[key] = rec
    if use_cache:
        db = _load()
        db[key] = rec
        _store(db)
    if report is not None:
        report.update(rec)
    return best


# ---------------------------------------------------------------- batched path
TILE2_SHAPES = {"fp32": (0, 1, 2, 3, 4), "fp64": (0, 1, 2, 4)}


def _probe_plan(plan, steps: int, nrec: int):
    """A plan of the same mesh, sponge, precision and adjoint mode as `plan`,
    `steps` steps long, full forward history, with `nrec` surface receivers."""
    from .solver import Plan
    d = plan.desc
    sides = tuple(k for k, bit in (("xmin", 1), ("xmax", 2), ("zmin", 4), ("zmax", 8)) if d.sponge_sides & bit)
    p = Plan(d.nelem_x, d.nelem_z, d.xmin, d.xmax, d.zmin, d.zmax, d.dt, steps, 4, d.sponge_width, d.sponge_cmax,
             d.sponge_R, sides, precision=plan.precision, adjoint_mode=plan.adjoint_mode, checkpoint_every=0,
             device=plan.device, tile_rows=plan.tile_rows)
    n = max(int(nrec), 1)
    rs = p.add_receivers(np.stack([np.linspace(d.xmin, d.xmax, n), np.full(n, d.zmin)], axis=1))
    return p, rs


def _time_solves(p, rs, n: int, steps: int, repeats: int = 1) -> float:
    """Seconds per solve-step of a forward+adjoint batch of n solves (the
    fastest of `repeats` timed runs after one warm-up run)."""
    dev = p.device
    c2 = torch.full((n, p.N), 2500.0 ** 2, dtype=p.dtype, device=dev)
    w = p.wavelet(15.0)
    src = [p.source_node(0.5 * (p.desc.xmin + p.desc.xmax), p.desc.zmin + 50.0)] * n
    pitch = p.rec_sizes[rs]
    obs = torch.zeros((n, steps, pitch), dtype=torch.float64, device=dev)
    p.forward_adjoint(c2, w, src, obs, 1e-9, [rs] * n)
    st = torch.cuda.current_stream(dev)
    best = float("inf")
    for _ in range(repeats):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        p.forward_adjoint(c2, w, src, obs, 1e-9, [rs] * n)
        e1.record(st)
        torch.cuda.synchronize(dev)
        best = min(best, e0.elapsed_time(e1) / 1e3)
    return best / (n * steps)


def tune_batch(plan, batch: int, pitch: int, probe_steps: int = 192, groups=(4, 5, 6, 7, 8), lanes=(2, 3),
               chunks=(16, 32, 64), max_probe_batch: int = 512, use_cache: bool = True,
               report: Optional[dict] = None, log=None) -> dict:
    """Tune how ``plan`` runs forward+adjoint batches of `batch` solves: the
    two-step kernel's tile shape, the full-history group size and lane count,
    and the gradient-correlation chunk.  Candidates are those whose workspace
    fits the plan's memory budget at its real step count; each is timed on a
    probe plan of the same mesh and precision, `probe_steps` steps long, on
    the real batch size (at most `max_probe_batch` solves, so that the
    rounds per lane and the tail of the last round are those of the real
    run; the fastest of two timed runs), by coordinate descent (shape, then
    group x lanes, the shape again at that schedule, then chunk).  The
    result is cached (JSON, keyed by device, mesh, precision, adjoint mode,
    batch and steps) and applied to the plan (set_tile2, set_corr_chunk,
    plan.tuned for schedule_auto).  Before timing, the in-process table, the
    cache file and the configurations shipped for MI355X (autotune_mi355x.json)
    are consulted in that order; an entry is used only if it still fits the
    current budget.  Returns the configuration with its timings and its
    source ('memory', 'cache', 'shipped' or 'timed')."""
    dev = plan.device
    budget = plan.memory_budget()
    saved = plan.get_schedule()
    # the library planner's own choice is always a candidate (with half its
    # group): small meshes want a few large groups, not groups of 4-8
    try:
        plan.tuned = None
        lib = plan.schedule_auto(batch, pitch, budget)
    except Exception:
        lib = {"checkpoint_every": 100, "group": 0, "lanes": 1}
    cand_g, cand_l = set(groups), set(lanes)
    if lib["checkpoint_every"] <= 0:
        g_lib = lib["group"] if lib["group"] > 0 else batch
        cand_g |= {g_lib, max(1, g_lib // 2)}
        cand_l.add(lib["lanes"])
    feasible = []
    for g in sorted(cand_g):
        for ln in sorted(cand_l):
            if (ln - 1) * g >= batch:  # fewer than `ln` groups of at most g solves
                continue
            plan.set_schedule(0, g, ln)
            if plan.workspace_bytes(batch, pitch, True) <= budget:
                feasible.append((g, ln))
    plan.set_schedule(saved["checkpoint_every"], saved["group"], saved["lanes"])
    if not feasible:
        return {}
    d = plan.desc
    key = (f"{device_tag(dev)}|batch|{d.nelem_x}x{d.nelem_z}|{plan.precision}|"
           f"{plan.adjoint_mode}|b{batch}|nt{plan.nt}|p{probe_steps}")
    shapes = TILE2_SHAPES[plan.precision]

    def usable(r):
        return bool(r) and r.get("tile2") in shapes and (r.get("group"), r.get("lanes")) in \
            [tuple(f) for f in feasible] and r.get("chunk") in chunks

    rec, source = None, "timed"
    if use_cache and os.environ.get("MUN_AMD_AUTOTUNE") != "retune":
        for source, table in (("memory", _MEM), ("cache", None), ("shipped", None)):
            if table is None:
                table = _load() if source == "cache" else _shipped()
            if usable(table.get(key)):
                rec = table[key]
                break
        else:
            source = "timed"
    cached = rec is not None
    if rec is None:
        p, rs = _probe_plan(plan, probe_steps, pitch)
        times = {}

        def run(shape, g, ln, ck):
            k = f"tile2={shape} group={g} lanes={ln} chunk={ck}"
            if k not in times:
                p.set_tile2(shape)
                p.set_corr_chunk(ck)
                p.set_schedule(0, g, ln)
                times[k] = _time_solves(p, rs, min(batch, max_probe_batch), probe_steps, repeats=2)
                if log:
                    log(f"[autotune] {k}: {times[k] * 1e6:.3f} us per solve-step")
            return times[k]

        g_lib = lib["group"] if lib["group"] > 0 else batch
        start = (g_lib, lib["lanes"]) if (g_lib, lib["lanes"]) in feasible else None
        g0, l0 = start or max(feasible, key=lambda f: (f[0] * f[1], f[1]))
        shape = min(shapes, key=lambda s_: run(s_, g0, l0, 64))
        g0, l0 = min(feasible, key=lambda f: run(shape, f[0], f[1], 64))
        # the shape again at the chosen schedule (the ranking of shapes can
        # depend on the solves per launch), then the schedule once more if it moved
        shape2 = min(shapes, key=lambda s_: run(s_, g0, l0, 64))
        if shape2 != shape:
            shape = shape2
            g0, l0 = min(feasible, key=lambda f: run(shape, f[0], f[1], 64))
        ck = min(chunks, key=lambda c_: run(shape, g0, l0, c_))
        rec = {"tile2": shape, "group": g0, "lanes": l0, "chunk": ck, "probe_steps": probe_steps,
               "us_per_solve_step": {k: round(v * 1e6, 4) for k, v in times.items()},
               "tuned_at": time.strftime("%Y-%m-%dT%H:%M:%S")}
        del p
        torch.cuda.empty_cache()
        if use_cache:
            db = _load()
            db[key] = rec
            _store(db)
    _MEM[key] = rec
    plan.set_tile2(rec["tile2"])
    plan.set_corr_chunk(rec["chunk"])
    plan.tuned = {"group": rec["group"], "lanes": rec["lanes"]}
    if report is not None:
        report.update(rec)
    return dict(rec, cached=cached, source=source)


def best_tile2(nelem_x: int, nelem_z: int, batch: int, precision: str = "fp32", device=None, steps: int = 30,
               use_cache: bool = True) -> int:
    """Fastest two-step tile shape for forward steps of `batch` solves on this
    mesh (the faster of two timed runs per shape, then cached): the serving
    lanes' choice when they advance two steps per launch.  A shape must beat
    the 19 x 13 tile (shape 1, the batch path's choice) by more than 3 % to
    replace it: the probe's forward-only launches rank shapes within a few
    per cent of each other differently from the lanes' slot launches."""
    from .solver import Plan
    dev = torch.device(device if device is not None else "cuda")
    key = _key(nelem_x, nelem_z, batch, precision, dev) + "|tile2"
    shapes = TILE2_SHAPES[precision]
    hit = _MEM.get(key) or (_load().get(key) if use_cache else None)
    if hit and hit.get("tile2") in shapes:
        _MEM[key] = hit
        return int(hit["tile2"])
    L = 25.0
    p = Plan(nelem_x, nelem_z, 0.0, L * nelem_x, 0.0, L * nelem_z, 2e-4, steps, 4, 4 * L, 3000.0, 1e-8,
             ("xmin", "xmax", "zmax"), precision=precision, checkpoint_every=None, device=dev)
    rs = p.add_receivers(np.array([[L, 0.0]]))
    c2 = torch.full((batch, p.N), 2000.0 ** 2, dtype=p.dtype, device=dev)
    w = torch.zeros(steps, dtype=p.dtype, device=dev)
    src = [p.source_node(L * nelem_x / 2, L)] * batch
    st = torch.cuda.current_stream(dev)
    times = {}
    for s_ in shapes:
        p.set_tile2(s_)
        p.forward(c2, w, src, [rs] * batch)
        runs = []
        for _ in range(2):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            p.forward(c2, w, src, [rs] * batch)
            e1.record(st)
            torch.cuda.synchronize(dev)
            runs.append(e0.elapsed_time(e1) / 1e3 / steps)
        times[s_] = min(runs)
    best = min(times, key=times.get)
    if 1 in times and times[best] > times[1] / 1.03:
        best = 1
    rec = {"tile2": best, "step_us": {str(k): round(v * 1e6, 2) for k, v in times.items()},
           "tuned_at": time.strftime("%Y-%m-%dT%H:%M:%S")}
    _MEM[key] = rec
    if use_cache:
        db = _load()
        db[key] = rec
        _store(db)
    return best


def main(argv=None):
    """Tune and cache the batched configuration of a workload ahead of use:
    python -m mun_amd.autotune batch200|elbo120|cpu40 (the cache file is
    cache_path(), or MUN_AMD_AUTOTUNE_CACHE)."""
    import argparse
    from . import workloads as W
    from .batch import BatchSolver
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=("batch200", "elbo120", "cpu40"))
    ap.add_argument("--retune", action="store_true", help="time the candidates even if a choice is cached or shipped")
    a = ap.parse_args(argv)
    if a.retune:
        os.environ["MUN_AMD_AUTOTUNE"] = "retune"
    wl = {"batch200": lambda: W.batch200(nmodel=256), "elbo120": W.elbo120, "cpu40": W.cpu40}[a.workload]()
    if a.workload == "cpu40":  # the benchmark's 256 models of the 40x40 workload
        wl.coef = wl.coef[0][None] + np.random.default_rng(11).normal(0.0, 50.0, (256,) + wl.coef.shape[1:])
    s = BatchSolver.from_workload(wl, precision="fp32", device="cuda:0")
    print(json.dumps({"workload": a.workload, "cache": cache_path(),
                      **{k: s.tuning[k] for k in ("tile2", "group", "lanes", "chunk", "source")}}))


if __name__ == "__main__":
    main()
