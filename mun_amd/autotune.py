"""Self-written autotuner for the step kernel's tile shape.

The step kernel has instances for tile heights of 15, 7 and 3 element rows
(256-, 128- and 64-thread workgroups; every tile is 16 owned elements wide so
rows stay 256-byte aligned).  Tall tiles waste less on the recomputed halo
row; short tiles give more workgroups, which small meshes and small batches
need to fill 256 CUs.  ``best_tile_rows`` times the forward step of every
instance on the actual mesh and batch size (HIP events around a run of
launches on the library's stream), keeps the fastest and caches the choice in
a JSON file keyed by device, mesh, batch and precision.  The tuned height
applies to one-step launches (checkpointed forward, recompute, slot launches,
snapshots); two-step launches use a fixed 19 x 26-element tile.
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
    return os.environ.get("MUN_AMD_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "mun_amd", "autotune.json"))


def _load() -> dict:
    p = cache_path()
    try:
        with open(p) as f:
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


def _key(nelem_x, nelem_z, batch, precision, device) -> str:
    name = torch.cuda.get_device_name(device).replace(" ", "_")
    return f"{name}|{nelem_x}x{nelem_z}|b{batch}|{precision}"


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
