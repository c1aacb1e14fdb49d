"""Profiling driver: runs the solver's kernels at a fixed shape under a profiler.

    python -m mun_amd.profile_kernels [--batch 256] [--nt 40] [--copy-mb 1024] [--tile-rows 15] [--tile2 1]
    python -m mun_amd.profile_kernels --summarize-pmc FETCH_CSV WRITE_CSV --batch B [--out JSON]
    python -m mun_amd.profile_kernels --summarize-trace KERNEL_TRACE_CSV [--top 30]
    python -m mun_amd.profile_kernels --summarize-gaps KERNEL_TRACE_CSV
    python -m mun_amd.profile_kernels --summarize-counters SQ1_CSV SQ2_CSV SQ3_CSV

Order of GPU work (so counter runs can attribute dispatches):
  1. a plain device-to-device copy of --copy-mb MiB, three times
     (calibrates FETCH_SIZE / WRITE_SIZE against known bytes)
  2. a forward-only pass: --nt launches of the forward step kernel
  3. a forward+adjoint pass with --nt steps (step, adjoint, residual kernels)
on the 200 x 200 element, degree-4 batch workload.
"""
from __future__ import annotations

import argparse
import re
import sys

import torch

from . import workloads as W
from .batch import BatchSolver


def summarize_pmc(fetch_csv, write_csv, batch, nodes=641601, copy_mb=1024, tiles_per_solve=195, threads=256,
                  tiles_per_solve2=117, threads2=512):
    """Bytes per step-kernel launch from two rocprofv3 counter passes
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by what the same
    counter reads for the calibration copy (the runtime's copyBuffer kernel
    moving copy_mb MiB), against the launch's algorithmic bytes per node:
    one step per launch (step_kernel) reads 12 and writes 8; two steps per
    launch (step2_kernel) read 12 and write 8 in the forward-only pass (the
    launches of `batch` solves) and 12 where u^{n+1} is stored too (the
    full-history forward of the lane groups, the adjoint).  Solves per launch
    follow from each dispatch's grid: tiles_per_solve workgroups of `threads`
    (195 tiles of the 15-row step kernel on the 801 x 801-node mesh), or
    tiles_per_solve2 of threads2 for the two-step kernel (117 tiles of the
    default 19 x 26 shape; other shapes are read from the instance's name).
    `forward` / `adjoint` report the launches of `batch` solves (else the
    largest), preferring the two-step kernel; `kernels` all."""
    import csv
    import statistics

    def per_kernel(path, counter):
        vals = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            name = r["Kernel_Name"].split("(")[0].replace("void ", "")
            vals.setdefault((name, int(r["Grid_Size"])), []).append(float(r["Counter_Value"]))
        return vals

    copy_bytes = copy_mb << 20
    out = {"batch": batch, "nodes": nodes}
    recs = {}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        vals = per_kernel(path, counter)
        # calibration: the largest copyBuffer dispatches (the copy_mb MiB copies)
        copies = [v for (n, g), vs in vals.items() if "copyBuffer" in n for v in vs]
        cal_kib = statistics.median(sorted(copies)[-3:])
        factor = copy_bytes / (cal_kib * 1024.0)
        key = counter.split("_")[0].lower()
        out[f"{key}_factor"] = round(factor, 4)
        for (name, grid), vs in sorted(vals.items()):
            two = name.startswith("sem::step2_kernel")
            if not (two or name.startswith("sem::step_kernel")) or ", 2, " in name:
                continue
            mode = "forward" if ", 0, " in name else "adjoint"
            per2 = threads2 * tiles_per_solve2
            shape = re.search(r"step2_kernel<[^>]*, (\d+), (\d+)>", name) if two else None
            if shape:
                # the tile shape is in the instance's template arguments
                ex, ez = int(shape.group(1)), int(shape.group(2))
                nel = int(round((nodes ** 0.5 - 1) / 4))
                per2 = (ex * ez + 63) // 64 * 64 * (-(-nel // (ex - 3))) * (-(-nel // (ez - 3)))
            solves = max(1, round(grid / (per2 if two else threads * tiles_per_solve)))
            steps = 2 if two else 1
            wr = 12 if two and (mode == "adjoint" or solves != batch) else 8
            rec = recs.setdefault((mode, steps, solves), {"mode": mode, "kernel": name, "steps_per_launch": steps,
                                                          "grid": grid, "solves": solves, "dispatches": len(vs)})
            rec[f"{key}_kib_raw_median"] = statistics.median(vs)
            rec[f"{key}_bytes"] = statistics.median(vs) * 1024.0 * factor
            rec[f"{key}_algorithmic_bytes"] = (12 if key == "fetch" else wr) * nodes * solves
    for rec in recs.values():
        if "fetch_bytes" in rec and "write_bytes" in rec:
            rec["bytes_per_launch"] = int(rec["fetch_bytes"] + rec["write_bytes"])
            rec["algorithmic_bytes_per_launch"] = rec["fetch_algorithmic_bytes"] + rec["write_algorithmic_bytes"]
            rec["ratio"] = round(rec["bytes_per_launch"] / rec["algorithmic_bytes_per_launch"], 4)
    out["kernels"] = sorted(recs.values(), key=lambda r: (r["mode"], -r["steps_per_launch"], r["solves"]))
    for mode in ("forward", "adjoint"):
        mine = [r for r in out["kernels"] if r["mode"] == mode and "ratio" in r]
        if mine:
            top = max(r["steps_per_launch"] for r in mine)
            mine = [r for r in mine if r["steps_per_launch"] == top]
            pick = [r for r in mine if r["solves"] == batch] or [max(mine, key=lambda r: r["solves"])]
            out[mode] = pick[0]
    if "forward" in out:
        out["bytes_per_launch"] = out["forward"].get("bytes_per_launch")
    return out


def summarize_trace(trace_csv, top=30):
    """Markdown table of a rocprofv3 kernel trace (`--kernel-trace
    --output-format csv`): per (kernel, grid) the call count, average and
    total duration and share of the summed kernel time, with the process
    span.  Durations of kernels that overlap (lanes on separate streams) are
    summed as recorded."""
    import collections
    import csv

    grp = collections.defaultdict(lambda: [0, 0])
    t0 = t1 = None
    for r in csv.DictReader(open(trace_csv)):
        s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        t0 = s if t0 is None else min(t0, s)
        t1 = e if t1 is None else max(t1, e)
        key = (r["Kernel_Name"].split("(")[0].replace("void ", ""), int(r["Grid_Size_X"]), r["VGPR_Count"],
               r["LDS_Block_Size"])
        grp[key][0] += 1
        grp[key][1] += e - s
    total = sum(v[1] for v in grp.values()) or 1
    lines = [f"span {(t1 - t0) / 1e9:.3f} s, kernel time {total / 1e9:.3f} s" if t0 is not None else "empty trace",
             "| kernel | grid | VGPR | LDS B | calls | avg us | total ms | % of kernel time |",
             "|---|---|---|---|---|---|---|---|"]
    for (name, grid, vgpr, lds), (n, ns) in sorted(grp.items(), key=lambda kv: -kv[1][1])[:top]:
        lines.append(f"| `{name}` | {grid} | {vgpr} | {lds} | {n} | {ns / n / 1e3:.1f} | {ns / 1e6:.1f} | "
                     f"{100 * ns / total:.2f} |")
    return "\n".join(lines)


def summarize_gaps(trace_csv, min_gap_us=5.0, top=12):
    """Idle time between consecutive kernels of each stream in a rocprofv3
    kernel trace: per stream the kernels, busy time and span, and the gaps of
    at least min_gap_us attributed to the kernel that ended each gap (what
    the stream waited to start), largest total first; then the time no
    kernel of any stream ran while the streams with a two-step kernel were
    between their first and last one."""
    import collections
    import csv

    rows = list(csv.DictReader(open(trace_csv)))
    sid = "Stream_Id" if rows and "Stream_Id" in rows[0] else "Queue_Id"
    streams = collections.defaultdict(list)
    for r in rows:
        name = r["Kernel_Name"].split("(")[0].replace("void ", "")
        streams[r[sid]].append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), name))
    lines = [f"| {sid.lower()} | kernels | busy ms | span ms | idle in gaps >= {min_gap_us:g} us (ms) | "
             f"gaps | largest gaps by the kernel that ended them (count, total ms, max us) |",
             "|---|---|---|---|---|---|---|"]
    windows = []
    for key, ks in sorted(streams.items(), key=lambda kv: -len(kv[1])):
        ks.sort()
        busy = sum(e - s_ for s_, e, _ in ks)
        by = collections.defaultdict(lambda: [0, 0, 0])
        idle = n = 0
        for (s0, e0, _), (s1, e1, nm) in zip(ks, ks[1:]):
            g = s1 - e0
            if g >= min_gap_us * 1e3:
                idle += g
                n += 1
                v = by[nm]
                v[0] += 1
                v[1] += g
                v[2] = max(v[2], g)
        worst = sorted(by.items(), key=lambda kv: -kv[1][1])[:3]
        desc = "; ".join(f"`{nm}` ({c}, {t / 1e6:.2f}, {m / 1e3:.0f})" for nm, (c, t, m) in worst)
        lines.append(f"| {key} | {len(ks)} | {busy / 1e6:.1f} | {(ks[-1][1] - ks[0][0]) / 1e6:.1f} | "
                     f"{idle / 1e6:.2f} | {n} | {desc} |")
        two = [k for k in ks if "step2_kernel" in k[2]]
        if two:
            windows.append((two[0][0], max(k[1] for k in two)))
        if len(lines) >= top + 2:
            break
    if windows:
        w0, w1 = min(w[0] for w in windows), max(w[1] for w in windows)
        iv = sorted((s_, e) for ks in streams.values() for s_, e, _ in ks if e > w0 and s_ < w1)
        covered, cur0, cur1 = 0, None, None
        for s_, e in iv:
            s_, e = max(s_, w0), min(e, w1)
            if cur1 is None or s_ > cur1:
                if cur1 is not None:
                    covered += cur1 - cur0
                cur0, cur1 = s_, e
            else:
                cur1 = max(cur1, e)
        if cur1 is not None:
            covered += cur1 - cur0
        lines.append("")
        lines.append(f"two-step window {(w1 - w0) / 1e6:.1f} ms: no kernel running for "
                     f"{(w1 - w0 - covered) / 1e6:.2f} ms ({100 * (w1 - w0 - covered) / max(w1 - w0, 1):.2f} %)")
    return "\n".join(lines)


def summarize_counters(csvs, kernels=("step2_kernel", "corr_kernel")):
    """Markdown table of SQ / TA / TCC counter passes of this driver
    (rocprofv3 --pmc ... --output-format csv, one CSV per pass): per (kernel,
    grid) of the named kernels the medians over dispatches, as instructions
    per wave (SQ_INSTS_*), fractions of wave cycles (SQ_WAIT_*, SQ_ACTIVE_*),
    the texture addresser's stalled fraction and the L2 hit fraction."""
    import collections
    import csv
    import statistics

    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in csvs:
        for r in csv.DictReader(open(path)):
            name = r["Kernel_Name"].split("(")[0].replace("void ", "")
            if any(k in name for k in kernels):
                vals[(name, int(r["Grid_Size"]))][r["Counter_Name"]].append(float(r["Counter_Value"]))
    rows = [("dispatches", None, None), ("SQ_INSTS_VALU / wave", "SQ_INSTS_VALU", "SQ_WAVES"),
            ("SQ_INSTS_VMEM_RD / wave", "SQ_INSTS_VMEM_RD", "SQ_WAVES"),
            ("SQ_INSTS_VMEM_WR / wave", "SQ_INSTS_VMEM_WR", "SQ_WAVES"),
            ("SQ_INSTS_LDS / wave", "SQ_INSTS_LDS", "SQ_WAVES"),
            ("SQ_WAIT_ANY / SQ_WAVE_CYCLES", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES"),
            ("SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES", "SQ_WAIT_INST_ANY", "SQ_WAVE_CYCLES"),
            ("SQ_ACTIVE_INST_VALU / SQ_WAVE_CYCLES", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES"),
            ("TA_ADDR_STALLED_BY_TC / TA_TA_BUSY", "TA_ADDR_STALLED_BY_TC_CYCLES", "TA_TA_BUSY"),
            ("TCC hit fraction", "TCC_HIT", None), ("waves", "SQ_WAVES", 1), ("TA_TA_BUSY", "TA_TA_BUSY", 1)]
    keys = sorted(vals, key=lambda k: (k[0], -k[1]))
    lines = ["| counter | " + " | ".join(f"`{n}` grid {g}" for n, g in keys) + " |",
             "|---|" + "---|" * len(keys)]
    for label, num, den in rows:
        cells = []
        for k in keys:
            med = {c: statistics.median(v) for c, v in vals[k].items()}
            if num is None:
                cells.append(str(max(len(v) for v in vals[k].values())))
            elif num not in med:
                cells.append("-")
            elif num == "TCC_HIT":
                cells.append(f"{med['TCC_HIT'] / (med['TCC_HIT'] + med['TCC_MISS']):.2f}" if "TCC_MISS" in med else "-")
            elif den == 1:
                cells.append(f"{med[num]:.4g}")
            elif den in med:
                q = med[num] / med[den]
                cells.append(f"{q:.1f}" if den == "SQ_WAVES" else f"{q:.2f}")
            else:
                cells.append("-")
        lines.append(f"| {label} | " + " | ".join(cells) + " |")
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--nt", type=int, default=40)
    ap.add_argument("--copy-mb", type=int, default=1024)
    ap.add_argument("--checkpoint-every", type=int, default=10)
    ap.add_argument("--tile-rows", type=int, default=15, help="step-kernel tile height (15, 7 or 3)")
    ap.add_argument("--tile2", type=int, default=1, help="two-step kernel tile shape (sem_plan_set_tile2)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarize two counter passes of this driver (no GPU work)")
    ap.add_argument("--out", help="JSON file for --summarize-pmc")
    ap.add_argument("--summarize-trace", metavar="KERNEL_TRACE_CSV",
                    help="per-kernel table of a rocprofv3 kernel trace (no GPU work)")
    ap.add_argument("--top", type=int, default=30, help="rows of --summarize-trace")
    ap.add_argument("--summarize-gaps", metavar="KERNEL_TRACE_CSV",
                    help="idle gaps between the kernels of each stream of a kernel trace (no GPU work)")
    ap.add_argument("--summarize-counters", nargs="+", metavar="COUNTER_CSV",
                    help="per-kernel medians of SQ / TA / TCC counter passes of this driver (no GPU work)")
    a = ap.parse_args(argv)
    if a.summarize_counters:
        print(summarize_counters(a.summarize_counters))
        return 0
    if a.summarize_trace:
        print(summarize_trace(a.summarize_trace, a.top))
        return 0
    if a.summarize_gaps:
        print(summarize_gaps(a.summarize_gaps))
        return 0
    if a.summarize_pmc:
        import json
        res = summarize_pmc(*a.summarize_pmc, batch=a.batch, copy_mb=a.copy_mb)
        text = json.dumps(res, indent=1)
        if a.out:
            with open(a.out, "w") as f:
                f.write(text + "\n")
        print(text)
        return 0
    dev = torch.device("cuda:0")
    n = a.copy_mb * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[profile] copy {a.copy_mb} MiB x3 done", flush=True)
    wl = W.batch200(nmodel=a.batch, nt=a.nt)
    wl.mesh.checkpoint_every = a.checkpoint_every
    s = BatchSolver.from_workload(wl, tile_rows=a.tile_rows, precision="fp32", device=dev, tune=False)
    s.plan.set_tile2(a.tile2)
    coef = torch.as_tensor(wl.coef, device=dev)
    _, c2 = s.velocity(coef)
    srcn, model_of = s._expand(a.batch)
    seis = s.plan.forward(c2, s.wavelet, srcn, [s.rec_set] * len(srcn), model_of=model_of)
    torch.cuda.synchronize()
    print(f"[profile] forward: {a.nt} step launches, batch {a.batch}, N={s.plan.N}", flush=True)
    s.loglik_grad(coef, seis[0].view(1, a.nt, -1) * 0.5, 1e-9)
    torch.cuda.synchronize()
    print("[profile] forward+adjoint done", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
