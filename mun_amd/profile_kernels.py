"""Profiling driver: runs the solver's kernels at a fixed shape under a profiler.

    python -m mun_amd.profile_kernels [--batch 256] [--nt 40] [--copy-mb 1024] [--tile-rows 15]
    python -m mun_amd.profile_kernels --stream-ceiling

Order of GPU work (so counter runs can attribute dispatches):
  1. a plain device-to-device copy of --copy-mb MiB, three times
     (calibrates FETCH_SIZE / WRITE_SIZE against known bytes)
  2. a forward-only pass: --nt launches of the forward step kernel
  3. a forward+adjoint pass with --nt steps (step, adjoint, residual kernels)
on the 200 x 200 element, degree-4 batch workload.
"""
from __future__ import annotations

import argparse
import sys

import torch

from . import workloads as W
from .batch import BatchSolver


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--nt", type=int, default=40)
    ap.add_argument("--copy-mb", type=int, default=1024)
    ap.add_argument("--checkpoint-every", type=int, default=10)
    ap.add_argument("--tile-rows", type=int, default=15, help="step-kernel tile height (15, 7 or 3)")
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="run the HBM streaming-ceiling probe (mun_amd/stream_ceiling) instead")
    a = ap.parse_args(argv)
    if a.stream_ceiling:
        import os
        import subprocess
        exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stream_ceiling")
        if not os.path.exists(exe):
            raise FileNotFoundError(f"{exe} is missing: build it with `python -m mun_amd.build`")
        return subprocess.run([exe], check=True).returncode
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
    s = BatchSolver.from_workload(wl, tile_rows=a.tile_rows, precision="fp32", device=dev)
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
