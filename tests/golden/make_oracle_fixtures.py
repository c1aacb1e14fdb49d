"""Write FP64-oracle fixtures of the full-size workloads (CPU, minutes each).

The GPU parity tests compare the FP32 HIP path against these at the real
mesh sizes and step counts, where running the oracle inside a test would take
too long.  Each fixture is a small ``.npz`` (< 1 MB):

  single200   200x200 deg-4 mesh (801^2 nodes), 6000 steps at 0.25 ms, 15 Hz,
              200 receivers, layered model + 16x16 B-spline perturbation
              N(0, 100 m/s), reference-mode adjoint, checkpoints every 100
              steps (the single large solve).
  batch200m0  model 0 of the 256-model batch on the same mesh (perturbation
              N(0, 150 m/s)), reference-mode adjoint.
  elbo120     (sample, shot) pairs of the 64 x 8 ELBO batch: 120x120 deg-4
              mesh (481^2 nodes), 4000 steps at 0.4 ms, 10 Hz, 100 receivers.

Stored per solve: seismogram rows every ``ROW_STEP`` steps, the observation
rows at the same steps (the tests regenerate observations with the FP64
kernels and check them against these first), the log-likelihood and misfit,
d loglik / d(B-spline node values), and d loglik / dc on every 8th node along
each axis.  Observations come from an oracle forward solve of the workload's
seeded "true" model (post-update sampling, as run_forward_and_adjoint
compares them).

    python tests/golden/make_oracle_fixtures.py single200|batch200m0|elbo120 [--nt NT] [--out DIR]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from mun_amd import workloads as W  # noqa: E402

ROW_STEP = 60        # seismogram / observation rows kept
NODE_STEP = 8        # nodal gradient subsampling along each axis
ELBO_PAIRS = ((0, 0), (17, 3), (63, 7))


def _disc(wl, shot, nt):
    m = wl.mesh
    prob = O.Problem(m.xmin, m.xmax, m.zmin, m.zmax, m.nelem_x, m.nelem_z, 4, m.dt, nt, tuple(wl.sources[shot]),
                     wl.f0, wl.receivers, sponge_width=m.sponge_width, sponge_cmax=m.sponge_cmax,
                     sponge_sides=m.sponge_sides)
    return O.discretize(prob)


def solve(wl, model, shot, nt, log):
    """Oracle forward of the true model (observations), then forward +
    reference-mode adjoint of `model` with checkpoints every 100 steps."""
    disc = _disc(wl, shot, nt)
    mesh = disc.mesh
    t0 = time.perf_counter()
    ct, _, _ = O.tensor_bspline_velocity(mesh, wl.true_coef, wl.base)
    yobs, _ = O.forward(disc, ct, record="post", history=None)
    log(f"  observations: {time.perf_counter() - t0:.0f} s")
    c, Bz, Bx = O.tensor_bspline_velocity(mesh, wl.coef[model], wl.base)
    t0 = time.perf_counter()
    ref = O.forward_adjoint(disc, c, yobs, wl.noise_std, mode="reference", checkpoint_every=100)
    log(f"  forward + adjoint: {time.perf_counter() - t0:.0f} s")
    g = ref["dloglik_dc"].reshape(mesh.NZ, mesh.NX)
    misfit = 0.5 * float(np.sum(ref["residual"] ** 2))
    return {
        "seis_rows": ref["seis"][::ROW_STEP].copy(),
        "yobs_rows": yobs[::ROW_STEP].copy(),
        "loglik": np.float64(ref["loglik"]),
        "misfit": np.float64(misfit),
        "grad_nodes": Bz.T @ g @ Bx,
        "grad_c_sub": g[::NODE_STEP, ::NODE_STEP].copy(),
        "src_node": np.int64(disc.src_node),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("case", choices=("single200", "batch200m0", "elbo120"))
    ap.add_argument("--nt", type=int, default=0, help="override the step count (quick checks only)")
    ap.add_argument("--out", default=HERE)
    args = ap.parse_args()

    def log(msg):
        print(msg, flush=True)

    meta = {"row_step": ROW_STEP, "node_step": NODE_STEP, "mode": "reference", "checkpoint_every": 100}
    if args.case == "single200":
        wl = W.single200()
        nt = args.nt or wl.mesh.nt
        log(f"single200: nt={nt}")
        out = solve(wl, 0, 0, nt, log)
        out.update({f"meta_{k}": np.asarray(v) for k, v in meta.items()})
        out["meta_nt"] = np.int64(nt)
        name = "oracle_single200.npz"
    elif args.case == "batch200m0":
        wl = W.batch200(nmodel=1)
        nt = args.nt or wl.mesh.nt
        log(f"batch200 model 0: nt={nt}")
        out = solve(wl, 0, 0, nt, log)
        out.update({f"meta_{k}": np.asarray(v) for k, v in meta.items()})
        out["meta_nt"] = np.int64(nt)
        name = "oracle_batch200m0.npz"
    else:
        wl = W.elbo120()
        nt = args.nt or wl.mesh.nt
        out = {f"meta_{k}": np.asarray(v) for k, v in meta.items()}
        out["meta_nt"] = np.int64(nt)
        out["pairs"] = np.asarray(ELBO_PAIRS, dtype=np.int64)
        for s, k in ELBO_PAIRS:
            log(f"elbo120 sample {s} shot {k}: nt={nt}")
            one = solve(wl, s, k, nt, log)
            out.update({f"s{s}k{k}_{key}": v for key, v in one.items()})
        name = "oracle_elbo120.npz"
    if args.nt:
        name = name.replace(".npz", f"_nt{nt}.npz")
    path = os.path.join(args.out, name)
    np.savez_compressed(path, **out)
    log(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


if __name__ == "__main__":
    main()
