"""On-disk formats of the reference's SEM drivers.

Files a user of the reference already has (or whose readers they have) keep
their layout:

  sem_forward_results.npz   receiver_data (nt, nrec), dt, nt, receiver_coords
                            (Bspline_Coupling/SEM/sem_forward.py:71-79)
  wavefield_snapshots.npz   snapshots (nsnap, n_roi) float32, dt, nt, xmin,
                            xmax, zmin, zmax (Bspline_Coupling/SEM/sem_waveform/core.py:1564-1581)
  vp_*.npy + vp_*.json      gridded background model[nz, nx] and its grid
                            metadata (CO2_plume/one_stage/single_cubic/initial1/
                            background/background_pre_pml.py:64-115)
  observation cache         noisy_data_<tag>.npy, clean_data_<tag>.npy,
                            obs_meta_<tag>.json with {"obs_dt", "obs_nt"} and a
                            12-hex SHA1 tag of the acquisition
                            (CO2_plume/one_stage/single_cubic/RQS_NF.py:1705-1754)

Only numpy arrays and JSON are read (``numpy.load`` without pickles).
"""
from __future__ import annotations

import hashlib
import json
import os
from pathlib import Path
from typing import Optional, Sequence, Tuple

import numpy as np


# ---------------------------------------------------------------- results
def save_forward_results(output_dir: str, results: dict, name: str = "sem_forward_results.npz") -> str:
    """np.savez of run()'s seismograms as sem_forward.py:71-79 writes them."""
    os.makedirs(output_dir, exist_ok=True)
    path = os.path.join(output_dir, name)
    np.savez(path, receiver_data=np.asarray(results["receiver_data"]), dt=results["dt"], nt=results["nt"],
             receiver_coords=np.asarray(results["receiver_coords"]))
    return path


def load_forward_results(path: str) -> dict:
    with np.load(path, allow_pickle=False) as f:
        return {"receiver_data": f["receiver_data"], "dt": float(f["dt"]), "nt": int(f["nt"]),
                "receiver_coords": f["receiver_coords"]}


def save_wavefield_snapshots(output_dir: str, snapshots: Sequence[np.ndarray], dt: float, nt: int, xmin: float,
                             xmax: float, zmin: float, zmax: float) -> Optional[str]:
    """Stack the ROI snapshots as float32 and write wavefield_snapshots.npz
    (core.py:1564-1581); nothing is written for an empty list."""
    if snapshots is None or len(snapshots) == 0:
        return None
    os.makedirs(output_dir, exist_ok=True)
    stack = np.stack([np.asarray(s, dtype=np.float32) for s in snapshots], axis=0)
    path = os.path.join(output_dir, "wavefield_snapshots.npz")
    np.savez(path, snapshots=stack, dt=dt, nt=nt, xmin=xmin, xmax=xmax, zmin=zmin, zmax=zmax)
    return path


# ---------------------------------------------------------------- gridded backgrounds
def grid_json(x: np.ndarray, z: np.ndarray, dx: float, dz: float) -> dict:
    """Grid block of the background JSON (background_pre_pml.py:64-66)."""
    return {"xmin": float(x[0]), "xmax": float(x[-1]), "zmin": float(z[0]), "zmax": float(z[-1]),
            "dx": float(dx), "dz": float(dz), "nx": int(x.size), "nz": int(z.size)}


def save_vp_grid(workdir: str, model_name: str, vp: np.ndarray, x: np.ndarray, z: np.ndarray,
                 description: str = "", extra: Optional[dict] = None) -> Tuple[str, str]:
    """<model_name>.npy (model[nz, nx]) and <model_name>.json with the grid
    metadata, as background_pre_pml.py:102-115 writes them."""
    os.makedirs(workdir, exist_ok=True)
    x = np.asarray(x, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    vp = np.asarray(vp, dtype=np.float64)
    if vp.shape != (z.size, x.size):
        raise ValueError(f"vp shape {vp.shape} must be (nz, nx) = {(z.size, x.size)}")
    npy = os.path.join(workdir, f"{model_name}.npy")
    np.save(npy, vp)
    dx = float(x[1] - x[0]) if x.size > 1 else 0.0
    dz = float(z[1] - z[0]) if z.size > 1 else 0.0
    obj = {"model_name": model_name, "description": description,
           "array_layout": "model[nz, nx], z increases downward, x increases to the right",
           "array_file": f"{model_name}.npy", "grid": grid_json(x, z, dx, dz)}
    if extra:
        obj.update(extra)
    js = os.path.join(workdir, f"{model_name}.json")
    with open(js, "w", encoding="utf-8") as f:
        json.dump(obj, f, indent=2)
    return npy, js


def load_vp_grid(path: str) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(x, z, vp[nz, nx]) from a background JSON (its ``array_file`` next to
    it) or from a .npy with a sibling .json of the same stem."""
    p = Path(path)
    js = p if p.suffix == ".json" else p.with_suffix(".json")
    with open(js, "r", encoding="utf-8") as f:
        meta = json.load(f)
    arr = p if p.suffix == ".npy" else js.parent / meta.get("array_file", js.stem + ".npy")
    vp = np.load(arr, allow_pickle=False).astype(np.float64)
    g = meta["grid"]
    nx, nz = int(g.get("nx", vp.shape[1])), int(g.get("nz", vp.shape[0]))
    if vp.shape != (nz, nx):
        raise ValueError(f"{arr}: array shape {vp.shape} does not match the grid ({nz}, {nx})")
    x = np.linspace(float(g["xmin"]), float(g["xmax"]), nx)
    z = np.linspace(float(g["zmin"]), float(g["zmax"]), nz)
    return x, z, vp


def background_from_config(bg_cfg: dict, xmin: float, xmax: float, zmin: float, zmax: float):
    """Grid axes and values from velocity.background_2d, with the key
    precedence of CO2_plume/one_stage/single_cubic/sem_waveform/core.py:353-403:
    'vp' or 'vp_path'; 'x' | 'x_path' | (xmin, xmax, nx) metadata (defaulting
    to the domain); same for z; decreasing axes are flipped with the grid."""
    if bg_cfg.get("vp") is not None:
        vp = np.asarray(bg_cfg["vp"], dtype=np.float64)
    else:
        vp_path = bg_cfg.get("vp_path")
        if vp_path is None:
            raise ValueError("background_2d must provide either 'vp' array or 'vp_path'.")
        vp = np.load(vp_path, allow_pickle=False).astype(np.float64)

    def axis(key, lo_key, hi_key, n_key, lo, hi, n_default):
        if bg_cfg.get(key) is not None:
            return np.asarray(bg_cfg[key], dtype=np.float64).reshape(-1)
        if bg_cfg.get(key + "_path") is not None:
            return np.load(bg_cfg[key + "_path"], allow_pickle=False).astype(np.float64).reshape(-1)
        return np.linspace(float(bg_cfg.get(lo_key, lo)), float(bg_cfg.get(hi_key, hi)),
                           int(bg_cfg.get(n_key, n_default)), dtype=np.float64)

    xg = axis("x", "xmin", "xmax", "nx", xmin, xmax, vp.shape[1])
    zg = axis("z", "zmin", "zmax", "nz", zmin, zmax, vp.shape[0])
    if xg.size < 2 or zg.size < 2:
        raise ValueError("background_2d grid x/z must have at least 2 points each.")
    if xg[0] > xg[-1]:
        xg, vp = xg[::-1], vp[:, ::-1]
    if zg[0] > zg[-1]:
        zg, vp = zg[::-1], vp[::-1, :]
    return np.ascontiguousarray(xg), np.ascontiguousarray(zg), np.ascontiguousarray(vp)


# ---------------------------------------------------------------- observation cache
def stable_json_hash(obj) -> str:
    """First 12 hex digits of SHA1 over canonical JSON (RQS_NF.py:1707-1709)."""
    payload = json.dumps(obj, sort_keys=True, separators=(",", ":"), default=str)
    return hashlib.sha1(payload.encode("utf-8")).hexdigest()[:12]


def observation_tag(vp_background_path: str, sources, receivers, total_time: float, dt: float, f0: float,
                    amp: float, noise_std: float, pml: float, npol: int, nelem: Sequence[int], tau: float,
                    v_inside: float, blend: str, true_ctrl) -> str:
    """Cache tag of the reference's observations (RQS_NF.py:1711-1729)."""
    return (f"semtrue_{Path(vp_background_path).stem}"
            + f"_{len(sources)}s_{len(receivers)}r"
            + f"_T{total_time:g}_dt{dt:g}"
            + f"_f{f0:g}_a{amp:g}"
            + f"_noise{noise_std:g}"
            + "_" + stable_json_hash({
                "sources": sources, "receivers": receivers, "pml": pml, "npol": npol,
                "nelem": list(nelem), "tau": tau, "v_inside": v_inside, "blend": blend,
                "true_ctrl": np.asarray(true_ctrl, dtype=float).round(8).tolist()}))


def observation_paths(cache_dir: str, tag: str) -> Tuple[str, str, str]:
    return (os.path.join(cache_dir, f"noisy_data_{tag}.npy"), os.path.join(cache_dir, f"clean_data_{tag}.npy"),
            os.path.join(cache_dir, f"obs_meta_{tag}.json"))


def save_observations(cache_dir: str, tag: str, noisy: np.ndarray, clean: np.ndarray, obs_dt: float,
                      obs_nt: int) -> None:
    os.makedirs(cache_dir, exist_ok=True)
    n, c, m = observation_paths(cache_dir, tag)
    np.save(n, np.asarray(noisy))
    np.save(c, np.asarray(clean))
    with open(m, "w") as f:
        json.dump({"obs_dt": float(obs_dt), "obs_nt": int(obs_nt)}, f, indent=2)


def load_observations(cache_dir: str, tag: str):
    """(noisy, clean, obs_dt, obs_nt) if all three files exist, else None."""
    n, c, m = observation_paths(cache_dir, tag)
    if not (os.path.exists(n) and os.path.exists(c) and os.path.exists(m)):
        return None
    with open(m, "r") as f:
        meta = json.load(f)
    return (np.load(n, allow_pickle=False).astype(np.float64), np.load(c, allow_pickle=False).astype(np.float64),
            float(meta["obs_dt"]), int(meta["obs_nt"]))
