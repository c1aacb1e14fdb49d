"""Drop-in replacement for ``sem_waveform.core.SEMSimulation``.

Same constructor (one nested config dict), same ``run()`` and
``run_forward_and_adjoint()`` methods and return dictionaries as
Bspline_Coupling/SEM/sem_waveform/core.py:88-1371, with the time loops,
velocity map, adjoint and projection executed on the GPU through
libmun_sem.so.  Differences a caller can see are opt-in config keys:

  method.precision        'fp32' (default) | 'fp64'   state precision
  method.adjoint_mode     'reference' (default, the sponge is not adjointed,
                          core.py:1252-1263) | 'exact' (discrete adjoint)
  method.absorbing_sides  sides the sponge damps (default all four,
                          boundary.py:38-48)
  method.checkpoint_every wavefield checkpoint interval (0 keeps the full
                          history); default: the full history when it fits
                          in free device memory, else checkpoints every
                          100 steps (results are the same bits either way)
  method.cfl_adjust       True (default): shrink dt like the reference
                          (utils.py:96-111); False: use time.dt as given
  time.nt                 explicit number of steps (default ceil(T/dt))
  receivers.positions     (nrec, 2) explicit receivers
                          (CO2_plume/2D_model/inversion/sem_waveform/core.py:90-172)
  velocity.bspline_grid   {'coef': (nbz, nbx), 'base': (N,) or None}
                          tensor cubic B-spline velocity nodes; run() then
                          takes node values and run_forward_and_adjoint
                          accepts 'bspline_nodes' and returns
                          'grad_wrt_nodes'.
  velocity.curve          'clamped6' | 'periodic': closed-curve
                          parameterisation (default: 'periodic' when a
                          gridded background is configured, else 'clamped6')

Configurations of the CO2 variant of the solver
(CO2_plume/one_stage/single_cubic/sem_waveform/core.py) are accepted as
well: velocity.background_2d (or .background) gridded background with
velocity.anomaly {enabled, control_points, perturbations, v_inside, blend},
velocity.vmin/vmax, and method.misfit_window {enabled, tmin, tmax |
window_len, decim}.  run() keeps ROI wavefield snapshots when
output.save_wavefield is set (core.py:318-325, 400-401, 649-651, 1564-1581).
Node-indexed attributes (``global_coords``, ``velocity_model``) use the
reference's element-blocked node numbering (mesh.py:4-43).
"""
from __future__ import annotations

import math
import time
from typing import Optional

import numpy as np
import torch

from . import io as sem_io
from .solver import Plan, source_wavelet

SIDES_ALL = ("xmin", "xmax", "zmin", "zmax")


def reference_node_order(nex: int, nez: int, p: int) -> np.ndarray:
    """perm[k] = lexicographic node id of the reference's node k: elements are
    visited ez-major, ex-minor, local j then i, and a node is numbered at its
    first visit (mesh.py:19-38)."""
    NX = nex * p + 1
    NZ = nez * p + 1
    lex = np.arange(NX * NZ).reshape(NZ, NX)
    seen = np.zeros(NX * NZ, dtype=bool)
    out = []
    for ez in range(nez):
        # element row: rows p*ez .. p*ez+p; only the first element row owns its bottom line
        block = lex[p * ez:p * ez + p + 1]
        for ex in range(nex):
            sub = block[:, p * ex:p * ex + p + 1].reshape(-1)
            new = sub[~seen[sub]]
            seen[new] = True
            out.append(new)
    return np.concatenate(out)


def cfl_dt(c_max: float, p: int, dx: float, dz: float) -> float:
    """0.35 (min(dx,dz)/(p+1)) / (c_max (p+1) sqrt 2)  (utils.py:78-82)."""
    return 0.35 * (min(dx, dz) / (p + 1)) / (c_max * (p + 1) * math.sqrt(2.0))


class SEMSimulation:
    """GPU-backed SEM simulation with the reference's Python call surface."""

    REQUIRED = ("domain", "time", "source", "receivers", "method", "velocity", "output")

    def __init__(self, config: dict, device=None):
        self.config = config
        self.device = torch.device(device if device is not None else "cuda")
        self._validate_config()
        self._initialize_components()

    # -------------------------------------------------------------- set-up
    def _validate_config(self):
        """core.py:130-135."""
        for key in self.REQUIRED:
            if key not in self.config:
                raise ValueError(f"Missing required config key: {key}")

    def _initialize_components(self):
        cfg = self.config
        d, t, s, r, m, v = (cfg[k] for k in ("domain", "time", "source", "receivers", "method", "velocity"))
        self.xmin, self.xmax = float(d["xmin"]), float(d["xmax"])
        self.zmin, self.zmax = float(d["zmin"]), float(d["zmax"])
        self.nelem_x, self.nelem_z = int(d["nelem_x"]), int(d["nelem_z"])
        self.npol = int(m["polynomial_order"])
        if self.npol != 4:
            raise ValueError(f"polynomial_order {self.npol}: the GPU kernels implement order 4")
        self.ngll = self.npol + 1
        self.dx_elem = (self.xmax - self.xmin) / self.nelem_x
        self.dz_elem = (self.zmax - self.zmin) / self.nelem_z
        self.total_time = float(t["total_time"])
        self.dt = float(t["dt"])
        self.nt = int(t["nt"]) if "nt" in t else int(np.ceil(self.total_time / self.dt))
        self.src_x, self.src_z = (float(q) for q in s["position"])
        self.src_f0 = float(s["frequency"])
        self.src_amp = float(s.get("amplitude", 1.0))
        self.pml_thickness = float(m["pml_thickness"])
        self.precision = m.get("precision", "fp32")
        self.adjoint_mode = m.get("adjoint_mode", "reference")
        self.sides = tuple(m.get("absorbing_sides", SIDES_ALL))
        ck = m.get("checkpoint_every")
        self.checkpoint_every = 100 if ck is None else int(ck)
        self.schedule = "auto" if ck is None else "whole"
        self.cfl_adjust = bool(m.get("cfl_adjust", True))
        self.adj_history_dtype = m.get("adj_history_dtype", "float64")
        # receivers: explicit positions or the circular array around the source
        if "positions" in r and r["positions"] is not None:
            self.receiver_coords_sem = [tuple(map(float, q)) for q in np.asarray(r["positions"]).reshape(-1, 2)]
        else:
            n = int(r["num_receivers"])
            rad = float(r["radius"])
            ang = np.linspace(0, 2 * np.pi, n, endpoint=False)
            self.receiver_coords_sem = [(self.src_x + rad * np.cos(a), self.src_z + rad * np.sin(a)) for a in ang]
        self.num_receivers = len(self.receiver_coords_sem)
        # velocity parameterisation
        self.grid = v.get("bspline_grid")
        bg_cfg = v.get("background_2d", v.get("background"))
        self.background_cfg = bg_cfg if isinstance(bg_cfg, dict) else None
        self.v_background = None      # reference node order (numpy)
        self._bg_lex = None           # lexicographic (device)
        anom = v.get("anomaly", {}) if isinstance(v.get("anomaly", {}), dict) else {}
        if self.grid is None and self.background_cfg is not None:
            # CO2 variant (C1S core.py:276-293, 353-447)
            self.vmin = float(v.get("vmin", 2000.0))
            self.vmax = float(v.get("vmax", 3000.0))
            self.anomaly_enabled = bool(anom.get("enabled", True))
            cp = anom.get("control_points", v.get("control_points"))
            self.ctrl_pts_original = None if cp is None else np.asarray(cp, dtype=np.float64).reshape(-1, 2)
            pt = anom.get("perturbations", v.get("perturbations"))
            a = None if pt is None else np.asarray(pt, dtype=np.float64)
            self.true_offset = None if (a is None or a.size == 0) else a.reshape(-1, 2)
            self.v_inside = float(anom.get("v_inside", self.vmin))
            self.blend = str(anom.get("blend", "smooth"))
            self.curve_kind = str(v.get("curve", "periodic"))
            self.tau = float(v.get("tau", 10.0))
            self.spline_samples = int(v.get("spline_samples", 800))
        elif self.grid is None:
            self.vmin = float(v["inside_velocity"])
            self.vmax = float(v["outside_velocity"])
            self.anomaly_enabled = True
            cp = v.get("control_points")
            self.ctrl_pts_original = None if cp is None else np.asarray(cp, dtype=np.float64).reshape(-1, 2)
            pt = v.get("perturbations")
            self.true_offset = None if pt is None else np.asarray(pt, dtype=np.float64).reshape(-1, 2)
            self.v_inside = self.vmin
            self.blend = "smooth"
            self.curve_kind = str(v.get("curve", "clamped6"))
            self.tau = float(v.get("tau", 10.0))
            self.spline_samples = int(v.get("spline_samples", 800))
        else:
            self.grid_coef = np.asarray(self.grid["coef"], dtype=np.float64)
            self.grid_base_ref = self.grid.get("base")  # node values in the reference numbering
        # output options (core.py:217-222, 318-325)
        o = cfg["output"]
        self.output_dir = o.get("output_dir", "sem_output")
        self.save_wavefield = bool(o.get("save_wavefield", False))
        self.save_seismograms = bool(o.get("save_seismograms", True))
        self.roi_xmin = float(o.get("roi_xmin", self.xmin + 100.0))
        self.roi_xmax = float(o.get("roi_xmax", self.xmax - 100.0))
        self.roi_zmin = float(o.get("roi_zmin", self.zmin + 100.0))
        self.roi_zmax = float(o.get("roi_zmax", self.zmax - 100.0))
        self._snapshot_interval_cfg = o.get("snapshot_interval")
        self.misfit_window_cfg = m.get("misfit_window")
        self._plans = {}
        # the plan needs c_max for the sponge; a gridded background raises it
        # to the background maximum (C1S core.py:417-418)
        self.c_max = float(self.vmax) if self.grid is None else \
            float(v.get("outside_velocity", v.get("vmax", 3000.0)))
        if self.background_cfg is not None and self.grid is None:
            xg, zg, vpg = sem_io.background_from_config(self.background_cfg, self.xmin, self.xmax, self.zmin,
                                                        self.zmax)
            self._bg_grid = (xg, zg, vpg, str(self.background_cfg.get("interp", "linear")))
            self.vmin = float(min(self.vmin, np.nanmin(vpg)))
            self.vmax = float(max(self.vmax, np.nanmax(vpg)))
            self.c_max = self.vmax
        p0 = self._plan("post")
        self.npoints = p0.N
        self._perm = reference_node_order(self.nelem_x, self.nelem_z, self.npol)
        self.global_coords = p0.coords[self._perm]
        self.src_node = int(np.nonzero(self._perm == p0.source_node(self.src_x, self.src_z))[0][0])
        self.src_pos = self.global_coords[self.src_node]
        if self.background_cfg is not None and self.grid is None:
            xg, zg, vpg, meth = self._bg_grid
            self._bg_lex = p0.grid_to_nodes(xg, zg, vpg, meth)
            self.v_background = self._bg_lex.cpu().numpy()[self._perm]
        self.grid_base = None
        if self.grid is not None and self.grid_base_ref is not None:
            self.grid_base = np.empty(self.npoints)
            self.grid_base[self._perm] = np.asarray(self.grid_base_ref, dtype=np.float64).reshape(-1)
        # initial velocity and CFL adjustment, as __init__ does (core.py:241-282)
        if self.grid is None:
            c, _, _ = self._model_velocity(self._ctrl_with_offset(None))
        else:
            c, _ = self._grid_velocity(self.grid_coef)
        self._set_velocity(c)
        self._adjust_dt(c)

    def _ctrl_with_offset(self, offsets):
        """Absolute control points: configured ones + offsets (or the configured
        perturbations when offsets is None); None without an anomaly."""
        if self.ctrl_pts_original is None or not self.anomaly_enabled:
            return None
        if offsets is None:
            offsets = self.true_offset if self.true_offset is not None else 0.0
        off = np.asarray(_np(offsets), dtype=np.float64)
        return self.ctrl_pts_original + (off.reshape(-1, 2) if off.ndim else off)

    def _window(self):
        """Step window of method.misfit_window at the current dt/nt, or None
        (C1S core.py:1440-1466)."""
        mw = self.misfit_window_cfg
        if not (mw and mw.get("enabled", False)):
            return None
        tmin = float(mw.get("tmin", 0.0) if mw.get("tmin") is not None else 0.0)
        tmax = mw.get("tmax")
        if tmax is None and mw.get("window_len") is not None:
            tmax = tmin + float(mw["window_len"])
        tmax = float(tmax) if tmax is not None else float(self.dt * (self.nt - 1))
        if tmax <= tmin:
            raise ValueError(f"misfit_window invalid: tmax({tmax}) <= tmin({tmin})")
        decim = max(int(mw.get("decim", 1)), 1)
        it0 = max(0, min(int(round(tmin / self.dt)), self.nt - 1))
        it1 = max(it0 + 1, min(int(round(tmax / self.dt)) + 1, self.nt))
        return it0, it1, decim

    def _plan(self, record: str, window=None) -> Plan:
        key = (record, self.dt, self.nt, window)
        if key not in self._plans:
            p = Plan(self.nelem_x, self.nelem_z, self.xmin, self.xmax, self.zmin, self.zmax, self.dt, self.nt,
                     self.npol, self.pml_thickness, self.c_max, 1e-8, self.sides, precision=self.precision,
                     adjoint_mode=self.adjoint_mode, record=record, checkpoint_every=self.checkpoint_every,
                     device=self.device, misfit_window=window, schedule=self.schedule)
            p.add_receivers(np.asarray(self.receiver_coords_sem, dtype=np.float64))
            self._plans[key] = p
        return self._plans[key]

    def _adjust_dt(self, c: torch.Tensor):
        """CFL shrink of dt/nt (utils.py:96-111; quirk Q1), unless disabled."""
        if not bool(torch.isfinite(c).all()):
            raise ValueError("[CFL] velocity_model contains non-finite values (NaN/Inf).")
        if not self.cfl_adjust:
            return
        dtm = cfl_dt(float(c.max()), self.npol, self.dx_elem, self.dz_elem)
        if self.dt > dtm:
            self.dt = float(dtm)
            self.nt = int(np.ceil(self.total_time / self.dt))

    def _set_velocity(self, c: torch.Tensor):
        self.velocity_model = c[0].cpu().numpy()[self._perm]

    # -------------------------------------------------------------- velocity
    def _curve_args(self):
        bg = self._bg_lex
        return dict(kind=self.curve_kind, v_in=self.v_inside, v_out=self.vmax, tau=self.tau, background=bg,
                    blend=self.blend)

    def _model_velocity(self, ctrl):
        """Velocity of an absolute control-point set (None: background or
        constant v_out); returns (c [1,N], c2 [1,N], aux or None)."""
        plan = self._plan("post")
        if ctrl is None:
            if self._bg_lex is not None:
                c = self._bg_lex[None].clone()
            else:
                c = torch.full((1, plan.N), self.vmax, dtype=torch.float64, device=self.device)
            return c, (c * c).to(plan.dtype), None
        a = self._curve_args()
        return plan.curve_velocity(np.asarray(ctrl, dtype=np.float64).reshape(1, -1, 2), a["kind"], a["v_in"],
                                   a["v_out"], a["tau"], self.spline_samples, 7, a["background"], a["blend"])

    def _curve_velocity(self, ctrl):
        return self._model_velocity(np.asarray(ctrl, dtype=np.float64).reshape(-1, 2))

    def _curve_project(self, dL_dv, aux, nctrl):
        plan = self._plan("post")
        a = self._curve_args()
        g = plan.curve_project(dL_dv, aux, nctrl, a["kind"], a["v_in"], a["v_out"], a["tau"], a["background"],
                               a["blend"])
        return g[0].cpu().numpy()

    def _grid_velocity(self, coef):
        plan = self._plan("post")
        base = None if self.grid_base is None else torch.as_tensor(self.grid_base, device=self.device)
        return plan.bspline_map(torch.as_tensor(np.asarray(coef, dtype=np.float64))[None], base)

    # -------------------------------------------------------------- solves
    def _align_observation_data(self, y_obs, obs_dt, obs_nt):
        """Linear resampling of y_obs onto this simulation's time axis, zero
        outside (core.py:109-128)."""
        y_obs = np.asarray(y_obs, dtype=np.float64)
        if y_obs.shape[0] > obs_nt:
            y_obs = y_obs[:obs_nt]
        src_t = np.arange(obs_nt, dtype=np.float64) * float(obs_dt)
        dst_t = np.arange(self.nt, dtype=np.float64) * float(self.dt)
        if y_obs.shape[0] == self.nt and np.allclose(src_t[:self.nt], dst_t, atol=1e-12, rtol=0.0):
            return y_obs
        out = np.zeros((self.nt, y_obs.shape[1]))
        for i in range(y_obs.shape[1]):
            out[:, i] = np.interp(dst_t, src_t, y_obs[:, i], left=0.0, right=0.0)
        return out

    def _wavelet(self, plan):
        w = source_wavelet(self.nt, self.dt, self.src_f0, 0.1, 0.3, self.src_amp)
        return torch.as_tensor(w, dtype=plan.dtype, device=self.device)

    def _snapshot_interval(self) -> int:
        if self._snapshot_interval_cfg is not None:
            return max(int(self._snapshot_interval_cfg), 1)
        return max(200, self.nt // 100)  # core.py:400-401

    def _roi(self, plan):
        """ROI node box and the ROI nodes in the reference's order (the mask
        of core.py:318-325 applied to its node numbering)."""
        ix = np.nonzero((plan.x1d >= self.roi_xmin) & (plan.x1d <= self.roi_xmax))[0]
        iz = np.nonzero((plan.z1d >= self.roi_zmin) & (plan.z1d <= self.roi_zmax))[0]
        if ix.size == 0 or iz.size == 0:
            return None, None
        box = (int(ix[0]), int(ix[-1]) + 1, int(iz[0]), int(iz[-1]) + 1)
        jz, jx = np.divmod(self._perm, plan.NX)
        inside = (jx >= box[0]) & (jx < box[1]) & (jz >= box[2]) & (jz < box[3])
        # position of each reference-ordered ROI node inside the box, row-major
        local = (jz[inside] - box[2]) * (box[1] - box[0]) + (jx[inside] - box[0])
        return box, local

    def run(self, ctrl_offsets_flat=None, use_torch: bool = True) -> dict:
        """Forward simulation; seismograms sampled before each update
        (core.py:334-543, 598-659).  For a closed-curve model the argument is
        the control-point offset (2K,) (default: the configured
        perturbations); for a B-spline grid it is the node values (nbz, nbx)
        (default: the configured ones).  With output.save_wavefield the ROI
        snapshots are returned and written to wavefield_snapshots.npz."""
        t0 = time.perf_counter()
        if self.grid is None:
            c, c2, _ = self._model_velocity(self._ctrl_with_offset(ctrl_offsets_flat))
        else:
            coef = self.grid_coef if ctrl_offsets_flat is None else np.asarray(_np(ctrl_offsets_flat))
            c, c2 = self._grid_velocity(coef)
        self._adjust_dt(c)
        self._set_velocity(c)
        plan = self._plan("pre")
        src = [plan.source_node(self.src_x, self.src_z)]
        snaps = None
        if self.save_wavefield:
            box, local = self._roi(plan)
            if box is not None:
                seis, s = plan.forward_snapshots(c2, self._wavelet(plan), src, self._snapshot_interval(), box, [0])
                s = s[0].reshape(s.shape[1], -1).cpu().numpy()
                snaps = [np.asarray(row[local], dtype=np.float32) for row in s]
            else:
                seis = plan.forward(c2, self._wavelet(plan), src, [0])
                snaps = []
        else:
            seis = plan.forward(c2, self._wavelet(plan), src, [0])
        data = seis[0].double().cpu().numpy()
        results = {"receiver_data": data if self.save_seismograms else None,
                   "wavefield_snapshots": snaps if self.save_wavefield else None,
                   "computation_time": time.perf_counter() - t0,
                   "velocity_model": self.velocity_model, "global_coords": self.global_coords,
                   "dt": float(self.dt), "nt": int(self.nt), "receiver_coords": self.receiver_coords_sem}
        if self.save_wavefield and snaps:
            sem_io.save_wavefield_snapshots(self.output_dir, snaps, self.dt, self.nt, self.xmin, self.xmax,
                                            self.zmin, self.zmax)
        return results

    def _ctrl_param(self, params_dict):
        ctrl = np.asarray(_np(params_dict["bspline_ctrl"]), dtype=np.float64)
        if ctrl.ndim == 1:
            if ctrl.size % 2:
                raise ValueError(f"bspline_ctrl 1D array length {ctrl.size} is not even; cannot reshape to (K,2)")
            ctrl = ctrl.reshape(-1, 2)
        elif not (ctrl.ndim == 2 and ctrl.shape[1] == 2):
            raise ValueError(f"bspline_ctrl must be shape (2*K,) or (K,2), got {ctrl.shape}")
        if self.curve_kind == "clamped6" and ctrl.shape[0] != 6:
            raise ValueError(f"the clamped closed curve takes 6 control points, got {ctrl.shape[0]}")
        return ctrl

    def _observations(self, plan, y_obs, obs_dt, obs_nt, window):
        """Observed data on the plan's nt rows.  Without a window: linear
        resampling onto the simulation's time axis (core.py:109-128).  With a
        window: y_obs must hold the windowed, decimated samples (dt * decim,
        (it1 - it0 + decim - 1) // decim rows), the only case the windowed
        reference compares directly (C1S core.py:1497-1503); they are placed on
        rows it0, it0 + decim, ..."""
        if window is None:
            return self._align_observation_data(y_obs, obs_dt, obs_nt)
        it0, it1, decim = window
        nt_dec = (it1 - it0 + decim - 1) // decim
        y = np.asarray(y_obs, dtype=np.float64)
        if not (abs(self.dt * decim - obs_dt) < 1e-12 and nt_dec == obs_nt and y.shape[0] == nt_dec):
            raise ValueError(f"misfit window: y_obs must have {nt_dec} rows at dt={self.dt * decim:g} "
                             f"(got {y.shape[0]} rows, obs_dt={obs_dt:g}, obs_nt={obs_nt})")
        out = np.zeros((self.nt, y.shape[1]))
        out[it0:it1:decim] = y
        return out

    def run_forward_and_adjoint(self, params_dict: dict) -> dict:
        """Forward + adjoint log-likelihood and gradient (core.py:1067-1371;
        C1S core.py:1297-1705 with a misfit window).

        Closed-curve model: params 'bspline_ctrl' (2K,) or (K,2) absolute
        control points -> 'grad_wrt_ctrl' (2K,) interleaved [x0,z0,...].
        B-spline grid model: params 'bspline_nodes' (nbz, nbx) ->
        'grad_wrt_nodes' (nbz, nbx).  Both return 'loglik', 'clean_data'
        (nt or window rows, nrec), 'dt', 'nt'."""
        y_obs = np.asarray(_np(params_dict["y_obs"]), dtype=np.float64)
        obs_dt = float(params_dict["obs_dt"])
        obs_nt = int(params_dict["obs_nt"])
        sigma = float(params_dict.get("noise_std", 0.0))
        aux = None
        if self.grid is None:
            ctrl = self._ctrl_param(params_dict)
            c, c2, aux = self._model_velocity(ctrl)
        else:
            c, c2 = self._grid_velocity(_np(params_dict.get("bspline_nodes", self.grid_coef)))
        self._adjust_dt(c)
        self._set_velocity(c)
        window = self._window()
        plan = self._plan("post", window)
        y = self._observations(plan, y_obs, obs_dt, obs_nt, window)
        obs = torch.as_tensor(y, dtype=torch.float64, device=self.device)[None]
        out = plan.forward_adjoint(c2, self._wavelet(plan), [plan.source_node(self.src_x, self.src_z)], obs,
                                   sigma, [0])
        d_syn = out["seis"][0].double().cpu().numpy()[:, :self.num_receivers]
        if window is not None:
            d_syn = d_syn[window[0]:window[1]]
        res = {"loglik": float(out["loglik"][0]), "clean_data": d_syn, "dt": float(self.dt), "nt": int(self.nt)}
        g_c = out["grad_c"]  # d loglik / dc
        if self.grid is None:
            # the reference hands dL/dv = -dJ_dv to the projection (core.py:1278-1287)
            res["grad_wrt_ctrl"] = self._curve_project(-g_c, aux, ctrl.shape[0])
        else:
            nbz, nbx = self.grid_coef.shape
            res["grad_wrt_nodes"] = plan.bspline_project(g_c, nbz, nbx)[0].cpu().numpy()
        self.last_grad_c = g_c[0].cpu().numpy()[self._perm]
        return res

    def run_batch(self, params) -> list:
        """Forward runs of M parameter sets in batched launches, sampled like
        run(): closed-curve models take control-point offsets [M, 2K] (added
        to the configured points), B-spline grids node values [M, nbz, nbx].
        Models whose CFL-adjusted (dt, nt) coincide share one batch (the
        reference adjusts dt per call, Q1).  Returns a list of M dicts with
        'receiver_data' (nt, nrec), 'dt', 'nt'.  Replaces the per-chain calls
        of SEMSimulation(config).run() in the MCMC likelihood
        (Bspline_Coupling/MCMC/MCMC.py:192-213, 426-433)."""
        M, c2, groups, _ = self._batch_models(params)
        out = [None] * M
        for (dt, nt), idx, plan, seis in self._forward_groups(groups, c2):
            data = seis.double().cpu().numpy()[:, :, :self.num_receivers]
            for k, m in enumerate(idx):
                out[m] = {"receiver_data": data[k], "dt": float(dt), "nt": int(nt)}
        return out

    def _forward_groups(self, groups, c2):
        """Batched pre-update forward runs per (dt, nt) group; yields
        ((dt, nt), model indices, plan, seismograms on the device)."""
        dt0, nt0 = self.dt, self.nt
        try:
            for (dt, nt), idx in groups.items():
                self.dt, self.nt = dt, nt
                plan = self._plan("pre")
                src = plan.source_node(self.src_x, self.src_z)
                seis = plan.forward(c2[idx].to(plan.dtype), self._wavelet(plan), [src] * len(idx), [0] * len(idx))
                yield (dt, nt), idx, plan, seis
        finally:
            self.dt, self.nt = dt0, nt0

    def _batch_models(self, params, ctrl_base=None, sequential_cfl: bool = False):
        """Velocities of M parameter sets in one batched map, grouped by their
        CFL-adjusted (dt, nt).  Closed-curve models: params are offsets [M, 2K]
        added to ``ctrl_base`` (default: the configured control points).

        CFL (utils.py:96-111, quirk Q1): by default each model's dt is this
        simulation's dt shrunk for that model alone, as a fresh simulation per
        model computes it.  ``sequential_cfl`` gives the dt that one cached
        simulation called model after model would use (its dt only ever
        shrinks, C1S core.py:1381-1390): the running minimum in model order,
        which then becomes this simulation's dt.
        Returns (M, c2 [M, N], {(dt, nt): [model indices]}, aux [M, 4, N] of
        the curve projection or None)."""
        params = np.asarray(_np(params), dtype=np.float64)
        M = params.shape[0]
        plan0 = self._plan("post")
        aux = None
        if self.grid is None:
            base_pts = self.ctrl_pts_original if ctrl_base is None else np.asarray(ctrl_base, dtype=np.float64)
            if base_pts is None or not self.anomaly_enabled:
                raise ValueError("batched runs need a closed-curve anomaly or a B-spline grid")
            ctrl = base_pts.reshape(1, -1, 2) + params.reshape(M, -1, 2)
            a = self._curve_args()
            c, c2, aux = plan0.curve_velocity(ctrl, a["kind"], a["v_in"], a["v_out"], a["tau"], self.spline_samples,
                                              7, a["background"], a["blend"])
        else:
            base = None if self.grid_base is None else torch.as_tensor(self.grid_base, device=self.device)
            c, c2 = plan0.bspline_map(torch.as_tensor(params.reshape(M, *self.grid_coef.shape)), base)
        if not bool(torch.isfinite(c).all()):
            raise ValueError("[CFL] velocity_model contains non-finite values (NaN/Inf).")
        groups = {}
        cmax = c.max(dim=1).values.cpu().numpy()
        run_dt, run_nt = self.dt, self.nt
        for m in range(M):
            dt, nt = (run_dt, run_nt) if sequential_cfl else (self.dt, self.nt)
            if self.cfl_adjust:
                dtm = cfl_dt(float(cmax[m]), self.npol, self.dx_elem, self.dz_elem)
                if dt > dtm:
                    dt, nt = float(dtm), int(np.ceil(self.total_time / dtm))
            run_dt, run_nt = dt, nt
            groups.setdefault((dt, nt), []).append(m)
        if sequential_cfl:
            self.dt, self.nt = run_dt, run_nt
        return M, c2, groups, aux

    def log_likelihood_batch(self, params, y_obs, sigma: float) -> np.ndarray:
        """sum Normal(simulated, sigma).log_prob(y_obs) for M parameter sets
        (the likelihood of Bspline_Coupling/MCMC/MCMC.py:426-433) from one
        batched forward; y_obs (nt, nrec) on the simulation's time axis.  The
        squared residual is reduced on the GPU in double."""
        M, c2, groups, _ = self._batch_models(params)
        y = np.asarray(_np(y_obs), dtype=np.float64)
        out = np.empty(M)
        const = math.log(float(sigma)) + 0.5 * math.log(2 * math.pi)
        for (dt, nt), idx, plan, seis in self._forward_groups(groups, c2):
            if y.shape != (nt, self.num_receivers):
                raise ValueError(f"y_obs shape {y.shape} != simulated {(nt, self.num_receivers)}")
            obs = torch.zeros((1,) + tuple(seis.shape[1:]), dtype=torch.float64, device=self.device)
            obs[0, :, :self.num_receivers] = torch.as_tensor(y, dtype=torch.float64)
            mis = plan.misfit(seis, obs, self.num_receivers).cpu().numpy()
            for k, m in enumerate(idx):
                out[m] = -mis[k] / float(sigma) ** 2 - nt * self.num_receivers * const
        return out

    def run_forward_only_for_loglik(self, ctrl12, sigma2, obs_dt, obs_nt, y_obs) -> float:
        """log-likelihood of a forward run with pre-update sampling, as the
        reference's finite-difference check computes it (core.py:1440-1492)."""
        if self.grid is None:
            c, c2, _ = self._model_velocity(np.asarray(_np(ctrl12), dtype=np.float64).reshape(-1, 2))
        else:
            c, c2 = self._grid_velocity(_np(ctrl12))
        plan = self._plan("pre")
        seis = plan.forward(c2, self._wavelet(plan), [plan.source_node(self.src_x, self.src_z)], [0])
        d = seis[0].double().cpu().numpy()
        r = d - self._align_observation_data(y_obs, obs_dt, obs_nt)
        return -0.5 * float(np.sum(r * r)) / float(sigma2)


def _p(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def _s(device):
    import ctypes
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _np(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return x
