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
  method.checkpoint_every wavefield checkpoint interval (default 100;
                          0 keeps the full history)
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
Node-indexed attributes (``global_coords``, ``velocity_model``) use the
reference's element-blocked node numbering (mesh.py:4-43).
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np
import torch

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
        self.checkpoint_every = int(m.get("checkpoint_every", 100))
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
        if self.grid is None:
            self.vmin = float(v["inside_velocity"])
            self.vmax = float(v["outside_velocity"])
            cp = v.get("control_points")
            self.ctrl_pts_original = None if cp is None else np.asarray(cp, dtype=np.float64).reshape(6, 2)
            pt = v.get("perturbations")
            self.true_offset = None if pt is None else np.asarray(pt, dtype=np.float64).reshape(6, 2)
            self.tau = float(v.get("tau", 10.0))
            self.spline_samples = int(v.get("spline_samples", 800))
            self.c_max = self.vmax
        else:
            self.grid_coef = np.asarray(self.grid["coef"], dtype=np.float64)
            self.grid_base_ref = self.grid.get("base")  # node values in the reference numbering
            self.c_max = float(v.get("outside_velocity", v.get("vmax", 3000.0)))
        self._plans = {}
        p0 = self._plan("post")
        self.npoints = p0.N
        self._perm = reference_node_order(self.nelem_x, self.nelem_z, self.npol)
        self.global_coords = p0.coords[self._perm]
        self.src_node = int(np.nonzero(self._perm == p0.source_node(self.src_x, self.src_z))[0][0])
        self.src_pos = self.global_coords[self.src_node]
        self.grid_base = None
        if self.grid is not None and self.grid_base_ref is not None:
            self.grid_base = np.empty(self.npoints)
            self.grid_base[self._perm] = np.asarray(self.grid_base_ref, dtype=np.float64).reshape(-1)
        # initial velocity and CFL adjustment, as __init__ does (core.py:241-282)
        if self.grid is None and self.ctrl_pts_original is not None:
            ctrl = self.ctrl_pts_original + (self.true_offset if self.true_offset is not None else 0.0)
            c, _, _ = self._curve_velocity(ctrl)
        elif self.grid is None:
            c = torch.full((1, self.npoints), self.vmax, dtype=torch.float64, device=self.device)
        else:
            c, _ = self._grid_velocity(self.grid_coef)
        self._set_velocity(c)
        self._adjust_dt(c)

    def _plan(self, record: str) -> Plan:
        key = (record, self.dt, self.nt)
        if key not in self._plans:
            p = Plan(self.nelem_x, self.nelem_z, self.xmin, self.xmax, self.zmin, self.zmax, self.dt, self.nt,
                     self.npol, self.pml_thickness, self.c_max, 1e-8, self.sides, precision=self.precision,
                     adjoint_mode=self.adjoint_mode, record=record, checkpoint_every=self.checkpoint_every,
                     device=self.device)
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
    def _curve_velocity(self, ctrl6):
        from . import _lib as L
        plan = self._plan("post")
        ctrl = torch.as_tensor(np.asarray(ctrl6, dtype=np.float64).reshape(1, 12), device=self.device)
        c = torch.empty((1, plan.N), dtype=torch.float64, device=self.device)
        c2 = torch.empty((1, plan.N), dtype=plan.dtype, device=self.device)
        aux = torch.empty((1, 4, plan.N), dtype=torch.float64, device=self.device)
        L.check(plan.lib.sem_curve_velocity(plan.h, 1, _p(ctrl), self.vmin, self.vmax, self.tau,
                                            self.spline_samples, 7, _p(c), _p(c2), _p(aux),
                                            _s(self.device)))
        return c, c2, aux

    def _curve_project(self, dL_dv, aux):
        from . import _lib as L
        plan = self._plan("post")
        g = torch.empty((1, 12), dtype=torch.float64, device=self.device)
        L.check(plan.lib.sem_curve_project(plan.h, 1, _p(dL_dv.contiguous()), _p(aux), self.vmin, self.vmax,
                                           self.tau, _p(g), _s(self.device)))
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

    def run(self, ctrl_offsets_flat=None, use_torch: bool = True) -> dict:
        """Forward simulation; seismograms sampled before each update
        (core.py:334-435, 598-659).  For a closed-curve model the argument is
        the (12,) control-point offset; for a B-spline grid it is the node
        values (nbz, nbx) (default: the configured ones)."""
        if self.grid is None:
            if ctrl_offsets_flat is None and self.true_offset is not None:
                ctrl_offsets_flat = self.true_offset
            off = np.zeros((6, 2)) if ctrl_offsets_flat is None else \
                np.asarray(_np(ctrl_offsets_flat), dtype=np.float64).reshape(6, 2)
            c, c2, _ = self._curve_velocity(self.ctrl_pts_original + off)
        else:
            coef = self.grid_coef if ctrl_offsets_flat is None else np.asarray(_np(ctrl_offsets_flat))
            c, c2 = self._grid_velocity(coef)
        self._adjust_dt(c)
        self._set_velocity(c)
        plan = self._plan("pre")
        seis = plan.forward(c2, self._wavelet(plan), [plan.source_node(self.src_x, self.src_z)], [0])
        return {"dt": float(self.dt), "nt": int(self.nt),
                "receiver_data": seis[0].double().cpu().numpy(),
                "receiver_coords": self.receiver_coords_sem}

    def run_forward_and_adjoint(self, params_dict: dict) -> dict:
        """Forward + adjoint log-likelihood and gradient (core.py:1067-1371).

        Closed-curve model: params 'bspline_ctrl' (12,) or (6,2) absolute
        control points -> 'grad_wrt_ctrl' (12,) interleaved [x0,z0,...].
        B-spline grid model: params 'bspline_nodes' (nbz, nbx) ->
        'grad_wrt_nodes' (nbz, nbx).  Both return 'loglik', 'clean_data'
        (nt, nrec), 'dt', 'nt'."""
        y_obs = np.asarray(params_dict["y_obs"], dtype=np.float64)
        obs_dt = float(params_dict["obs_dt"])
        obs_nt = int(params_dict["obs_nt"])
        sigma = float(params_dict.get("noise_std", 0.0))
        aux = None
        if self.grid is None:
            ctrl = np.asarray(_np(params_dict["bspline_ctrl"]), dtype=np.float64).reshape(6, 2)
            c, c2, aux = self._curve_velocity(ctrl)
        else:
            c, c2 = self._grid_velocity(_np(params_dict.get("bspline_nodes", self.grid_coef)))
        self._adjust_dt(c)
        self._set_velocity(c)
        plan = self._plan("post")
        y = self._align_observation_data(y_obs, obs_dt, obs_nt)
        obs = torch.as_tensor(y, dtype=plan.dtype, device=self.device)[None]
        out = plan.forward_adjoint(c2, self._wavelet(plan), [plan.source_node(self.src_x, self.src_z)], obs,
                                   sigma, [0])
        res = {"loglik": float(out["loglik"][0]), "clean_data": out["seis"][0].double().cpu().numpy(),
               "dt": float(self.dt), "nt": int(self.nt)}
        g_c = out["grad_c"]  # d loglik / dc
        if self.grid is None:
            # the reference hands dL/dv = -dJ_dv to the projection (core.py:1278-1287)
            res["grad_wrt_ctrl"] = self._curve_project(-g_c, aux)
        else:
            nbz, nbx = self.grid_coef.shape
            res["grad_wrt_nodes"] = plan.bspline_project(g_c, nbz, nbx)[0].cpu().numpy()
        self.last_grad_c = g_c[0].cpu().numpy()[self._perm]
        return res

    def run_forward_only_for_loglik(self, ctrl12, sigma2, obs_dt, obs_nt, y_obs) -> float:
        """log-likelihood of a forward run with pre-update sampling, as the
        reference's finite-difference check computes it (core.py:1440-1492)."""
        if self.grid is None:
            c, c2, _ = self._curve_velocity(np.asarray(_np(ctrl12), dtype=np.float64).reshape(6, 2))
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
