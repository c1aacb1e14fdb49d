"""Python host layer over the C ABI: plans, batches and device buffers.

All bulk data lives in torch tensors on the GPU; this module only hands their
pointers to libmun_sem.so.  There is no CPU compute path: every method raises
if the library is missing or the tensors are not on a GPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib as L

PRECISIONS = {"fp32": (L.SEM_FP32, torch.float32), "fp64": (L.SEM_FP64, torch.float64)}
ADJ_MODES = {"reference": L.SEM_ADJ_REFERENCE, "exact": L.SEM_ADJ_EXACT}
RECORD = {"post": L.SEM_RECORD_POST, "pre": L.SEM_RECORD_PRE}
CURVE_KINDS = {"clamped6": L.SEM_CURVE_CLAMPED6, "periodic": L.SEM_CURVE_PERIODIC}
BLENDS = {"smooth": L.SEM_BLEND_SMOOTH, "replace": L.SEM_BLEND_REPLACE}


def ricker(t, f0, t0=0.1):
    """(1 - 2 a) exp(-a), a = (pi f0 (t - t0))^2
    (Bspline_Coupling/SEM/sem_waveform/sources.py:3-5)."""
    a = (np.pi * f0 * (t - t0)) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def source_wavelet(nt: int, dt: float, f0: float, t0: float = 0.1, gate_time: float = 0.3,
                   amp: float = 1.0) -> np.ndarray:
    """Ricker * amp with the DC term trapz(w)/(t_end - t_0) removed, then zero
    for t >= gate_time (sources.py:7-21)."""
    t = np.arange(nt) * dt
    w = ricker(t, f0, t0) * amp
    if nt > 1:
        trap = float(np.sum(0.5 * (w[1:] + w[:-1]) * np.diff(t)))
        w = w - trap / (t[-1] - t[0])
    w[t >= gate_time] = 0.0
    return w


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream_ptr(device) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class Plan:
    """One mesh + time axis + sponge + precision; reusable across batches.

    Mirrors the set-up half of SEMSimulation.__init__
    (Bspline_Coupling/SEM/sem_waveform/core.py:151-325)."""

    def __init__(self, nelem_x: int, nelem_z: int, xmin: float, xmax: float, zmin: float, zmax: float,
                 dt: float, nt: int, order: int = 4, sponge_width: float = 200.0, sponge_cmax: float = 3000.0,
                 sponge_R: float = 1e-8, sponge_sides: Sequence[str] = ("xmin", "xmax", "zmin", "zmax"),
                 precision: str = "fp32", adjoint_mode: str = "reference", record: str = "post",
                 checkpoint_every: Optional[int] = 100, lagrange_eps: float = 1e-12, device=None,
                 tile_rows: int = 0, misfit_window: Optional[Sequence[int]] = None, schedule: str = "whole"):
        if not torch.cuda.is_available():
            raise RuntimeError("mun_amd.Plan needs a ROCm GPU (torch.cuda.is_available() is False)")
        if schedule not in ("whole", "auto"):
            raise ValueError("schedule must be 'whole' or 'auto'")
        self.schedule = schedule
        self.device = torch.device(device if device is not None else "cuda")
        self.lib = L.lib()
        code, self.dtype = PRECISIONS[precision]
        sides = 0
        for s in sponge_sides:
            sides |= L.SIDE_BITS[s]
        wb, we, wd = (0, 0, 1) if misfit_window is None else (int(q) for q in misfit_window)
        d = L.SemDesc(nelem_x, nelem_z, order, xmin, xmax, zmin, zmax, dt, nt, sponge_width, sponge_cmax,
                      sponge_R, sides, code, ADJ_MODES[adjoint_mode], RECORD[record],
                      0 if checkpoint_every is None else int(checkpoint_every), lagrange_eps, int(tile_rows),
                      wb, we, wd)
        self.misfit_window = None if misfit_window is None else (wb, we or int(nt), max(wd, 1))
        self.desc = d
        self.precision, self.adjoint_mode, self.record = precision, adjoint_mode, record
        self.tile_rows = int(tile_rows) or 15
        self.nt, self.dt = int(nt), float(dt)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            L.check(self.lib.sem_plan_create(C.byref(d), C.byref(h)))
        self.h = h
        NX, NZ, N = C.c_int(), C.c_int(), C.c_long()
        L.check(self.lib.sem_plan_dims(h, C.byref(NX), C.byref(NZ), C.byref(N)))
        self.NX, self.NZ, self.N = NX.value, NZ.value, N.value
        self.x1d = np.zeros(self.NX)
        self.z1d = np.zeros(self.NZ)
        L.check(self.lib.sem_plan_axes(h, self.x1d.ctypes.data_as(C.POINTER(C.c_double)),
                                       self.z1d.ctypes.data_as(C.POINTER(C.c_double))))
        self.rec_sizes = []
        self._ws = None
        self.tuned = None  # tuned batch configuration (autotune.tune_batch), used by schedule_auto

    @classmethod
    def from_mesh(cls, m, **kw):
        """Build from a ``workloads.MeshSpec``."""
        kw.setdefault("checkpoint_every", m.checkpoint_every)
        return cls(m.nelem_x, m.nelem_z, m.xmin, m.xmax, m.zmin, m.zmax, m.dt, m.nt, m.order,
                   m.sponge_width, m.sponge_cmax, 1e-8, m.sponge_sides, **kw)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.lib.sem_plan_destroy(h)
            except Exception:
                pass
            self.h = None

    # ------------------------------------------------------------ set-up
    @property
    def coords(self) -> np.ndarray:
        X, Z = np.meshgrid(self.x1d, self.z1d, indexing="xy")
        return np.stack([X.reshape(-1), Z.reshape(-1)], axis=1)

    def add_receivers(self, xz) -> int:
        xz = np.ascontiguousarray(np.asarray(xz, dtype=np.float64).reshape(-1, 2))
        sid = C.c_int()
        L.check(self.lib.sem_receiver_set_add(self.h, xz.shape[0], xz.ctypes.data_as(C.POINTER(C.c_double)),
                                              C.byref(sid)))
        self.rec_sizes.append(xz.shape[0])
        return sid.value

    def source_node(self, x: float, z: float) -> int:
        n = C.c_int()
        L.check(self.lib.sem_source_node(self.h, float(x), float(z), C.byref(n)))
        return n.value

    def wavelet(self, f0: float, amp: float = 1.0, t0: float = 0.1, gate_time: float = 0.3) -> torch.Tensor:
        w = source_wavelet(self.nt, self.dt, f0, t0, gate_time, amp)
        return torch.as_tensor(w, dtype=self.dtype, device=self.device)

    # ------------------------------------------------------------ velocity
    def bspline_map(self, coef: torch.Tensor, base: Optional[torch.Tensor] = None, want_c: bool = True):
        """c = base + Bz coef Bx^T on the GPU; returns (c float64 [M,N] or None, c2 [M,N])."""
        coef = coef.to(device=self.device, dtype=torch.float64).contiguous()
        if coef.dim() == 2:
            coef = coef[None]
        M, nbz, nbx = coef.shape
        c = torch.empty((M, self.N), dtype=torch.float64, device=self.device) if want_c else None
        c2 = torch.empty((M, self.N), dtype=self.dtype, device=self.device)
        if base is not None:
            base = base.to(device=self.device, dtype=torch.float64).contiguous()
        L.check(self.lib.sem_bspline_map(self.h, M, nbz, nbx, _ptr(coef), _ptr(base), _ptr(c), _ptr(c2),
                                         _stream_ptr(self.device)))
        return c, c2

    def grid_to_nodes(self, xg, zg, vp, method: str = "linear") -> torch.Tensor:
        """Regular-grid background vp[nzg, nxg] on increasing axes xg, zg onto
        the nodes (bilinear or nearest, coordinates clipped into the grid);
        float64 [N] in lexicographic node order."""
        dev = self.device
        xg = torch.as_tensor(np.asarray(xg, dtype=np.float64).reshape(-1), device=dev)
        zg = torch.as_tensor(np.asarray(zg, dtype=np.float64).reshape(-1), device=dev)
        vp = torch.as_tensor(np.asarray(vp, dtype=np.float64), device=dev).contiguous()
        if vp.shape != (zg.numel(), xg.numel()):
            raise ValueError(f"vp grid shape {tuple(vp.shape)} must be (len(zg), len(xg))")
        if bool((xg[1:] <= xg[:-1]).any()) or bool((zg[1:] <= zg[:-1]).any()):
            raise ValueError("grid axes must be strictly increasing")
        meth = {"linear": L.SEM_INTERP_LINEAR, "nearest": L.SEM_INTERP_NEAREST}[method]
        out = torch.empty(self.N, dtype=torch.float64, device=dev)
        L.check(self.lib.sem_grid_to_nodes(self.h, xg.numel(), zg.numel(), _ptr(xg), _ptr(zg), _ptr(vp), meth,
                                           _ptr(out), _stream_ptr(dev)))
        return out

    def _curve_desc(self, kind, nctrl, v_in, v_out, tau, samples, newton, background, blend):
        cd = L.SemCurveDesc(CURVE_KINDS[kind], int(nctrl), int(samples), int(newton), float(v_in), float(v_out),
                            float(tau), _ptr(background), BLENDS[blend])
        cd._keep = background
        return cd

    def curve_velocity(self, ctrl, kind: str = "clamped6", v_in: float = 2000.0, v_out: float = 3000.0,
                       tau: float = 10.0, samples: int = 800, newton_steps: int = 7,
                       background: Optional[torch.Tensor] = None, blend: str = "smooth"):
        """Closed-curve anomaly velocity for ctrl [M, K, 2] (absolute control
        points); returns (c float64 [M,N], c2 [M,N], aux float64 [M,4,N])."""
        ctrl = np.asarray(ctrl.detach().cpu() if isinstance(ctrl, torch.Tensor) else ctrl, dtype=np.float64)
        ctrl = ctrl.reshape(1, -1, 2) if ctrl.ndim < 3 else ctrl
        M, K = ctrl.shape[0], ctrl.shape[1]
        ctrl = torch.as_tensor(np.ascontiguousarray(ctrl), device=self.device)
        if background is not None:
            background = torch.as_tensor(background, dtype=torch.float64, device=self.device).contiguous()
        cd = self._curve_desc(kind, K, v_in, v_out, tau, samples, newton_steps, background, blend)
        c = torch.empty((M, self.N), dtype=torch.float64, device=self.device)
        c2 = torch.empty((M, self.N), dtype=self.dtype, device=self.device)
        aux = torch.empty((M, 4, self.N), dtype=torch.float64, device=self.device)
        L.check(self.lib.sem_curve_velocity_ex(self.h, C.byref(cd), M, _ptr(ctrl), _ptr(c), _ptr(c2), _ptr(aux),
                                               _stream_ptr(self.device)))
        return c, c2, aux

    def curve_project(self, dL_dv, aux, nctrl: int, kind: str = "clamped6", v_in: float = 2000.0,
                      v_out: float = 3000.0, tau: float = 10.0, background: Optional[torch.Tensor] = None,
                      blend: str = "smooth") -> torch.Tensor:
        """Analytic projection of dL/dv [M,N] onto the control points: [M, 2K]
        interleaved [x0, z0, x1, z1, ...]."""
        dL_dv = torch.as_tensor(dL_dv, dtype=torch.float64, device=self.device).reshape(-1, self.N).contiguous()
        M = dL_dv.shape[0]
        if background is not None:
            background = torch.as_tensor(background, dtype=torch.float64, device=self.device).contiguous()
        cd = self._curve_desc(kind, nctrl, v_in, v_out, tau, 2, 0, background, blend)
        g = torch.empty((M, 2 * int(nctrl)), dtype=torch.float64, device=self.device)
        L.check(self.lib.sem_curve_project_ex(self.h, C.byref(cd), M, _ptr(dL_dv), _ptr(aux.contiguous()), _ptr(g),
                                              _stream_ptr(self.device)))
        return g

    def bspline_project(self, g_c: torch.Tensor, nbz: int, nbx: int) -> torch.Tensor:
        """g_coef = Bz^T g_c Bx (chain rule of ``bspline_map``)."""
        g_c = g_c.to(device=self.device, dtype=torch.float64).contiguous()
        if g_c.dim() == 1:
            g_c = g_c[None]
        M = g_c.shape[0]
        out = torch.empty((M, nbz, nbx), dtype=torch.float64, device=self.device)
        L.check(self.lib.sem_bspline_project(self.h, M, nbz, nbx, _ptr(g_c), _ptr(out),
                                             _stream_ptr(self.device)))
        return out

    def layout(self):
        """(pitch, elements per solve) of the padded nodal arrays."""
        pitch, elems = C.c_int(), C.c_long()
        L.check(self.lib.sem_plan_layout(self.h, C.byref(pitch), C.byref(elems)))
        return pitch.value, elems.value

    # ------------------------------------------------------------ schedule
    def set_schedule(self, checkpoint_every: int, group: int = 0, lanes: int = 1):
        """How forward_adjoint runs a batch: groups of at most `group` solves
        (0: the whole batch), `lanes` groups at a time on streams of their own,
        checkpoints every `checkpoint_every` steps (0: the full forward
        history).  Results do not depend on it."""
        L.check(self.lib.sem_plan_set_schedule(self.h, int(checkpoint_every), int(group), int(lanes)))

    def set_streaming(self, mode: str = "auto"):
        """State-stream memory hints of the step kernels: 'auto' (plain when
        the solves in flight fit the last-level cache, else non-temporal),
        'plain' or 'nontemporal' (for measurement; results do not change)."""
        code = {"auto": L.SEM_STREAM_AUTO, "plain": L.SEM_STREAM_PLAIN, "nontemporal": L.SEM_STREAM_NONTEMPORAL}[mode]
        L.check(self.lib.sem_plan_set_streaming(self.h, code))

    def set_step_fusion(self, steps: int):
        """Time steps per step-kernel launch: 2 (the default; forward,
        snapshot, recompute and adjoint passes, not slot launches) or 1.
        Results do not depend on it."""
        L.check(self.lib.sem_plan_set_step_fusion(self.h, int(steps)))

    def get_step_fusion(self) -> int:
        v = C.c_int()
        L.check(self.lib.sem_plan_get_step_fusion(self.h, C.byref(v)))
        return v.value

    def set_tile2(self, shape: int):
        """Tile shape of the two-step kernel (0: 19 x 26 element threads,
        1: 19 x 13, 2: 35 x 14, 3: 35 x 28 (FP32 only), 4: 23 x 11).  Results
        do not depend on it; set it before receivers are reserved for
        captured rounds."""
        L.check(self.lib.sem_plan_set_tile2(self.h, int(shape)))

    def get_tile2(self):
        """(shape, element threads along x, along z) of the two-step tile."""
        sh, ex, ez = C.c_int(), C.c_int(), C.c_int()
        L.check(self.lib.sem_plan_get_tile2(self.h, C.byref(sh), C.byref(ex), C.byref(ez)))
        return sh.value, ex.value, ez.value

    def set_corr_chunk(self, steps: int):
        """Adjoint steps per gradient-correlation chunk, at most (default 64).
        Results do not depend on it; the workspace does."""
        L.check(self.lib.sem_plan_set_corr_chunk(self.h, int(steps)))

    def get_corr_chunk(self) -> int:
        v = C.c_int()
        L.check(self.lib.sem_plan_get_corr_chunk(self.h, C.byref(v)))
        return v.value

    def get_schedule(self):
        ck, g, ln = C.c_int(), C.c_int(), C.c_int()
        L.check(self.lib.sem_plan_get_schedule(self.h, C.byref(ck), C.byref(g), C.byref(ln)))
        return {"checkpoint_every": ck.value, "group": g.value, "lanes": ln.value}

    def release_workspace(self) -> None:
        """Drop the cached device workspace (the next solve allocates it again)."""
        self._ws = None

    def memory_budget(self, reserve_frac: float = 0.01, reserve_min: int = 3 << 30) -> int:
        """Device bytes a workspace may take: free memory plus what this plan's
        workspace and torch's cache already hold, less a reserve."""
        free, total = torch.cuda.mem_get_info(self.device)
        held = self._ws.numel() if self._ws is not None else 0
        cached = torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
        return max(0, free + held + cached - max(reserve_min, int(reserve_frac * total)))

    def schedule_auto(self, batch: int, pitch: int, budget: Optional[int] = None):
        """Pick a schedule whose workspace fits `budget` bytes (default:
        memory_budget()).  With a tuned configuration (``self.tuned``, set by
        ``autotune.tune_batch``): its full-history groups and lanes when the
        batch is at least one round of them and their workspace fits, else
        the library's planner: the full forward history when it fits, else
        lanes of full-history groups, else checkpoints."""
        budget = self.memory_budget() if budget is None else int(budget)
        t = getattr(self, "tuned", None)
        if t and batch >= t["group"] * t["lanes"]:
            self.set_schedule(0, t["group"], t["lanes"])
            if self.workspace_bytes(batch, pitch, True) <= budget:
                return self.get_schedule()
        L.check(self.lib.sem_plan_schedule_auto(self.h, int(batch), int(pitch), C.c_size_t(budget)))
        return self.get_schedule()

    # ------------------------------------------------------------ solves
    def workspace(self, batch: int, pitch: int, adjoint: bool) -> torch.Tensor:
        n = C.c_size_t()
        L.check(self.lib.sem_workspace_bytes(self.h, batch, pitch, int(adjoint), C.byref(n)))
        if self._ws is None or self._ws.numel() < n.value:
            self._ws = None
            self._ws = torch.empty(max(n.value, 256), dtype=torch.uint8, device=self.device)
        return self._ws

    def workspace_bytes(self, batch: int, pitch: int, adjoint: bool) -> int:
        n = C.c_size_t()
        L.check(self.lib.sem_workspace_bytes(self.h, batch, pitch, int(adjoint), C.byref(n)))
        return n.value

    def _batch(self, c2, wavelet, src_nodes, rec_sets, model_of, seis, pitch):
        B = len(rec_sets)
        rs = (C.c_int * B)(*[int(s) for s in rec_sets])
        wstride = 0 if wavelet.dim() == 1 else wavelet.shape[-1]
        bt = L.SemBatch(B, c2.shape[0], _ptr(c2), _ptr(model_of), _ptr(wavelet), wstride, _ptr(src_nodes), rs, pitch,
                        _ptr(seis))
        bt._keep = rs
        return bt

    def _prep(self, c2, wavelet, src_nodes, rec_sets, model_of):
        dev = self.device
        c2 = c2.to(device=dev, dtype=self.dtype).contiguous()
        if c2.dim() == 1:
            c2 = c2[None]
        wavelet = wavelet.to(device=dev, dtype=self.dtype).contiguous()
        src = torch.as_tensor(np.asarray(src_nodes, dtype=np.int32).reshape(-1), device=dev)
        B = src.numel()
        rec_sets = list(rec_sets) if rec_sets is not None else [0] * B
        if len(rec_sets) == 1 and B > 1:
            rec_sets = rec_sets * B
        if len(rec_sets) != B:
            raise ValueError("rec_sets and src_nodes must have the same length")
        if model_of is not None:
            model_of = torch.as_tensor(np.asarray(model_of, dtype=np.int32).reshape(-1), device=dev)
            if model_of.numel() != B or int(model_of.max()) >= c2.shape[0] or int(model_of.min()) < 0:
                raise ValueError("model_of must hold one model index per solve")
        elif c2.shape[0] != B:
            raise ValueError(f"c2 holds {c2.shape[0]} models for {B} solves; pass model_of")
        if wavelet.dim() == 2 and wavelet.shape[0] != B:
            raise ValueError("per-solve wavelets need one row per solve")
        if wavelet.shape[-1] != self.nt:
            raise ValueError(f"wavelet must have nt={self.nt} samples")
        if not bool(torch.isfinite(c2).all()):
            raise ValueError("velocity model contains non-finite values (NaN/Inf)")
        pitch = max(self.rec_sizes[s] for s in rec_sets)
        return c2, wavelet, src, rec_sets, model_of, pitch

    def forward(self, c2, wavelet, src_nodes, rec_sets=None, model_of=None) -> torch.Tensor:
        """Batched forward solve; returns seismograms [B, nt, nrec_max]."""
        c2, wavelet, src, rec_sets, model_of, pitch = self._prep(c2, wavelet, src_nodes, rec_sets, model_of)
        B = len(rec_sets)
        seis = torch.empty((B, self.nt, pitch), dtype=self.dtype, device=self.device)
        ws = self.workspace(B, pitch, False)
        bt = self._batch(c2, wavelet, src, rec_sets, model_of, seis, pitch)
        L.check(self.lib.sem_forward(self.h, C.byref(bt), _ptr(ws), ws.numel(), _stream_ptr(self.device)))
        return seis

    def forward_snapshots(self, c2, wavelet, src_nodes, every: int, box, rec_sets=None, model_of=None):
        """Forward solve that also keeps u after steps 0, every, 2 every, ...
        on the node box (ix0, ix1, iz0, iz1); returns (seis [B, nt, nrec_max],
        snapshots [B, nsnap, iz1-iz0, ix1-ix0])."""
        c2, wavelet, src, rec_sets, model_of, pitch = self._prep(c2, wavelet, src_nodes, rec_sets, model_of)
        B = len(rec_sets)
        ix0, ix1, iz0, iz1 = (int(q) for q in box)
        every = int(every)
        if every < 1:
            raise ValueError("every must be >= 1")
        nsnap = (self.nt + every - 1) // every
        seis = torch.empty((B, self.nt, pitch), dtype=self.dtype, device=self.device)
        snaps = torch.empty((B, nsnap, max(iz1 - iz0, 0), max(ix1 - ix0, 0)), dtype=self.dtype, device=self.device)
        ws = self.workspace(B, pitch, False)
        bt = self._batch(c2, wavelet, src, rec_sets, model_of, seis, pitch)
        L.check(self.lib.sem_forward_snapshots(self.h, C.byref(bt), every, ix0, ix1, iz0, iz1, _ptr(snaps),
                                               _ptr(ws), ws.numel(), _stream_ptr(self.device)))
        return seis, snaps

    def misfit(self, seis: torch.Tensor, obs: torch.Tensor, nrec: int) -> torch.Tensor:
        """0.5 * ||seis - obs||^2 per solve over the first nrec columns, the
        difference formed in double (seis: [B, nt, pitch] in the plan's
        precision, obs: [B or 1, nt, pitch], taken in double)."""
        seis = seis.to(device=self.device, dtype=self.dtype).contiguous()
        obs = obs.to(device=self.device, dtype=torch.float64).expand(seis.shape).contiguous()
        B, nt, pitch = seis.shape
        res = torch.empty_like(seis[0])  # adjoint source, plan precision
        out = torch.empty(B, dtype=torch.float64, device=self.device)
        for b in range(B):
            L.check(self.lib.sem_residual(self.h, _ptr(seis[b]), _ptr(obs[b]), _ptr(res), _ptr(out[b:b + 1]),
                                          int(nrec), nt, pitch, _stream_ptr(self.device)))
        return out

    def forward_adjoint(self, c2, wavelet, src_nodes, obs, noise_std, rec_sets=None, model_of=None,
                        want_grad_c: bool = True, want_grad_c2: bool = False):
        """Forward + misfit + adjoint gradient for a batch.

        Returns dict(loglik [B] float64, misfit [B], grad_c [B,N] (d loglik/dc),
        grad_c2 (optional), seis [B,nt,pitch])."""
        c2, wavelet, src, rec_sets, model_of, pitch = self._prep(c2, wavelet, src_nodes, rec_sets, model_of)
        B = len(rec_sets)
        # observations stay in double: the residual is formed in double
        obs = obs.to(device=self.device, dtype=torch.float64)
        if obs.dim() == 2:
            obs = obs[None]
        if obs.shape[0] != B or obs.shape[1] != self.nt or obs.shape[2] > pitch:
            raise ValueError(f"obs must be [B={B}, nt={self.nt}, <= {pitch}]")
        if obs.shape[2] < pitch:
            obs = torch.nn.functional.pad(obs, (0, pitch - obs.shape[2]))
        obs = obs.contiguous()
        seis = torch.empty((B, self.nt, pitch), dtype=self.dtype, device=self.device)
        misfit = torch.empty(B, dtype=torch.float64, device=self.device)
        gc = torch.empty((B, self.N), dtype=torch.float64, device=self.device) if want_grad_c else None
        gc2 = torch.empty((B, self.N), dtype=torch.float64, device=self.device) if want_grad_c2 else None
        sig = np.asarray(noise_std, dtype=np.float64).reshape(-1)
        if sig.size > 1 and not np.all(sig == sig[0]):
            raise ValueError("one noise_std per batch")
        inv_s2 = 1.0 / (float(sig[0]) ** 2 + 1e-24)
        if self.schedule == "auto":
            self.schedule_auto(B, pitch)
        ws = self.workspace(B, pitch, True)
        bt = self._batch(c2, wavelet, src, rec_sets, model_of, seis, pitch)
        L.check(self.lib.sem_forward_adjoint(self.h, C.byref(bt), _ptr(obs), inv_s2, _ptr(misfit), _ptr(gc),
                                             _ptr(gc2), _ptr(ws), ws.numel(), _stream_ptr(self.device)))
        return {"misfit": misfit, "loglik": -misfit * inv_s2, "grad_c": gc, "grad_c2": gc2, "seis": seis,
                "inv_sigma2": inv_s2}
