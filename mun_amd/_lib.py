"""ctypes binding of libmun_sem.so (the C ABI in include/sem_capi.h).

The library is the only compute path: importing this module on a machine
without the built library, or calling into it without a GPU, raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmun_sem.so")

SEM_OK = 0
SEM_FP32, SEM_FP64 = 0, 1
SEM_ADJ_REFERENCE, SEM_ADJ_EXACT = 0, 1
SEM_RECORD_POST, SEM_RECORD_PRE = 0, 1
SIDE_BITS = {"xmin": 1, "xmax": 2, "zmin": 4, "zmax": 8}
ERROR_NAMES = {-1: "SEM_E_INVALID", -2: "SEM_E_UNSUPPORTED", -3: "SEM_E_NOMEM", -4: "SEM_E_DEVICE",
               -5: "SEM_E_NONFINITE"}

EXPORTS = ["sem_plan_create", "sem_plan_destroy", "sem_plan_dims", "sem_plan_axes", "sem_receiver_set_add",
           "sem_receiver_set_size", "sem_source_node", "sem_workspace_bytes", "sem_forward",
           "sem_forward_adjoint", "sem_bspline_map", "sem_bspline_project", "sem_curve_velocity",
           "sem_curve_project", "sem_plan_layout", "sem_slots_step", "sem_slots_record_final", "sem_pad_nodal",
           "sem_receivers_reserve", "sem_receivers_sync", "sem_residual", "sem_gradient_finalize",
           "sem_forward_snapshots", "sem_grid_to_nodes", "sem_curve_velocity_ex", "sem_curve_project_ex",
           "sem_slots_correlate", "sem_plan_gradient_elems", "sem_plan_set_schedule", "sem_plan_get_schedule",
           "sem_plan_schedule_auto", "sem_plan_set_streaming", "sem_plan_set_step_fusion",
           "sem_plan_get_step_fusion", "sem_slots_step2", "sem_bspline_project_noalloc", "sem_plan_set_tile2",
           "sem_plan_get_tile2", "sem_plan_set_corr_chunk", "sem_plan_get_corr_chunk", "sem_loop_create",
           "sem_loop_destroy", "sem_loop_chunk", "sem_loop_push", "sem_loop_round", "sem_loop_poll",
           "sem_loop_counts", "sem_stream_create", "sem_stream_destroy", "sem_last_error"]

SEM_INTERP_LINEAR, SEM_INTERP_NEAREST = 0, 1
SEM_STREAM_AUTO, SEM_STREAM_PLAIN, SEM_STREAM_NONTEMPORAL = 0, 1, 2
SEM_CURVE_CLAMPED6, SEM_CURVE_PERIODIC = 0, 1
SEM_BLEND_SMOOTH, SEM_BLEND_REPLACE = 0, 1
SEM_CURVE_MAX_CTRL = 32


class SemDesc(C.Structure):
    _fields_ = [("nelem_x", C.c_int), ("nelem_z", C.c_int), ("order", C.c_int),
                ("xmin", C.c_double), ("xmax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double),
                ("dt", C.c_double), ("nt", C.c_int),
                ("sponge_width", C.c_double), ("sponge_cmax", C.c_double), ("sponge_R", C.c_double),
                ("sponge_sides", C.c_uint), ("precision", C.c_int), ("adjoint_mode", C.c_int),
                ("record_mode", C.c_int), ("checkpoint_every", C.c_int), ("lagrange_eps", C.c_double),
                ("tile_rows", C.c_int), ("win_begin", C.c_int), ("win_end", C.c_int), ("win_decim", C.c_int)]


class SemCurveDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("nctrl", C.c_int), ("samples", C.c_int), ("newton_steps", C.c_int),
                ("v_inside", C.c_double), ("v_outside", C.c_double), ("tau", C.c_double),
                ("background", C.c_void_p), ("blend", C.c_int)]


class SemBatch(C.Structure):
    _fields_ = [("batch", C.c_int), ("nmodel", C.c_int), ("c2", C.c_void_p), ("model_of", C.c_void_p), ("wavelet", C.c_void_p),
                ("wavelet_stride", C.c_long), ("src_node", C.c_void_p), ("rec_set", C.POINTER(C.c_int)),
                ("rec_pitch", C.c_int), ("seis", C.c_void_p)]


SLOT_IDLE, SLOT_FORWARD, SLOT_RECOMPUTE, SLOT_ADJOINT, SLOT_FORWARD_HIST = 0, 1, 2, 3, 4


class SemSlot(C.Structure):
    _fields_ = [("phase", C.c_int), ("n0", C.c_int), ("len", C.c_int), ("src_node", C.c_int),
                ("rec_set", C.c_int), ("record", C.c_int), ("u0", C.c_void_p), ("u1", C.c_void_p),
                ("w", C.c_void_p), ("c2", C.c_void_p), ("wavelet", C.c_void_p), ("seis", C.c_void_p),
                ("ck_u", C.c_void_p), ("ck_w", C.c_void_p), ("seg", C.c_void_p), ("seg_stride", C.c_long),
                ("resid", C.c_void_p), ("G", C.c_void_p), ("ring", C.c_void_p), ("ring_slots", C.c_int),
                ("pair_parity", C.c_int), ("w2", C.c_void_p)]


class SemRequest(C.Structure):
    _fields_ = [("id", C.c_longlong), ("nt", C.c_int), ("full_history", C.c_int), ("src_node", C.c_int),
                ("rec_set", C.c_int), ("nrec", C.c_int), ("reserved", C.c_int), ("inv_sigma2", C.c_double)] + \
               [(n, C.c_void_p) for n in ("c2", "wavelet", "seis", "resid", "obs", "misfit", "G", "grad_c", "w",
                                          "w2", "om", "om2", "ring", "hist", "u0", "u1", "cku", "ckw", "seg")]


class SemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """Load libmun_sem.so (fails loudly when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m mun_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i, ip, d, dp = C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_double, C.POINTER(C.c_double)
    sz = C.c_size_t
    L.sem_last_error.restype = C.c_char_p
    L.sem_plan_create.argtypes = [C.POINTER(SemDesc), C.POINTER(vp)]
    L.sem_plan_destroy.argtypes = [vp]
    L.sem_plan_dims.argtypes = [vp, ip, ip, C.POINTER(C.c_long)]
    L.sem_plan_axes.argtypes = [vp, dp, dp]
    L.sem_receiver_set_add.argtypes = [vp, i, dp, ip]
    L.sem_receiver_set_size.argtypes = [vp, i, ip]
    L.sem_source_node.argtypes = [vp, d, d, ip]
    L.sem_workspace_bytes.argtypes = [vp, i, i, i, C.POINTER(sz)]
    L.sem_plan_set_schedule.argtypes = [vp, i, i, i]
    L.sem_plan_get_schedule.argtypes = [vp, ip, ip, ip]
    L.sem_plan_schedule_auto.argtypes = [vp, i, i, sz]
    L.sem_plan_set_streaming.argtypes = [vp, i]
    L.sem_plan_set_step_fusion.argtypes = [vp, i]
    L.sem_plan_get_step_fusion.argtypes = [vp, ip]
    L.sem_forward.argtypes = [vp, C.POINTER(SemBatch), vp, sz, vp]
    L.sem_forward_adjoint.argtypes = [vp, C.POINTER(SemBatch), vp, d, vp, vp, vp, vp, sz, vp]
    L.sem_bspline_map.argtypes = [vp, i, i, i, vp, vp, vp, vp, vp]
    L.sem_bspline_project.argtypes = [vp, i, i, i, vp, vp, vp]
    L.sem_bspline_project_noalloc.argtypes = [vp, i, i, i, vp, vp, vp]
    L.sem_plan_set_tile2.argtypes = [vp, i]
    L.sem_plan_get_tile2.argtypes = [vp, ip, ip, ip]
    L.sem_plan_set_corr_chunk.argtypes = [vp, i]
    L.sem_plan_get_corr_chunk.argtypes = [vp, ip]
    L.sem_curve_velocity.argtypes = [vp, i, vp, d, d, d, i, i, vp, vp, vp, vp]
    L.sem_curve_project.argtypes = [vp, i, vp, vp, d, d, d, vp, vp]
    L.sem_plan_layout.argtypes = [vp, ip, C.POINTER(C.c_long)]
    L.sem_plan_gradient_elems.argtypes = [vp, C.POINTER(C.c_long)]
    L.sem_slots_step.argtypes = [vp, vp, i, i, C.c_uint, i, vp]
    L.sem_slots_step2.argtypes = [vp, vp, i, i, C.c_uint, i, vp]
    L.sem_slots_record_final.argtypes = [vp, vp, i, C.c_uint, i, vp]
    L.sem_slots_correlate.argtypes = [vp, vp, i, i, i, vp]
    L.sem_pad_nodal.argtypes = [vp, vp, vp, i, vp]
    L.sem_receivers_reserve.argtypes = [vp, C.c_long, i]
    L.sem_receivers_sync.argtypes = [vp]
    L.sem_residual.argtypes = [vp, vp, vp, vp, vp, i, i, i, vp]
    L.sem_gradient_finalize.argtypes = [vp, vp, vp, d, vp, vp, vp]
    L.sem_forward_snapshots.argtypes = [vp, C.POINTER(SemBatch), i, i, i, i, i, vp, vp, sz, vp]
    L.sem_grid_to_nodes.argtypes = [vp, i, i, vp, vp, vp, i, vp, vp]
    L.sem_curve_velocity_ex.argtypes = [vp, C.POINTER(SemCurveDesc), i, vp, vp, vp, vp, vp]
    L.sem_curve_project_ex.argtypes = [vp, C.POINTER(SemCurveDesc), i, vp, vp, vp, vp]
    ll, llp = C.c_longlong, C.POINTER(C.c_longlong)
    L.sem_loop_create.argtypes = [vp, i, i, i, i, i, C.POINTER(vp)]
    L.sem_loop_destroy.argtypes = [vp]
    L.sem_loop_chunk.argtypes = [vp, ip]
    L.sem_loop_push.argtypes = [vp, C.POINTER(SemRequest), vp]
    L.sem_loop_round.argtypes = [vp, vp]
    L.sem_loop_poll.argtypes = [vp, ll, llp, llp, i, llp]
    L.sem_loop_counts.argtypes = [vp, llp, llp, llp, ip]
    L.sem_stream_create.argtypes = [C.POINTER(vp)]
    L.sem_stream_destroy.argtypes = [vp]
    for name in EXPORTS:
        if name != "sem_last_error":
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc: int):
    if rc != SEM_OK:
        raise SemError(rc, lib().sem_last_error().decode())
