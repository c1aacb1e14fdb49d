"""Continuous-batching server for forward+adjoint SEM solves of uneven sizes.

Requests (mesh size class, step count, velocity B-spline nodes, source,
receivers, observations) arrive over time.  Each size class is served by a
*lane* (two where one solve already fills the GPU, lanes_for_mesh): one
plan, one HIP stream of its own and a device-resident slot table
(``sem_slot``, see include/sem_capi.h).  A solve occupies one slot from
admission to completion and walks through rounds of S steps.  When its full forward history fits the
HBM budget at admission it keeps it (no recompute):

    FORWARD_HIST seg 0 .. K-1  ->  (final sample, residual, misfit)
    -> ADJOINT seg K-1 -> ... -> ADJOINT seg 0
    -> (gradient, projection onto the request's B-spline nodes) -> done

otherwise it keeps checkpoints of (u, w) every S steps:

    FORWARD seg 0 .. K-1  ->  (final sample, residual, misfit)
    -> RECOMPUTE seg K-1 -> ADJOINT seg K-1 -> ... -> RECOMPUTE seg 0 -> ADJOINT seg 0
    -> (gradient, projection) -> done

Every round advances *all* occupied slots of a lane together, two steps per
launch by default: the launches of a step pair cover the slots of every
phase (forward, recompute and adjoint slots in one launch in the device
loop), so solves at different phases and time steps share launches.  Requests that arrive while others run are admitted
at the next round boundary of their lane as soon as a slot is free and their
HBM footprint (state + checkpoints + segment buffer + seismograms) fits the
budget; lanes run on separate streams, so different size classes overlap on
the device.  The budget is one device allocation made when the server starts
(its arena), and each admitted request takes one block of it, so serving
allocates no device memory per request; its inputs reach the device through
pinned staging, and requests with the same receiver positions share the
lane's receiver set.  A round can be replayed from a captured HIP graph (its
launch arguments only point at the lane's device slot table, which is
rewritten between rounds).  close() destroys the lanes' streams.

Two engines drive the rounds.  With ``loop="device"`` (the default) the
request queue, the slot states and the round's slot table live on the device
(``sem_loop_*`` in include/sem_capi.h): the host pushes each admitted request
into the lane's device queue, replays one captured graph per round (the
device picks queued requests into free slots, runs the segment, forms
residuals, assembles completed gradients, advances every slot's phase and
frees completed slots), and polls a completion list in host-mapped memory;
only the projection onto the request's B-spline nodes and the copies of the
results are enqueued per completed solve, on a stream of their own.  With
``loop="host"`` the host keeps the slot states, writes each round's slot
table and enqueues the transitions, residuals and finishing work itself,
with one captured graph per set of phases present.  Either way rounds are
pipelined: up to ROUNDS_AHEAD rounds per lane are enqueued ahead of the
host, so no round waits for it.

Each slot's arithmetic is the batch path's, so a result does not depend on
which other requests share its launches.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _lib as L
from .solver import Plan, source_wavelet

ROUNDS_AHEAD = 2  # rounds a lane keeps enqueued ahead of the host


def staged(x, dtype, device) -> torch.Tensor:
    """x on the device, copied from pinned host memory without blocking the
    host: a copy from pageable memory would wait for the work already queued
    on the stream (the lane's rounds in flight) before it returns."""
    host = torch.as_tensor(np.ascontiguousarray(x), dtype=dtype).pin_memory()
    return host.to(device, non_blocking=True)


def owned_stream(device):
    """A new stream of the caller's own (sem_stream_create), wrapped for
    torch, and its handle for release_stream.  Lanes take these rather than
    torch's pooled streams, which are handed out in turn from a fixed set and
    so return to later servers: of eight servers created one after another
    in one process, some served the 48-request stream 10-45 % slower on
    pooled streams, and the third and later ones 5-12 % slower on streams
    handed on from server to server; on new streams, released when the
    server closed, all eight served it at 118-123 solves/s saturated."""
    dev = torch.device(device)
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    h = C.c_void_p()
    with torch.cuda.device(idx):
        L.check(L.lib().sem_stream_create(C.byref(h)))
    return torch.cuda.ExternalStream(h.value, device=idx), h


def release_stream(handle):
    """Destroy a stream of owned_stream (after its work).  Everything that
    may still record an event on it (pinned buffers of non-blocking copies
    on it) must have been freed first."""
    if handle is not None and handle.value:
        L.check(L.lib().sem_stream_destroy(handle))


@dataclass
class SolveRequest:
    """One forward+adjoint solve on a square-element mesh."""
    rid: int
    nelem_x: int
    nelem_z: int
    xmax: float
    zmax: float
    dt: float
    nt: int
    coef: np.ndarray              # (nbz, nbx) B-spline velocity nodes
    src_xz: tuple
    f0: float
    rec_xz: np.ndarray            # (nrec, 2)
    obs: Optional[np.ndarray]     # (nt, nrec) observations, or None for forward-only
    noise_std: float = 1e-9
    sponge_width: float = 200.0
    sponge_cmax: float = 3500.0
    arrival: float = 0.0
    base: Optional[np.ndarray] = None

    @property
    def lane_key(self):
        return (self.nelem_x, self.nelem_z, self.xmax, self.zmax, self.dt, self.sponge_width, self.sponge_cmax)


@dataclass
class SolveResult:
    rid: int
    loglik: float
    misfit: float
    grad: np.ndarray              # d loglik / d coef, (nbz, nbx)
    seis: np.ndarray              # (nt, nrec) synthetic seismograms
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self):
        return self.finished - self.arrival


def corr_chunk(S: int) -> int:
    """Adjoint steps per gradient-correlation chunk in a lane: at most 50, a
    segment of S steps split into near-equal chunks, rounded up to an even
    length when shorter than the segment (two-step launches never straddle an
    even chunk boundary); the device loop's sem_loop_chunk is the same rule.
    Longer chunks mean fewer correlation launches per round and a longer
    adjoint ring (chunk + 1 states) per request: with 100-step segments, 50
    served 32 200x200 solves through one lane 4 % faster than 20 and the
    mixed-size stream as fast; 100 left the stream 8 % slower."""
    n = -(-S // 50)
    c = -(-S // n)
    return c + 1 if c % 2 and c < S else c


def request_footprint(NX, NZ, pitch_elems, nt, nrec_pitch, S, elem_bytes=4, grad_elems=None,
                      full_history=False, two_step=False):
    """HBM bytes one solve holds while resident: c2, w, either u0, u1, 2K
    checkpoint states and S+1 segment states or the nt+1 states of the full
    history, the adjoint ring and omega, the FP64 gradient accumulator
    (grad_elems doubles; default: one nodal array), traces (seismograms and
    residuals in the state precision, observations in double); with two steps
    per launch the second buffers of w and omega."""
    E = pitch_elems
    K = max(1, math.ceil(nt / S))
    GE = E if grad_elems is None else grad_elems
    states = ((nt + 1) if full_history else (2 + 2 * K + (S + 1))) + (2 if two_step else 0)
    nodal = E * elem_bytes * (1 + 1 + states + (corr_chunk(S) + 1) + 1) + GE * 8
    traces = 2 * nt * nrec_pitch * elem_bytes + nt * nrec_pitch * 8 + nt * elem_bytes
    return nodal + traces


def lanes_for_mesh(nelem_x: int, nelem_z: int, compute_units: int) -> int:
    """Lanes (streams) for a size class by default: 2 where one solve
    already fills the GPU (its two-step launch on the 19 x 13 tile, 16 x 10
    owned elements, has at least one workgroup per compute unit), so that one
    lane's correlation and round boundaries overlap the other's steps, as the
    batch path's two lanes do; 1 below that.  32 full-history 200x200 solves
    through 16 slots: 28.6 solves/s as 2 x 8 against 25.1 as one lane;
    splitting the 20-160^2 classes of the mixed stream too served it 4-10 %
    slower."""
    tiles = -(-nelem_x // 16) * -(-nelem_z // 10)
    return 2 if tiles >= compute_units else 1


ALIGN = 256  # bytes: every buffer of a request block starts on this boundary


def request_layout(pitch_elems, nt, nrec_pitch, S, chunk, grad_elems, node_elems=0, full_history=False,
                   two_step=False, elem_bytes=4):
    """The buffers of one request as laid out in its arena block: a list of
    (name, shape, dtype, zeroed, byte offset) and the block's size.  Buffers
    a launch reads before writing are zeroed at admission; the full history
    needs only u^0 cleared (done separately).  node_elems: doubles of the
    nodal gradient the device loop writes at completion (0: none)."""
    E = pitch_elems
    K = max(1, math.ceil(nt / S))
    f = torch.float32 if elem_bytes == 4 else torch.float64
    d = torch.float64
    bufs = [("w", (E,), f, True), ("om", (E,), f, True), ("ring", (chunk + 1, E), f, True), ("c2p", (E,), f, True),
            ("G", (grad_elems,), d, True), ("seis", (nt, nrec_pitch), f, True), ("res", (nt, nrec_pitch), f, True),
            ("obs", (nt, nrec_pitch), d, True), ("misfit", (1,), d, True), ("wav", (nt,), f, False)]
    if two_step:
        bufs += [("w2", (E,), f, True), ("om2", (E,), f, True)]
    if full_history:
        bufs += [("hist", (nt + 1, E), f, False)]
    else:
        bufs += [("u0", (E,), f, True), ("u1", (E,), f, True), ("cku", (K, E), f, True), ("ckw", (K, E), f, True),
                 ("seg", (S + 1, E), f, True)]
    if node_elems:
        bufs += [("gc", (node_elems,), d, False)]
    out, off = [], 0
    for name, shape, dt, zero in bufs:
        out.append((name, shape, dt, zero, off))
        nb = math.prod(shape) * (4 if dt == torch.float32 else 8)
        off += -(-nb // ALIGN) * ALIGN
    return out, off


class BlockAllocator:
    """First-fit allocation of ALIGN-multiple blocks in [0, size), freed
    blocks merged with free neighbours.  Host bookkeeping only."""

    def __init__(self, size: int):
        self.size = size // ALIGN * ALIGN
        self.free = [(0, self.size)]  # (offset, length), sorted by offset, never adjacent

    def alloc(self, nbytes: int) -> Optional[int]:
        n = -(-max(int(nbytes), 1) // ALIGN) * ALIGN
        for i, (o, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    del self.free[i]
                else:
                    self.free[i] = (o + n, ln - n)
                return o
        return None

    def release(self, off: int, nbytes: int):
        n = -(-max(int(nbytes), 1) // ALIGN) * ALIGN
        i = 0
        while i < len(self.free) and self.free[i][0] < off:
            i += 1
        if i > 0:
            po, pl = self.free[i - 1]
            if po + pl > off:
                raise ValueError("block released twice or overlapping a free range")
        if i < len(self.free) and off + n > self.free[i][0]:
            raise ValueError("block released twice or overlapping a free range")
        self.free.insert(i, (off, n))
        if i + 1 < len(self.free) and off + n == self.free[i + 1][0]:  # merge right
            self.free[i] = (off, n + self.free[i + 1][1])
            del self.free[i + 1]
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] == off:  # merge left
            self.free[i - 1] = (self.free[i - 1][0], self.free[i - 1][1] + self.free[i][1])
            del self.free[i]

    def free_bytes(self) -> int:
        return sum(ln for _, ln in self.free)


class DeviceArena:
    """The serving budget as one device allocation made when the server is
    created: requests take blocks of it (BlockAllocator), so serving itself
    allocates no device memory.  A fresh device allocation can block its
    caller for seconds while the driver prepares the pages; inside a
    request's admission that stall would land in its latency."""

    def __init__(self, nbytes: int, device):
        self.alloc_ = BlockAllocator(int(nbytes))
        self.buf = torch.empty(self.alloc_.size, dtype=torch.uint8, device=device)

    def take(self, nbytes: int) -> Optional[int]:
        return self.alloc_.alloc(nbytes)

    def give(self, off: int, nbytes: int):
        self.alloc_.release(off, nbytes)

    def view(self, off: int, shape, dtype) -> torch.Tensor:
        n = math.prod(shape) * (4 if dtype == torch.float32 else 8)
        return self.buf[off:off + n].view(dtype).view(shape)


@dataclass
class _Active:
    req: SolveRequest
    slot: int
    K: int
    phase: int
    k: int
    buf: Dict[str, torch.Tensor] = field(default_factory=dict)
    footprint: int = 0
    admitted: float = 0.0
    rec_set: int = 0
    nrec: int = 0
    src_node: int = 0
    hist: bool = False            # full forward history (no recompute)
    block: int = 0                # offset of the request's block in the server's arena
    wflip: bool = False           # two-step rounds: w^n of the forward is in w2, not w
    omflip: bool = False          # two-step rounds: omega of the adjoint is in om2, not om


class Lane:
    """The slots, plan and stream of one size class (or of one of its lanes:
    `key` is the class's lane_key, followed by the lane's index when the
    class has more than one)."""

    def __init__(self, server: "SEMServer", key, capacity: int, nt_max: int, rec_pitch: int):
        nex, nez, xmax, zmax, dt, sw, scm = key[:7]
        self.server = server
        self.key = key
        self.C = capacity
        self.S = server.S
        self.chunk = corr_chunk(self.S)
        self.rec_pitch = rec_pitch
        tile_rows = 15
        if server.autotune:
            from .autotune import best_tile_rows
            tile_rows = best_tile_rows(nex, nez, capacity, "fp32", server.device)
        self.tile_rows = tile_rows
        self.plan = Plan(nex, nez, 0.0, xmax, 0.0, zmax, dt, nt_max, 4, sw, scm, 1e-8, server.sides,
                         precision="fp32", adjoint_mode=server.adjoint_mode, record="post",
                         checkpoint_every=self.S, device=server.device, tile_rows=tile_rows)
        # two-step tile shape of the lane (before the receiver tables are
        # reserved): the server's, else the autotuner's when rounds take two
        # steps per launch, else the default
        self.tile2 = server.tile2
        if self.tile2 is None:
            self.tile2 = 1
            if server.two_step and server.autotune:
                from .autotune import best_tile2
                self.tile2 = best_tile2(nex, nez, capacity, "fp32", server.device)
        self.plan.set_tile2(self.tile2)
        pitch, elems = C.c_int(), C.c_long()
        L.check(self.plan.lib.sem_plan_layout(self.plan.h, C.byref(pitch), C.byref(elems)))
        self.E = elems.value
        gelems = C.c_long()
        L.check(self.plan.lib.sem_plan_gradient_elems(self.plan.h, C.byref(gelems)))
        self.GE = gelems.value  # doubles in one solve's gradient accumulator
        # receiver tables: one set per admitted request; each receiver is listed in <= 4 tiles
        self.max_sets = server.max_sets
        self.rec_set_ids = {}  # (count, positions as bytes) -> receiver set id
        self._reserve()
        self.stream, self._stream_h = owned_stream(server.device)
        self.table_host = (L.SemSlot * capacity)()
        # two pinned staging copies of the table: the copy for round r + 2 is
        # written only after round r (and its table upload) has completed
        self.table_pinned = [torch.empty(C.sizeof(self.table_host), dtype=torch.uint8).pin_memory()
                             for _ in range(ROUNDS_AHEAD)]
        self.table_dev = torch.zeros(C.sizeof(self.table_host), dtype=torch.uint8, device=server.device)
        self.slots: List[Optional[_Active]] = [None] * capacity
        self.queue = deque()
        self.inflight = deque()   # completion events of the rounds enqueued and not yet seen done
        self.finishing = []       # (solve, event, host copies) of solves whose results are on their way
        self.graphs = {}  # phase mask -> captured round
        self.rounds = 0

    def close(self):
        """Destroy the lane's stream after its work, once the buffers that
        record events on it (the pinned slot tables, results on their way)
        are freed; the lane is not used afterwards."""
        h, self._stream_h = getattr(self, "_stream_h", None), None
        if h is None:
            return
        self.stream.synchronize()
        self.graphs = {}
        self.table_pinned = []
        self.table_dev = None
        self.finishing = []
        self.inflight.clear()
        release_stream(h)

    def _reserve(self):
        L.check(self.plan.lib.sem_receivers_reserve(self.plan.h, 4 * self.rec_pitch * self.max_sets, self.max_sets))

    def receiver_set(self, rec_xz) -> int:
        """The lane's receiver set for these positions, added on first use:
        requests with the same receiver positions share one, so a server that
        keeps serving requests of known layouts stops adding sets (a plan
        never drops one)."""
        xz = np.ascontiguousarray(np.asarray(rec_xz, dtype=np.float64).reshape(-1, 2))
        key = (xz.shape[0], xz.tobytes())
        sid = self.rec_set_ids.get(key)
        if sid is None:
            self.ensure_set_capacity()
            sid = self.plan.add_receivers(xz)
            self.rec_set_ids[key] = sid
        return sid

    def ensure_set_capacity(self):
        """Grow the receiver table before it would move under a captured graph."""
        if len(self.plan.rec_sizes) + 1 > self.max_sets:
            self.max_sets *= 2
            self._reserve()
            self.graphs = {}  # captured arguments point at the old table

    # -------------------------------------------------------------- slots
    def free_slot(self):
        for i, a in enumerate(self.slots):
            if a is None:
                return i
        return None

    def occupied(self):
        return [a for a in self.slots if a is not None]

    def _desc(self, a: _Active) -> L.SemSlot:
        r, b, S = a.req, a.buf, self.S
        s0 = a.k * S
        ln = min(S, r.nt - s0)
        d = L.SemSlot()
        d.phase, d.n0, d.len = a.phase, s0, ln
        d.src_node, d.rec_set, d.record = a.src_node, a.rec_set, 1
        d.c2, d.wavelet = b["c2p"].data_ptr(), b["wav"].data_ptr()
        d.seis, d.resid, d.G = b["seis"].data_ptr(), b["res"].data_ptr(), b["G"].data_ptr()
        d.seg_stride = self.E
        d.seg = b["hist"][s0].data_ptr() if a.hist else b["seg"].data_ptr()
        if a.phase == L.SLOT_FORWARD_HIST:
            d.w = b["w"].data_ptr()
            if "w2" in b:
                d.w, d.w2 = (b["w2"].data_ptr(), b["w"].data_ptr()) if a.wflip else (d.w, b["w2"].data_ptr())
        elif a.phase == L.SLOT_FORWARD:
            d.u0, d.u1, d.w = b["u0"].data_ptr(), b["u1"].data_ptr(), b["w"].data_ptr()
            if "w2" in b:  # two steps per launch: u by step pair, the increment flips
                d.w, d.w2 = (b["w2"].data_ptr(), b["w"].data_ptr()) if a.wflip else (d.w, b["w2"].data_ptr())
                d.pair_parity = 1
            if a.k + 1 < a.K:
                d.ck_u, d.ck_w = b["cku"][a.k + 1].data_ptr(), b["ckw"][a.k + 1].data_ptr()
        elif a.phase == L.SLOT_RECOMPUTE:
            d.u0, d.w = b["cku"][a.k].data_ptr(), b["ckw"][a.k].data_ptr()
            if "w2" in b:  # the segment's increment is not needed after it: no flip to track
                d.w2 = b["w2"].data_ptr()
        elif a.phase == L.SLOT_ADJOINT:
            d.w = b["om"].data_ptr()
            if "om2" in b:
                d.w, d.w2 = (b["om2"].data_ptr(), b["om"].data_ptr()) if a.omflip else (d.w, b["om2"].data_ptr())
            d.ring, d.ring_slots = b["ring"].data_ptr(), self.chunk + 1
        return d

    def write_table(self):
        for i in range(self.C):
            a = self.slots[i]
            self.table_host[i] = self._desc(a) if a is not None else L.SemSlot()
        pinned = self.table_pinned[self.rounds % ROUNDS_AHEAD]
        pinned.numpy()[:] = np.frombuffer(self.table_host, dtype=np.uint8)
        self.table_dev.copy_(pinned, non_blocking=True)

    def busy(self) -> bool:
        return bool(self.occupied() or self.queue or self.inflight or self.finishing)

    def outstanding(self) -> int:
        """Admitted solves not yet completed."""
        return len(self.occupied())

    def phase_mask(self) -> int:
        """Phases present in the lane (bit 1 << SLOT_*): a round launches the
        forward kernel only when a slot runs forward or recompute steps, the
        adjoint kernel and the correlation only when a slot runs adjoint steps."""
        mask = 0
        for a in self.occupied():
            mask |= 1 << a.phase
        return mask

    def _launch_round(self, mask: int):
        lib, h = self.plan.lib, self.plan.h
        sp = C.c_void_p(self.stream.cuda_stream)
        tab = C.c_void_p(self.table_dev.data_ptr())
        adjoint = bool(mask & (1 << L.SLOT_ADJOINT))
        # two steps per launch when the server runs two-step rounds (step
        # pairs m, m + 1 never straddle a correlation chunk: S and the chunk
        # are even), else one
        two = mask & TWO_STEP_PHASES if self.server.two_step else 0
        one = mask & ~two
        for m in range(self.S):
            if one:
                L.check(lib.sem_slots_step(h, tab, self.C, m, one, self.rec_pitch, sp))
            if two and m % 2 == 0:
                L.check(lib.sem_slots_step2(h, tab, self.C, m, two, self.rec_pitch, sp))
            if adjoint and ((m + 1) % self.chunk == 0 or m == self.S - 1):
                # gradient correlation of the adjoint slots' chunk that just ended
                L.check(lib.sem_slots_correlate(h, tab, self.C, m // self.chunk, self.chunk, sp))

    def launch_round(self, use_graph: bool):
        mask = self.phase_mask()
        with torch.cuda.stream(self.stream):
            self.write_table()
            if use_graph:
                g = self.graphs.get(mask)
                if g is None:
                    L.check(self.plan.lib.sem_receivers_sync(self.plan.h))
                    torch.cuda.synchronize(self.server.device)
                    g = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(g, stream=self.stream):
                        self._launch_round(mask)
                    self.graphs[mask] = g
                g.replay()
            else:
                self._launch_round(mask)
            ev = torch.cuda.Event()
            ev.record(self.stream)
            self.inflight.append(ev)
        self.rounds += 1


TWO_STEP_PHASES = ((1 << L.SLOT_FORWARD) | (1 << L.SLOT_FORWARD_HIST) | (1 << L.SLOT_RECOMPUTE)
                   | (1 << L.SLOT_ADJOINT))


@dataclass
class _Pushed:
    req: SolveRequest
    buf: Dict[str, torch.Tensor]
    footprint: int
    admitted: float
    nrec: int
    block: int = 0


class LoopLane(Lane):
    """A size class served by the device-resident request loop (sem_loop_*):
    the host pushes admitted requests and replays the same captured round;
    the device does the per-round scheduling."""

    QUEUE_CAPACITY = 1024

    def __init__(self, server: "SEMServer", key, capacity: int, nt_max: int, rec_pitch: int):
        super().__init__(server, key, capacity, nt_max, rec_pitch)
        h = C.c_void_p()
        L.check(self.plan.lib.sem_loop_create(self.plan.h, capacity, self.S, 1 if server.two_step else 0, rec_pitch,
                                              self.QUEUE_CAPACITY, C.byref(h)))
        self.loop = h
        ck = C.c_int()
        L.check(self.plan.lib.sem_loop_chunk(h, C.byref(ck)))
        self.chunk = ck.value          # ring states per request: chunk + 1
        self.pushed: Dict[int, _Pushed] = {}
        self.npushed = 0
        self.ndone = 0                 # completions seen (sem_loop_poll)
        self.round_events = {}         # round number -> completion event of that round
        self.graph = None              # the captured round
        self._ids = (C.c_longlong * 256)()
        self._rounds = (C.c_longlong * 256)()

    def close(self):
        self.destroy_loop()
        self.graph = None
        self.round_events = {}
        super().close()

    def destroy_loop(self):
        if getattr(self, "loop", None):
            torch.cuda.synchronize(self.server.device)
            self.plan.lib.sem_loop_destroy(self.loop)
            self.loop = None

    def __del__(self):
        # the loop's device and host-mapped memory; a lane's stream is only
        # destroyed by close()
        try:
            self.destroy_loop()
        except Exception:
            pass

    def ensure_set_capacity(self):
        if len(self.plan.rec_sizes) + 1 > self.max_sets:
            self.max_sets *= 2
            self._reserve()
            self.graph = None  # captured arguments point at the old table

    def outstanding(self) -> int:
        return self.npushed - self.ndone

    def busy(self) -> bool:
        return bool(self.outstanding() or self.queue or self.inflight or self.finishing)

    def push(self, req: SolveRequest, buf: Dict[str, torch.Tensor], hist: bool, footprint: int, now: float,
             nrec: int, rec_set: int, src_node: int, block: int):
        b = buf
        r = L.SemRequest()
        r.id, r.nt, r.full_history = req.rid, req.nt, 1 if hist else 0
        r.src_node, r.rec_set, r.nrec = src_node, rec_set, nrec
        r.inv_sigma2 = 1.0 / (req.noise_std ** 2 + 1e-24)
        for name, key in (("c2", "c2p"), ("wavelet", "wav"), ("seis", "seis"), ("resid", "res"), ("obs", "obs"),
                          ("misfit", "misfit"), ("G", "G"), ("grad_c", "gc"), ("w", "w"), ("w2", "w2"),
                          ("om", "om"), ("om2", "om2"), ("ring", "ring"), ("hist", "hist"), ("u0", "u0"),
                          ("u1", "u1"), ("cku", "cku"), ("ckw", "ckw"), ("seg", "seg")):
            if key in b:
                setattr(r, name, b[key].data_ptr())
        with torch.cuda.stream(self.stream):
            L.check(self.plan.lib.sem_loop_push(self.loop, C.byref(r), C.c_void_p(self.stream.cuda_stream)))
        self.pushed[req.rid] = _Pushed(req, b, footprint, now, nrec, block)
        self.npushed += 1

    def launch_round(self, use_graph: bool):
        lib, sp = self.plan.lib, C.c_void_p(self.stream.cuda_stream)
        with torch.cuda.stream(self.stream):
            if use_graph:
                if self.graph is None:
                    L.check(lib.sem_receivers_sync(self.plan.h))
                    torch.cuda.synchronize(self.server.device)
                    g = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(g, stream=self.stream):
                        L.check(lib.sem_loop_round(self.loop, sp))
                    self.graph = g
                self.graph.replay()
            else:
                L.check(lib.sem_loop_round(self.loop, sp))
            ev = torch.cuda.Event()
            ev.record(self.stream)
        self.round_events[self.rounds] = ev
        self.inflight.append(ev)
        self.rounds += 1

    def counts(self) -> dict:
        """The loop's counters: requests pushed, rounds enqueued and completed
        on the device, slots occupied after the last completed round."""
        pushed, enq, done = C.c_longlong(), C.c_longlong(), C.c_longlong()
        active = C.c_int()
        L.check(self.plan.lib.sem_loop_counts(self.loop, C.byref(pushed), C.byref(enq), C.byref(done),
                                              C.byref(active)))
        return {"pushed": pushed.value, "rounds_enqueued": enq.value, "rounds_done": done.value,
                "active": active.value, "completed": self.ndone}

    def poll(self):
        """(request id, round) of the solves completed since the last poll."""
        out = []
        total = C.c_longlong()
        while True:
            L.check(self.plan.lib.sem_loop_poll(self.loop, self.ndone, self._ids, self._rounds, len(self._ids),
                                                C.byref(total)))
            n = min(len(self._ids), total.value - self.ndone)
            out += [(int(self._ids[i]), int(self._rounds[i])) for i in range(n)]
            self.ndone += n
            if self.ndone >= total.value:
                return out


class SEMServer:
    """Continuous-batching forward+adjoint SEM service on one GPU."""

    def __init__(self, hbm_budget_bytes: float = 64e9, segment: int = 100, capacity: Optional[dict] = None,
                 adjoint_mode: str = "reference", sides=("xmin", "xmax", "zmax"), use_graph: bool = True,
                 device=None, rec_pitch: int = 200, nt_max: int = 5000, max_sets: int = 512,
                 autotune: bool = True, history: str = "auto", two_step: bool = True,
                 tile2: Optional[int] = None, loop: str = "device", lanes_per_class: Optional[int] = None):
        """hbm_budget_bytes: the device memory that holds the admitted
        requests' buffers, allocated once here as the server's arena; each
        request takes one block of it (a request that finds no block large
        enough waits).  history: 'auto' admits a request with its full forward history
        when that fits the budget (else with checkpoints every `segment`
        steps); 'checkpoints' always uses checkpoints.  two_step: advance
        every slot two steps per launch (sem_slots_step2; the default where
        the segment and correlation chunk are even: on the 48-request
        benchmark stream the device loop served 118-119 solves/s saturated
        against 106-110 with one step per launch, p99 0.25 against 0.30-0.33
        s).  Results are the same bits either way.
        tile2: the two-step kernel's tile shape in every lane (default: tuned
        per lane when two_step and autotune, which keeps the 19 x 13 tile
        unless another is clearly faster, else shape 1, 19 x 13); results do
        not depend on it.  loop: 'device'
        keeps the request queue, slot states and round scheduling on the
        device (sem_loop_*), 'host' schedules the rounds from the host; the
        same bits either way.  lanes_per_class: lanes (streams) per size
        class, its capacity shared among them (default: class_lanes' rule,
        two where one solve fills the GPU); the same bits in any lane."""
        if history not in ("auto", "checkpoints"):
            raise ValueError("history must be 'auto' or 'checkpoints'")
        if loop not in ("device", "host"):
            raise ValueError("loop must be 'device' or 'host'")
        self.loop = loop
        self.history = history
        # two steps per launch (sem_slots_step2); needs an even segment and
        # correlation chunk
        self.two_step = bool(two_step) and int(segment) % 2 == 0 and corr_chunk(int(segment)) % 2 == 0
        self.device = torch.device(device if device is not None else "cuda")
        self.budget = float(hbm_budget_bytes)
        self.S = int(segment)
        self.capacity = capacity or {}
        self.adjoint_mode = adjoint_mode
        self.sides = tuple(sides)
        self.use_graph = use_graph
        self.rec_pitch = int(rec_pitch)
        self.nt_max = int(nt_max)
        self.max_sets = int(max_sets)
        self.autotune = bool(autotune)
        self.tile2 = None if tile2 is None else int(tile2)
        self.lanes: Dict[tuple, Lane] = {}
        self.class_of: Dict[tuple, List[Lane]] = {}  # size class -> its lanes
        self.lanes_per_class = None if lanes_per_class is None else int(lanes_per_class)
        self.finish_stream = None
        self._finish_h = None
        self.arena = DeviceArena(int(self.budget), self.device)
        self.in_use = 0
        self.results: Dict[int, SolveResult] = {}
        self.t0 = None
        self.log = []

    def close(self):
        """Release the lanes' loops and destroy their streams and the
        finishing stream, after their work; the server is not used
        afterwards.  A server that is deleted without close() keeps its
        streams until the process ends."""
        for lane in getattr(self, "lanes", {}).values():
            lane.close()
        h, self._finish_h = getattr(self, "_finish_h", None), None
        if h is not None:
            self.finish_stream.synchronize()
            self.finish_stream = None
            release_stream(h)

    def clock(self) -> float:
        return time.perf_counter() - self.t0

    def class_lanes(self, nelem_x: int, nelem_z: int, capacity: int) -> int:
        """Lanes (streams) a size class is split into: `lanes_per_class` when
        set, else lanes_for_mesh's rule for this device; at most one per slot."""
        n = self.lanes_per_class
        if n is None:
            n = lanes_for_mesh(nelem_x, nelem_z, torch.cuda.get_device_properties(self.device).multi_processor_count)
        return max(1, min(n, capacity))

    def lane(self, req: SolveRequest) -> Lane:
        """The lane that takes `req`: of its size class's lanes (created on
        first use, the class's capacity shared out among them), the one with
        the least queued and outstanding work per slot."""
        key = req.lane_key
        if key not in self.class_of:
            cap = self.capacity.get(req.nelem_x, max(4, min(64, 2560 // max(req.nelem_x, 1))))
            n = self.class_lanes(req.nelem_x, req.nelem_z, cap)
            cls = LoopLane if self.loop == "device" else Lane
            lanes = []
            for i in range(n):
                k = key if i == 0 else key + (i,)
                self.lanes[k] = cls(self, k, cap // n + (1 if i < cap % n else 0), self.nt_max, self.rec_pitch)
                lanes.append(self.lanes[k])
            self.class_of[key] = lanes
        return min(self.class_of[key], key=lambda l: (len(l.queue) + l.outstanding()) / l.C)

    # -------------------------------------------------------------- admission
    def _admit(self, lane: Lane, req: SolveRequest, now: float) -> bool:
        device_loop = isinstance(lane, LoopLane)
        if device_loop:
            # the device queue holds up to one lane's worth of requests waiting for a slot
            slot = None
            if lane.outstanding() >= min(2 * lane.C, lane.QUEUE_CAPACITY):
                return False
        else:
            slot = lane.free_slot()
            if slot is None:
                return False
        plan = lane.plan
        dev, E, S = self.device, lane.E, self.S
        K = max(1, math.ceil(req.nt / S))
        nrec = int(np.asarray(req.rec_xz).reshape(-1, 2).shape[0])
        if nrec > lane.rec_pitch:
            raise ValueError(f"request {req.rid}: {nrec} receivers exceed the lane pitch {lane.rec_pitch}")
        if not 1 <= req.nt <= self.nt_max:
            raise ValueError(f"request {req.rid}: {req.nt} time steps outside 1..nt_max = {self.nt_max}")
        # one arena block per request: its full forward history when that fits
        # (history='auto'), else checkpoints; none free yet -> it waits
        lay_args = (E, req.nt, lane.rec_pitch, S, lane.chunk, lane.GE, plan.N if device_loop else 0)
        hist, block = False, None
        if self.history == "auto":
            layout, fp = request_layout(*lay_args, full_history=True, two_step=self.two_step)
            block = self.arena.take(fp)
            hist = block is not None
        if block is None:
            layout, fp = request_layout(*lay_args, full_history=False, two_step=self.two_step)
            block = self.arena.take(fp)
            if block is None:
                return False
        with torch.cuda.stream(lane.stream):
            b = {}
            for name, shape, dt, zero, off in layout:
                b[name] = self.arena.view(block + off, shape, dt)
                if zero:
                    b[name].zero_()
            if hist:
                # u^0 .. u^nt: slot m + 1 is written before any launch reads
                # it, so only u^0 = 0 needs clearing
                b["hist"][0].zero_()
            if req.obs is not None:
                b["obs"][:, :nrec] = staged(req.obs, torch.float64, dev)
            w = source_wavelet(req.nt, req.dt, req.f0)
            b["wav"].copy_(staged(w, torch.float32, dev))
            base = None if req.base is None else staged(req.base, torch.float64, dev)
            _, c2 = plan.bspline_map(staged(req.coef, torch.float64, dev)[None], base,
                                     want_c=False)
            L.check(plan.lib.sem_pad_nodal(plan.h, C.c_void_p(c2.data_ptr()), C.c_void_p(b["c2p"].data_ptr()), 1,
                                           C.c_void_p(lane.stream.cuda_stream)))
            b["c2_flat"] = c2
        rec_set = lane.receiver_set(req.rec_xz)
        src_node = plan.source_node(*req.src_xz)
        L.check(plan.lib.sem_receivers_sync(plan.h))
        if device_loop:
            lane.push(req, b, hist, fp, now, nrec, rec_set, src_node, block)
        else:
            a = _Active(req=req, slot=slot, K=K, phase=L.SLOT_FORWARD_HIST if hist else L.SLOT_FORWARD, k=0, buf=b,
                        footprint=fp, admitted=now, hist=hist, rec_set=rec_set, nrec=nrec, src_node=src_node,
                        block=block)
            lane.slots[slot] = a
        self.in_use += fp
        return True

    # -------------------------------------------------------------- transitions
    def _after_enqueue(self, lane: Lane):
        """Work that follows the round just enqueued, in stream order: final
        samples and residuals of the slots whose forward pass it completes,
        the slots' next phases, and the finishing work of the solves it
        completes (their slots free up at once; their memory when the results
        have reached the host)."""
        plan = lane.plan
        lib, h = plan.lib, plan.h
        sp = C.c_void_p(lane.stream.cuda_stream)
        fwd_phases = (L.SLOT_FORWARD, L.SLOT_FORWARD_HIST)
        finished_fwd = [a for a in lane.occupied() if a.phase in fwd_phases and a.k == a.K - 1]
        with torch.cuda.stream(lane.stream):
            if finished_fwd:
                # final post-update samples for the slots that just completed their forward pass
                L.check(lib.sem_slots_record_final(h, C.c_void_p(lane.table_dev.data_ptr()), lane.C,
                                                   (1 << L.SLOT_FORWARD) | (1 << L.SLOT_FORWARD_HIST),
                                                   lane.rec_pitch, sp))
                for a in finished_fwd:
                    if a.k == a.K - 1:
                        b = a.buf
                        L.check(lib.sem_residual(h, C.c_void_p(b["seis"].data_ptr()), C.c_void_p(b["obs"].data_ptr()),
                                                 C.c_void_p(b["res"].data_ptr()), C.c_void_p(b["misfit"].data_ptr()),
                                                 a.nrec, a.req.nt, lane.rec_pitch, sp))
            done = []
            for a in lane.occupied():
                if self.two_step and a.phase != L.SLOT_RECOMPUTE:
                    # a segment of ln steps took ceil(ln / 2) launches, each of
                    # which moved the increment to the other buffer
                    ln = min(self.S, a.req.nt - a.k * self.S)
                    if (ln + 1) // 2 % 2 == 1:
                        if a.phase == L.SLOT_ADJOINT:
                            a.omflip = not a.omflip
                        else:
                            a.wflip = not a.wflip
                if a.phase in fwd_phases:
                    if a.k + 1 < a.K:
                        a.k += 1
                    else:
                        a.phase, a.k = (L.SLOT_ADJOINT if a.hist else L.SLOT_RECOMPUTE), a.K - 1
                elif a.phase == L.SLOT_RECOMPUTE:
                    a.phase = L.SLOT_ADJOINT
                elif a.phase == L.SLOT_ADJOINT:
                    if a.k > 0:
                        a.phase, a.k = (L.SLOT_ADJOINT if a.hist else L.SLOT_RECOMPUTE), a.k - 1
                    else:
                        done.append(a)
            for a in done:
                self._finish(lane, a)

    def _finish(self, lane: Lane, a: _Active):
        """Enqueue (on the lane's stream, inside its stream context) the
        gradient of a completed solve, its projection onto the request's
        B-spline nodes and the copies of the results to pinned host memory."""
        plan, b, req = lane.plan, a.buf, a.req
        inv_s2 = 1.0 / (req.noise_std ** 2 + 1e-24)
        gc = torch.empty(plan.N, dtype=torch.float64, device=self.device)
        L.check(plan.lib.sem_gradient_finalize(plan.h, C.c_void_p(b["G"].data_ptr()), C.c_void_p(b["c2p"].data_ptr()),
                                               inv_s2, C.c_void_p(gc.data_ptr()), None,
                                               C.c_void_p(lane.stream.cuda_stream)))
        nbz, nbx = np.asarray(req.coef).shape
        g = plan.bspline_project(gc[None], nbz, nbx)[0]
        host = {"g": torch.empty(g.shape, dtype=g.dtype, pin_memory=True),
                "seis": torch.empty((req.nt, a.nrec), dtype=b["seis"].dtype, pin_memory=True),
                "misfit": torch.empty(1, dtype=torch.float64, pin_memory=True)}
        host["g"].copy_(g, non_blocking=True)
        host["seis"].copy_(b["seis"][:, :a.nrec], non_blocking=True)
        host["misfit"].copy_(b["misfit"], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(lane.stream)
        lane.finishing.append((a, ev, host, (gc, g)))
        lane.slots[a.slot] = None

    def _finish_pushed(self, lane: "LoopLane", rid: int, rnd: int):
        """Results of a solve the device loop completed in round `rnd`: its
        projection onto the request's B-spline nodes and the copies to pinned
        host memory, on the server's finishing stream once that round is done."""
        a = lane.pushed.pop(rid)
        plan, b, req = lane.plan, a.buf, a.req
        if self.finish_stream is None:
            self.finish_stream, self._finish_h = owned_stream(self.device)
        fs = self.finish_stream
        fs.wait_event(lane.round_events[rnd])
        for k in [k for k in lane.round_events if k < lane.rounds - 64]:
            del lane.round_events[k]
        with torch.cuda.stream(fs):
            nbz, nbx = np.asarray(req.coef).shape
            g = plan.bspline_project(b["gc"][None], nbz, nbx)[0]
            host = {"g": torch.empty(g.shape, dtype=g.dtype, pin_memory=True),
                    "seis": torch.empty((req.nt, a.nrec), dtype=b["seis"].dtype, pin_memory=True),
                    "misfit": torch.empty(1, dtype=torch.float64, pin_memory=True)}
            host["g"].copy_(g, non_blocking=True)
            host["seis"].copy_(b["seis"][:, :a.nrec], non_blocking=True)
            host["misfit"].copy_(b["misfit"], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(fs)
        lane.finishing.append((a, ev, host, (b, g)))

    def _collect(self, lane: Lane) -> bool:
        """Results of finished solves whose copies have landed; their memory
        returns to the budget."""
        got = False
        for item in list(lane.finishing):
            a, ev, host, _ = item
            if not ev.query():
                continue
            req = a.req
            inv_s2 = 1.0 / (req.noise_std ** 2 + 1e-24)
            mis = float(host["misfit"][0])
            self.results[req.rid] = SolveResult(req.rid, -mis * inv_s2, mis, host["g"].numpy().copy(),
                                                host["seis"].numpy().copy(), req.arrival, a.admitted, self.clock())
            lane.finishing.remove(item)
            self.in_use -= a.footprint
            self.arena.give(a.block, a.footprint)
            got = True
        return got

    # -------------------------------------------------------------- driver
    def run(self, requests: List[SolveRequest], progress=None, realtime: bool = True) -> Dict[int, SolveResult]:
        """Serve `requests` (arrival times in seconds from start) to completion.
        Lanes, their captured rounds and tile choices persist across calls (a
        warm server); each call returns its own results."""
        self.results = {}
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        self.t0 = time.perf_counter()
        total = len(pending)
        last_report = 0.0
        while pending or any(l.busy() for l in self.lanes.values()):
            now = self.clock()
            while pending and (pending[0].arrival <= now or not realtime):
                r = pending.popleft()
                if not realtime:
                    r.arrival = now
                self.lane(r).queue.append(r)
            progressed = False
            lanes = list(self.lanes.values())
            for lane in lanes:
                while lane.inflight and lane.inflight[0].query():
                    lane.inflight.popleft()
                    progressed = True
                if isinstance(lane, LoopLane):
                    for rid, rnd in lane.poll():
                        self._finish_pushed(lane, rid, rnd)
                        progressed = True
                progressed |= self._collect(lane)
            # admission one request per lane in turn, so that no lane takes
            # the arena's free memory for requests that would wait in its
            # device queue while another lane has requests and free slots
            admitted = True
            while admitted:
                admitted = False
                for lane in lanes:
                    if lane.queue and self._admit(lane, lane.queue[0], self.clock()):
                        lane.queue.popleft()
                        admitted = progressed = True
            for lane in lanes:
                device_loop = isinstance(lane, LoopLane)
                # keep up to ROUNDS_AHEAD rounds enqueued: admit, enqueue the
                # next round (host loop: then the work that follows it)
                while len(lane.inflight) < ROUNDS_AHEAD:
                    while lane.queue and self._admit(lane, lane.queue[0], self.clock()):
                        lane.queue.popleft()
                    if not (lane.outstanding() if device_loop else lane.occupied()):
                        break
                    lane.launch_round(self.use_graph)
                    if not device_loop:
                        self._after_enqueue(lane)
                    progressed = True
            if progress and self.clock() - last_report > 5.0:
                last_report = self.clock()
                progress(f"[serve] t={last_report:.1f}s done {len(self.results)}/{total} "
                         f"in flight {sum(l.outstanding() for l in self.lanes.values())} "
                         f"hbm {self.in_use / 1e9:.1f} GB")
            if not progressed:
                time.sleep(0.0005)
        torch.cuda.synchronize(self.device)
        return self.results


def run_plain(requests: List[SolveRequest], device=None, adjoint_mode="reference",
              sides=("xmin", "xmax", "zmax"), segment=100, progress=None,
              plans: Optional[dict] = None) -> Dict[int, SolveResult]:
    """The plain path: one request at a time through the batch path at batch 1
    (one launch sequence per solve, no overlap), in arrival order.  `plans`
    (a dict) keeps the plans of earlier calls for reuse."""
    dev = torch.device(device if device is not None else "cuda")
    plans = {} if plans is None else plans
    out = {}
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda q: q.arrival):
        now = time.perf_counter() - t0
        if now < r.arrival:
            time.sleep(r.arrival - now)
        key = r.lane_key + (r.nt,)
        if key not in plans:
            plans[key] = Plan(r.nelem_x, r.nelem_z, 0.0, r.xmax, 0.0, r.zmax, r.dt, r.nt, 4, r.sponge_width,
                              r.sponge_cmax, 1e-8, sides, precision="fp32", adjoint_mode=adjoint_mode,
                              checkpoint_every=segment, device=dev, schedule="auto")
        p = plans[key]
        start = time.perf_counter() - t0
        # one receiver set per layout, as the lanes keep them
        xz = np.ascontiguousarray(np.asarray(r.rec_xz, dtype=np.float64).reshape(-1, 2))
        ids = plans.setdefault(key + ("receiver sets",), {})
        rs = ids.get(xz.tobytes())
        if rs is None:
            rs = ids[xz.tobytes()] = p.add_receivers(xz)
        _, c2 = p.bspline_map(torch.as_tensor(r.coef, dtype=torch.float64)[None],
                              None if r.base is None else torch.as_tensor(r.base))
        w = torch.as_tensor(source_wavelet(r.nt, r.dt, r.f0), dtype=torch.float32, device=dev)
        o = p.forward_adjoint(c2, w, [p.source_node(*r.src_xz)], torch.as_tensor(r.obs, device=dev)[None],
                              r.noise_std, [rs])
        nbz, nbx = np.asarray(r.coef).shape
        g = p.bspline_project(o["grad_c"], nbz, nbx)[0]
        torch.cuda.synchronize(dev)
        fin = time.perf_counter() - t0
        nrec = p.rec_sizes[rs]
        out[r.rid] = SolveResult(r.rid, float(o["loglik"][0]), float(o["misfit"][0]), g.cpu().numpy(),
                                 o["seis"][0, :, :nrec].cpu().numpy(), r.arrival, start, fin)
        if progress and len(out) % 8 == 0:
            progress(f"[plain] {len(out)}/{len(requests)} t={fin:.1f}s")
    return out


def latency_stats(results: Dict[int, SolveResult]):
    lat = np.array([r.latency for r in results.values()])
    span = max(r.finished for r in results.values()) - min(r.arrival for r in results.values())
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "solves_per_s": len(results) / span, "makespan_s": span,
            "n": len(results)}
