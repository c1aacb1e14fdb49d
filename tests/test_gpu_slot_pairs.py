"""Forward slots advanced two steps per launch (sem_slots_step2 with
sem_slot.pair_parity) against the same slots one step per launch
(sem_slots_step), called through the C ABI on hand-built slot tables: the
final state, increment, checkpoint and seismograms are the same bits; and
the tables the header names as invalid leave a slot untouched."""
import ctypes as C

import numpy as np
import pytest
import torch

from mun_amd import _lib as L

pytestmark = pytest.mark.gpu

NT, PITCH = 40, 32
FWD = 1 << L.SLOT_FORWARD


@pytest.fixture(scope="module")
def setup():
    from mun_amd.solver import Plan, source_wavelet
    dev = torch.device("cuda:0")
    nex, nez, lx, lz, dt = 12, 9, 300.0, 225.0, 4e-4
    plan = Plan(nex, nez, 0.0, lx, 0.0, lz, dt, NT, 4, 60.0, 3500.0, 1e-8, ("xmin", "xmax", "zmax"),
                precision="fp32", record="post", checkpoint_every=10, device=dev)
    rng = np.random.default_rng(4)
    _, c2 = plan.bspline_map(torch.as_tensor(rng.uniform(1800.0, 3000.0, (5, 5)))[None])
    pitch, elems = C.c_int(), C.c_long()
    L.check(plan.lib.sem_plan_layout(plan.h, C.byref(pitch), C.byref(elems)))
    E = elems.value
    c2p = torch.zeros(E, dtype=torch.float32, device=dev)
    sp = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    L.check(plan.lib.sem_pad_nodal(plan.h, C.c_void_p(c2.data_ptr()), C.c_void_p(c2p.data_ptr()), 1, sp))
    rec = np.stack([np.linspace(20.0, lx - 20.0, 9), np.full(9, 2.0)], axis=1)
    rset = plan.add_receivers(rec)
    wav = torch.as_tensor(source_wavelet(NT, dt, 12.0), dtype=torch.float32, device=dev)
    src = plan.source_node(0.45 * lx, 30.0)
    # the plan runs (and holds its receiver tables) before the slot launches
    seis0 = plan.forward(c2, wav, [src], [rset])
    assert seis0.abs().max() > 0
    # every segment starts from the same random nonzero state
    u_init = torch.as_tensor(rng.normal(0.0, 1e-3, E), dtype=torch.float32, device=dev)
    w_init = torch.as_tensor(rng.normal(0.0, 1e-4, E), dtype=torch.float32, device=dev)
    return dict(plan=plan, E=E, c2p=c2p, rset=rset, wav=wav, src=src, u_init=u_init, w_init=w_init, sp=sp,
                dev=dev, pitch=pitch.value, NX=4 * nex + 1, NZ=4 * nez + 1)


def _nodes(s, t):
    """The mesh nodes of a padded nodal array (kernels never write the
    padding, which keeps whatever the buffer held)."""
    return t[:s["NZ"] * s["pitch"]].view(s["NZ"], s["pitch"])[:, :s["NX"]]


def _buffers(s, pair, n0, ck):
    """Per-slot buffers with the segment's starting state where the slot's
    convention keeps u^{n0}: by step parity, or by step pair."""
    E, dev = s["E"], s["dev"]
    b = {k: torch.zeros(E, dtype=torch.float32, device=dev) for k in ("u0", "u1", "w", "w2")}
    hi = ((n0 + 1) >> 1) & 1 if pair else n0 & 1
    b["u1" if hi else "u0"].copy_(s["u_init"])
    b["w"].copy_(s["w_init"])
    if ck:
        b["ck_u"] = torch.zeros(E, dtype=torch.float32, device=dev)
        b["ck_w"] = torch.zeros(E, dtype=torch.float32, device=dev)
    b["seis"] = torch.zeros(NT, PITCH, dtype=torch.float32, device=dev)
    return b


def _slot(s, b, n0, ln, pair, w2=True):
    d = L.SemSlot()
    d.phase, d.n0, d.len, d.src_node, d.rec_set, d.record = L.SLOT_FORWARD, n0, ln, s["src"], s["rset"], 1
    d.c2, d.wavelet, d.seis = s["c2p"].data_ptr(), s["wav"].data_ptr(), b["seis"].data_ptr()
    d.u0, d.u1, d.w = b["u0"].data_ptr(), b["u1"].data_ptr(), b["w"].data_ptr()
    if w2:
        d.w2 = b["w2"].data_ptr()
    if "ck_u" in b:
        d.ck_u, d.ck_w = b["ck_u"].data_ptr(), b["ck_w"].data_ptr()
    d.pair_parity = 1 if pair else 0
    return d


def _table(s, descs):
    arr = (L.SemSlot * len(descs))(*descs)
    return torch.frombuffer(bytearray(arr), dtype=torch.uint8).to(s["dev"])


def _run(s, descs, two, maxlen, final=True):
    lib, h, sp = s["plan"].lib, s["plan"].h, s["sp"]
    tab = _table(s, descs)
    tp = C.c_void_p(tab.data_ptr())
    if two:
        for m in range(0, maxlen, 2):
            L.check(lib.sem_slots_step2(h, tp, len(descs), m, FWD, PITCH, sp))
    else:
        for m in range(maxlen):
            L.check(lib.sem_slots_step(h, tp, len(descs), m, FWD, PITCH, sp))
    if final:
        L.check(lib.sem_slots_record_final(h, tp, len(descs), FWD, PITCH, sp))
    torch.cuda.synchronize(s["dev"])


# (n0, len, checkpoint): a checkpointed even segment, an odd final segment
# (its last step taken alone), a one-step segment
SEGMENTS = [(0, 10, True), (10, 7, False), (4, 1, False)]


def test_paired_forward_slots_give_the_one_step_bits(setup):
    s = setup
    one = [_buffers(s, False, n0, ck) for n0, _, ck in SEGMENTS]
    two = [_buffers(s, True, n0, ck) for n0, _, ck in SEGMENTS]
    maxlen = max(ln for _, ln, _ in SEGMENTS)
    _run(s, [_slot(s, b, n0, ln, False, w2=False) for b, (n0, ln, _) in zip(one, SEGMENTS)], False, maxlen)
    _run(s, [_slot(s, b, n0, ln, True) for b, (n0, ln, _) in zip(two, SEGMENTS)], True, maxlen)
    for (n0, ln, ck), a, b in zip(SEGMENTS, one, two):
        n = n0 + ln
        ua = _nodes(s, a["u1" if n & 1 else "u0"])
        ub = _nodes(s, b["u1" if ((n + 1) >> 1) & 1 else "u0"])
        # ceil(len / 2) launches, each moving the increment to the other buffer
        wa = _nodes(s, a["w"])
        wb = _nodes(s, b["w2" if ((ln + 1) // 2) % 2 else "w"])
        print(f"  segment n0={n0} len={ln}: |u| {ua.abs().max().item():.3e}, "
              f"seis rows {int((a['seis'].abs().sum(1) > 0).sum())}", flush=True)
        assert ua.abs().max() > 0 and not torch.equal(ua, _nodes(s, s["u_init"]))
        assert torch.equal(ua, ub), (n0, ln)
        assert torch.equal(wa, wb), (n0, ln)
        assert torch.equal(a["seis"], b["seis"]), (n0, ln)
        if ck:
            for k in ("ck_u", "ck_w"):
                assert torch.equal(_nodes(s, a[k]), _nodes(s, b[k])), k
            assert torch.equal(_nodes(s, a["ck_u"]), ua) and torch.equal(_nodes(s, a["ck_w"]), wa)


@pytest.mark.parametrize("case", ["no_pair_parity", "odd_n0", "odd_len_with_checkpoint", "w2_missing"])
def test_invalid_paired_tables_leave_the_slot_untouched(setup, case):
    """sem_slots_step2 does not advance a forward slot without pair_parity,
    with an odd n0, with a checkpoint and an odd len, or without a second
    increment buffer."""
    s = setup
    n0, ln, ck, pair, w2 = {"no_pair_parity": (0, 6, False, False, True), "odd_n0": (3, 6, False, True, True),
                            "odd_len_with_checkpoint": (0, 7, True, True, True),
                            "w2_missing": (0, 6, False, True, False)}[case]
    b = _buffers(s, True, n0, ck)
    before = {k: v.clone() for k, v in b.items()}
    _run(s, [_slot(s, b, n0, ln, pair, w2=w2)], True, ln, final=False)
    for k in b:
        assert torch.equal(b[k], before[k]), (case, k)


def test_one_step_launches_leave_paired_slots_untouched(setup):
    """sem_slots_step does not advance a forward slot with pair_parity set
    (its state is kept by step pair, which one-step launches would break)."""
    s = setup
    b = _buffers(s, True, 0, False)
    before = {k: v.clone() for k, v in b.items()}
    _run(s, [_slot(s, b, 0, 6, True)], False, 6, final=False)
    for k in b:
        assert torch.equal(b[k], before[k]), k
