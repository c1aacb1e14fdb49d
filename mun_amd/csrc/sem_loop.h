// Device-resident request loop (sem_loop_* in include/sem_capi.h): the
// queue, the slot states and the per-round slot tables live in device memory,
// and small single-block kernels at the start and end of each round do what a
// host scheduler would otherwise do between rounds (pick queued requests into
// free slots, write the slot descriptors of the round, mark the slots whose
// forward pass or whole solve ends, move every slot to its next phase, publish
// completions).  The step, correlation, residual and gradient arithmetic is the
// batch path's, so a served solve gives the same bits alone, in any mix and on
// the batch path.
#pragma once

#include "sem_kernels.h"

namespace sem {

constexpr int LOOP_MAX_SLOTS = 256;
constexpr int LOOP_DONE_CAP = 4096;  // completion records kept in host memory

struct LoopSlot {
    sem_request req;
    int active;        // 1 while a request occupies the slot
    int phase, k, K;   // SEM_SLOT_* phase, current segment, segments
    int wflip, omflip;  // two-step loops: the increment of the forward / adjoint is in w2 / om2
    int fin_fwd;       // set at the start of a round: the slot's forward pass ends in it
    int done;          // set at the start of a round: the slot's solve completes in it
};

struct LoopCtl {
    long long qhead, qtail;  // requests popped / pushed
    long long round;         // rounds completed
    long long ndone;         // solves completed
};

// Host-mapped block (written by the device only; the host reads it).  ndone
// and round are published after the records they cover, behind a system-scope
// fence.
struct LoopHost {
    long long ndone;
    long long round;
    int active;
    int queued;
    long long ids[LOOP_DONE_CAP];
    long long id_round[LOOP_DONE_CAP];
};

struct LoopParams {
    int nslots, S, two_step, chunk, qcap, elem;
    long E;               // elements per padded nodal array
    sem_request* q;       // [qcap] queue ring
    LoopSlot* st;         // [nslots]
    sem_slot* table;      // [nslots] descriptors of the round's segments
    sem_slot* ftable;     // [nslots] the slots whose forward pass ended (others idle)
    LoopCtl* ctl;
    LoopHost* host;       // device address of the host-mapped block
};

__device__ __forceinline__ void* loop_off(void* p, long elems, int elem) {
    return p ? (void*)((char*)p + elems * (long)elem) : nullptr;
}

// Descriptor of a slot's current segment (the same fields a host scheduler
// fills, sem_slot in include/sem_capi.h).
__device__ inline sem_slot loop_desc(const LoopParams& p, const LoopSlot& s) {
    sem_slot d = {};
    if (!s.active) return d;
    const sem_request& r = s.req;
    const int s0 = s.k * p.S;
    const int ln = min(p.S, r.nt - s0);
    d.phase = s.phase;
    d.n0 = s0;
    d.len = ln;
    d.src_node = r.src_node;
    d.rec_set = r.rec_set;
    d.record = 1;
    d.c2 = r.c2;
    d.wavelet = r.wavelet;
    d.seis = r.seis;
    d.resid = r.resid;
    d.G = r.G;
    d.seg_stride = p.E;
    d.seg = r.full_history ? loop_off(r.hist, (long)s0 * p.E, p.elem) : r.seg;
    if (s.phase == SEM_SLOT_FORWARD_HIST) {
        d.w = r.w;
        if (p.two_step) {
            d.w = s.wflip ? r.w2 : r.w;
            d.w2 = s.wflip ? r.w : r.w2;
        }
    } else if (s.phase == SEM_SLOT_FORWARD) {
        d.u0 = r.u0;
        d.u1 = r.u1;
        d.w = r.w;
        if (p.two_step) {
            d.w = s.wflip ? r.w2 : r.w;
            d.w2 = s.wflip ? r.w : r.w2;
            d.pair_parity = 1;
        }
        if (s.k + 1 < s.K) {
            d.ck_u = loop_off(r.cku, (long)(s.k + 1) * p.E, p.elem);
            d.ck_w = loop_off(r.ckw, (long)(s.k + 1) * p.E, p.elem);
        }
    } else if (s.phase == SEM_SLOT_RECOMPUTE) {
        d.u0 = loop_off(r.cku, (long)s.k * p.E, p.elem);
        d.w = loop_off(r.ckw, (long)s.k * p.E, p.elem);
        if (p.two_step) d.w2 = r.w2;  // the segment's increment is not needed after it
    } else if (s.phase == SEM_SLOT_ADJOINT) {
        d.w = r.om;
        if (p.two_step) {
            d.w = s.omflip ? r.om2 : r.om;
            d.w2 = s.omflip ? r.om : r.om2;
        }
        d.ring = r.ring;
        d.ring_slots = p.chunk + 1;
    }
    return d;
}

// One thread appends the request to the queue ring (the host keeps fewer than
// qcap requests pushed and not completed).
__global__ void loop_push_kernel(LoopParams p, sem_request r) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const long long t = p.ctl->qtail;
    p.q[t % p.qcap] = r;
    p.ctl->qtail = t + 1;
}

// Start of a round: free slots take queued requests in arrival order (lowest
// slot first), then every slot's descriptor and end-of-round flags are
// written.  One block.
__global__ __launch_bounds__(LOOP_MAX_SLOTS) void loop_begin_kernel(LoopParams p) {
    if (threadIdx.x == 0) {
        long long h = p.ctl->qhead;
        const long long t = p.ctl->qtail;
        for (int i = 0; i < p.nslots && h < t; ++i) {
            LoopSlot& s = p.st[i];
            if (s.active) continue;
            s.req = p.q[h % p.qcap];
            ++h;
            s.active = 1;
            s.K = max(1, (s.req.nt + p.S - 1) / p.S);
            s.k = 0;
            s.phase = s.req.full_history ? SEM_SLOT_FORWARD_HIST : SEM_SLOT_FORWARD;
            s.wflip = s.omflip = 0;
            s.fin_fwd = s.done = 0;
        }
        p.ctl->qhead = h;
    }
    __syncthreads();
    // the round's segments, and which slots end their forward pass in it
    // (final sample and residual follow the steps: ftable lists them) or
    // complete (the gradient follows)
    for (int i = threadIdx.x; i < p.nslots; i += blockDim.x) {
        LoopSlot& s = p.st[i];
        const sem_slot d = loop_desc(p, s);
        const bool fwd = s.phase == SEM_SLOT_FORWARD || s.phase == SEM_SLOT_FORWARD_HIST;
        s.fin_fwd = s.active && fwd && s.k == s.K - 1;
        s.done = s.active && s.phase == SEM_SLOT_ADJOINT && s.k == 0;
        p.table[i] = d;
        p.ftable[i] = d;
        if (!s.fin_fwd) p.ftable[i].phase = SEM_SLOT_IDLE;
    }
}

// Residual and misfit of the slots whose forward pass ended: one block per
// slot, the batch path's residual_rows.
template <typename T>
__global__ __launch_bounds__(RESID_THREADS) void loop_residual_kernel(LoopParams p, int rec_pitch) {
    const LoopSlot& s = p.st[blockIdx.x];
    if (!s.fin_fwd) return;  // uniform per block
    __shared__ double red[RESID_THREADS];
    const sem_request& r = s.req;
    residual_rows<T>((const T*)r.seis, r.obs, (T*)r.resid, r.misfit, r.nrec, r.nt, rec_pitch, 0, r.nt, 1, red);
}

// d loglik / dc of the completed slots: the batch path's assembly and scaling
// (grad_finalize_kernel), blockIdx.y = slot.
template <typename T>
__global__ void loop_finalize_kernel(LoopParams p, long sb, long gplane, int nexp, int nex, int nez, int NX, int NZ,
                                     int pitch, double J) {
    const LoopSlot& s = p.st[blockIdx.y];
    if (!s.done) return;
    const sem_request& r = s.req;
    const double scale = J * r.inv_sigma2;
    const long N = (long)NX * NZ;
    const T* c2 = (const T*)r.c2;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        const int jz = (int)(k / NX), ix = (int)(k - (long)jz * NX);
        const double g2 = scale * assemble_node(r.G, sb, gplane, nexp, nex, nez, pitch, ix, jz);
        r.grad_c[k] = 2.0 * sqrt((double)c2[(long)jz * pitch + ix]) * g2;
    }
}

// End of a round: every slot moves to its next phase (two-step loops first
// note where the segment left the increment), completed slots are freed and
// their ids published to the host in slot order.  One block, one thread: the
// transitions are a few integer updates per slot.
__global__ void loop_advance_kernel(LoopParams p) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    long long nd = p.ctl->ndone;
    const long long round = p.ctl->round;
    int active = 0;
    for (int i = 0; i < p.nslots; ++i) {
        LoopSlot& s = p.st[i];
        if (!s.active) continue;
        if (p.two_step && s.phase != SEM_SLOT_RECOMPUTE) {
            // a segment of ln steps took ceil(ln / 2) launches, each of which
            // moved the increment to the other buffer
            const int ln = min(p.S, s.req.nt - s.k * p.S);
            if (((ln + 1) / 2) % 2 == 1) {
                if (s.phase == SEM_SLOT_ADJOINT)
                    s.omflip ^= 1;
                else
                    s.wflip ^= 1;
            }
        }
        bool done = false;
        if (s.phase == SEM_SLOT_FORWARD || s.phase == SEM_SLOT_FORWARD_HIST) {
            if (s.k + 1 < s.K) {
                ++s.k;
            } else {
                s.phase = s.req.full_history ? SEM_SLOT_ADJOINT : SEM_SLOT_RECOMPUTE;
                s.k = s.K - 1;
            }
        } else if (s.phase == SEM_SLOT_RECOMPUTE) {
            s.phase = SEM_SLOT_ADJOINT;
        } else if (s.phase == SEM_SLOT_ADJOINT) {
            if (s.k > 0) {
                s.phase = s.req.full_history ? SEM_SLOT_ADJOINT : SEM_SLOT_RECOMPUTE;
                --s.k;
            } else {
                done = true;
            }
        }
        if (done) {
            p.host->ids[nd % LOOP_DONE_CAP] = s.req.id;
            p.host->id_round[nd % LOOP_DONE_CAP] = round;
            ++nd;
            s.active = 0;
            s.phase = SEM_SLOT_IDLE;
        } else {
            ++active;
        }
    }
    p.ctl->ndone = nd;
    p.ctl->round = round + 1;
    p.host->active = active;
    p.host->queued = (int)(p.ctl->qtail - p.ctl->qhead);
    __threadfence_system();  // the records above reach the host before the counters
    p.host->round = round + 1;
    p.host->ndone = nd;
    __threadfence_system();
}

}  // namespace sem
