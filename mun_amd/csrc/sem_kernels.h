// Device kernels of the MI355X SEM solver (gfx950 / CDNA4).
//
// Discretisation (restated from Bspline_Coupling/SEM/sem_waveform/operators.py:6-42,
// 65-84, 193-242 and core.py:1197-1275):
//   element stiffness  Ke = Gx^T W Gx + Gz^T W Gz,  W = diag(w_a w_b J c^2)
//   Gx = (2/dx) kron(I, D), Gz = (2/dz) kron(D, I)   (GLL collocation: L = I)
//   lumped mass        M = J Wx(ix) Wz(jz)   (Wx, Wz: assembled 1-D GLL weights)
//   sponge             Dn = exp(-(sx + sz) dt) = Dx(ix) Dz(jz)
//
// Time stepping is in increment form.  With w^n = u^n - D u^{n-1} the
// reference recurrence u^{n+1} = D[2u^n - D u^{n-1} + a^n] becomes
//   w^{n+1} = D (w^n + a^n),   u^{n+1} = D u^n + w^{n+1},
//   a^n = dt^2 M^-1 (s_n M_src e_src - K u^n),
// which is the same recurrence in exact arithmetic but keeps the small
// increment at full FP32 precision.  The adjoint uses the same step with the
// receiver residual as source (and D = 1 in the reference mode).
//
// Memory layout: every nodal array is [solve][NZ][pitch] with pitch a
// multiple of 32 elements (128-byte rows) and 256-byte aligned solves.
//
// Work decomposition: one thread per element.  A workgroup covers a block of
// EX x EZ = 17 x 15 elements: 16 x 14 owned elements plus one halo column on
// the left and one halo row below.  Owned x-ranges are 64 nodes (256 bytes)
// and start on 256-byte boundaries, so owned rows are written as whole cache
// lines.  Each thread keeps its element's 5 x 5 values in registers and
// computes the 25 local contributions; element-edge contributions reach the
// owning element through a 9-value LDS exchange.  Every node is written by
// exactly one thread (owner computes, no atomics), so results are bitwise
// deterministic and independent of batch composition.
//
// Workgroups are numbered so that each XCD (workgroup id mod 8 under the
// hardware's round-robin dispatch) walks a contiguous range of (tile, solve)
// pairs: neighbouring tiles share halo lines in the same L2.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/sem_capi.h"

namespace sem {

constexpr int EX = 17;            // element threads along x per workgroup
constexpr int TX = EX - 1;        // owned elements along x per tile (64 nodes)
// Tile heights (element rows incl. the halo row) with a kernel instance:
// 15 rows -> 256-thread workgroups, 7 -> 128, 3 -> 64.
constexpr int EZ_DEFAULT = 15;
__host__ __device__ constexpr int tile_threads(int ez) { return (EX * ez + 63) / 64 * 64; }
constexpr int NXCD = 8;           // XCDs of an MI355X
// Last-level (Infinity) cache of an MI355X: launches whose working set fits
// run their state streams with plain accesses, larger ones non-temporally.
constexpr size_t LLC_BYTES = size_t(256) << 20;

// MODE_ADJS: adjoint step; its gradient correlation runs later per chunk of
// steps in corr_kernel.  MODE_ANY (slot launches only): each workgroup takes
// the forward or the adjoint step by its slot's phase, so one launch advances
// slots of both directions; with a fixed MODE the per-workgroup direction
// flags below are compile-time constants and fold away.
enum StepMode { MODE_FWD = 0, MODE_REC = 2, MODE_ADJS = 3, MODE_ANY = 4 };

// Global-memory qualified pointers: lets the compiler emit global_* instead of
// flat_* instructions for pointers that come from the slot table.
#define GMEM __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ GMEM T* gm(T* p) {
    return (GMEM T*)p;
}
template <typename T>
__device__ __forceinline__ const GMEM T* gm(const T* p) {
    return (const GMEM T*)p;
}

// a * b + c rounded once.  The element action, the update and the receiver
// sums are written with explicit fused multiply-adds and with automatic
// contraction off (`#pragma clang fp contract(off)` at the top of each), so
// that every kernel that shares them rounds identically, whatever the
// surrounding control flow lets the compiler fuse: one- and two-step launches
// give the same bits.
template <typename T>
__device__ __forceinline__ T fmadd(T a, T b, T c) {
    if constexpr (std::is_same_v<T, float>) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}

// y = M x for a 5 x 5 matrix with M[a][b] = -M[4-a][4-b] (the GLL
// differentiation matrix, and D^T scaled by the symmetric weights, are of this
// form): with e_b = x_b + x_{4-b}, o_b = x_b - x_{4-b},
//   y_a, y_{4-a} = S_a +- T_a,  S_a = s_a0 o_0 + s_a1 o_1,
//   T_a = t_a0 e_0 + t_a1 e_1 + t_a2 x_2   (a = 0, 1),
//   y_2 = m0 o_0 + m1 o_1,
// s_ab = (M[a][b] - M[a][4-b]) / 2, t_ab = (M[a][b] + M[a][4-b]) / 2,
// t_a2 = M[a][2], m_b = M[2][b]: 20 operations instead of 25.
template <typename T>
struct Fold5 {
    T s[2][2], t[2][3], m[2];
};

template <typename T>
__device__ __forceinline__ void apply5(const Fold5<T>& f, T x0, T x1, T x2, T x3, T x4, T (&y)[5]) {
#pragma clang fp contract(off)
    const T e0 = x0 + x4, o0 = x0 - x4, e1 = x1 + x3, o1 = x1 - x3;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const T S = fmadd(f.s[a][1], o1, f.s[a][0] * o0);
        const T D = fmadd(f.t[a][2], x2, fmadd(f.t[a][1], e1, f.t[a][0] * e0));
        y[a] = S + D;
        y[4 - a] = S - D;
    }
    y[2] = fmadd(f.m[1], o1, f.m[0] * o0);
}

// apply5 of the products x_b c_b, with the even / odd sums formed by fused
// multiply-adds (x_0 c_0 + x_4 c_4 rounded once per sum).
template <typename T>
__device__ __forceinline__ void apply5_scaled(const Fold5<T>& f, const T (&x)[5], const T (&c)[5], T (&y)[5]) {
#pragma clang fp contract(off)
    const T p4 = x[4] * c[4], p3 = x[3] * c[3], x2 = x[2] * c[2];
    const T e0 = fmadd(x[0], c[0], p4), o0 = fmadd(x[0], c[0], -p4);
    const T e1 = fmadd(x[1], c[1], p3), o1 = fmadd(x[1], c[1], -p3);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const T S = fmadd(f.s[a][1], o1, f.s[a][0] * o0);
        const T D = fmadd(f.t[a][2], x2, fmadd(f.t[a][1], e1, f.t[a][0] * e0));
        y[a] = S + D;
        y[4 - a] = S - D;
    }
    y[2] = fmadd(f.m[1], o1, f.m[0] * o0);
}

template <typename T, int P>
struct ElemConst {
    T D[P + 1][P + 1];   // D[a][k] = l_k'(xi_a): derivative at node a
    T Dw[P + 1][P + 1];  // Dw[a][i] = D[a][i] * w_a: weighted back-projection
    T ax[P + 1];         // (2/dx)^2 w_j : row factor of the x part
    T az[P + 1];         // (2/dz)^2 w_i : column factor of the z part
    T gx[P + 1][P + 1];  // (2/dx)^2 w_j w_q : gradient weight, x part
    T gz[P + 1][P + 1];  // (2/dz)^2 w_q w_i : gradient weight, z part
    T iw[P + 1];         // 1 / w_i          : inverse assembled weight inside an element
    T iw_shared;         // 1 / (w_0 + w_P)  : inverse weight on a shared element edge
    T dt2;               // dt^2
    // D and Dw^T folded for apply5 (both are centro-antisymmetric on GLL nodes)
    Fold5<T> fD, fDwT;
};

template <typename T, int P>
struct RecEntry {
    int ex, ez;          // element holding the receiver
    int r;               // receiver column within its set
    int owned;           // 1 in the tile that owns the element (writes the sample)
    T lx[P + 1];         // Lagrange weights along x
    T lz[P + 1];         // Lagrange weights along z
};

template <typename T, int P>
struct StepArgs {
    ElemConst<T, P> ec;
    int nex, nez, NX, NZ, pitch;
    int ntx, ntiles, total, nper;      // tiles, tiles*batch, workgroups per XCD
    long sb;                           // elements between solves of every nodal array
    const T* u_in;                     // u^n (or lambda_k), read with halo
    T* u_out;                          // u^{n+1} (or lambda_{k-1}), owned nodes
    T* w;                              // increment, updated in place
    const T* c2;                       // padded squared velocity, per model
    const int* model_of;
    const T* Dx;                       // [NX padded to 4]  exp(-sx dt)
    const T* Dz;                       // [NZ padded to 4]  exp(-sz dt)
    int damp;
    int damp_adj;                      // MODE_ANY launches: damp of the adjoint slots
    // point source (forward): a[src] += dt^2 s_n
    const int* src_node;               // lexicographic node id (jz * NX + ix)
    const T* wavelet; long wav_bs; int n;
    // receivers
    const int* rec_set_of;             // [batch] receiver set id
    const int* rec_ptr;                // [nsets][ntiles + 1] offsets into recs
    const RecEntry<T, P>* recs;
    T* seis;         long seis_bs; int rec_pitch; int rec_t;
    // checkpoint mirrors of (u^{n+1}, w^{n+1}), may be null
    T* ck_u;         T* ck_w;
    // adjoint: residual source
    const T* resid;  long resid_bs; int res_t; T invJ;
    // slot mode: per-solve pointers and steps come from a device slot table
    const sem_slot* slots;  // null: uniform batch mode
    int m;                  // step index within the slots' current segments
    unsigned phase_mask;    // slots advanced by this launch (bit 1 << SEM_SLOT_*)
    int post;               // record mode: 1 = sample after the update
    // two-step launches (step2_kernel): u^{n+1} (null: not stored), the
    // increment's second buffer (read a.w, write w_out), and the receiver
    // table and tile counts of the two-step tiling
    T* u_mid;
    T* w_out;
    T* u_copy;  // when set: u^{n+2} also here (plain stores), u_out then non-temporal (full-history forward)
    const int* rec_ptr2;
    const RecEntry<T, P>* recs2;
    int ntx2, ntiles2, total2, nper2;
};

// Per-workgroup pointers and indices of one solve at one step.
template <typename T>
struct View {
    const GMEM T* u_in;
    GMEM T* u_out;
    GMEM T* w;
    const GMEM T* c2;
    GMEM T* ck_u;
    GMEM T* ck_w;
    GMEM T* seis;
    const GMEM T* resid;
    const GMEM T* wav;  // wavelet sample of this step (null: no source)
    int src, set, rec_t, res_t;
};

// Resolve the view of solve b; returns false if this solve does not advance.
template <typename T, int P, int MODE>
__device__ __forceinline__ bool make_view(const StepArgs<T, P>& a, long b, View<T>& v) {
    if (!a.slots) {
        const long soff = b * a.sb;
        const long mdl = a.model_of ? a.model_of[b] : b;
        v.u_in = gm(a.u_in + soff);
        v.u_out = gm(a.u_out + soff);
        v.w = gm(a.w + soff);
        v.c2 = gm(a.c2 + mdl * a.sb);
        v.ck_u = a.ck_u ? gm(a.ck_u + soff) : nullptr;
        v.ck_w = a.ck_w ? gm(a.ck_w + soff) : nullptr;
        v.seis = a.seis ? gm(a.seis + b * a.seis_bs) : nullptr;
        v.resid = a.resid ? gm(a.resid + b * a.resid_bs) : nullptr;
        v.wav = (MODE == MODE_FWD && a.src_node) ? gm(a.wavelet + b * a.wav_bs + a.n) : nullptr;
        v.src = (MODE == MODE_FWD && a.src_node) ? a.src_node[b] : -1;
        v.set = a.rec_set_of ? a.rec_set_of[b] : 0;
        v.rec_t = a.rec_t;
        v.res_t = a.res_t;
        return true;
    }
    const sem_slot& d = a.slots[b];
    if (!((a.phase_mask >> d.phase) & 1u)) return false;
    v.c2 = gm((const T*)d.c2);
    v.set = d.rec_set;
    v.seis = gm((T*)d.seis);
    v.resid = gm((const T*)d.resid);
    v.ck_u = v.ck_w = nullptr;
    v.wav = nullptr;
    v.src = -1;
    v.rec_t = v.res_t = -1;
    if (MODE == MODE_REC) {  // final post-update sample of a forward slot
        const int n = d.n0 + d.len;
        const bool hi = d.pair_parity ? (((n + 1) >> 1) & 1) : (n & 1);
        v.u_in = d.phase == SEM_SLOT_FORWARD_HIST ? gm((const T*)d.seg + (long)d.len * d.seg_stride)
                                                  : gm((const T*)(hi ? d.u1 : d.u0));
        v.u_out = nullptr;
        v.w = nullptr;
        v.rec_t = n - 1;
        return (d.phase == SEM_SLOT_FORWARD || d.phase == SEM_SLOT_FORWARD_HIST) && d.record;
    }
    if (a.m >= d.len) return false;
    if (d.phase == SEM_SLOT_FORWARD || d.phase == SEM_SLOT_RECOMPUTE || d.phase == SEM_SLOT_FORWARD_HIST) {
        const int n = d.n0 + a.m;
        v.wav = gm((const T*)d.wavelet + n);
        v.src = d.src_node;
        if (d.phase == SEM_SLOT_FORWARD) {
            if (d.pair_parity) return false;  // two-step launches advance it
            v.u_in = gm((const T*)((n & 1) ? d.u1 : d.u0));
            v.u_out = gm((T*)((n & 1) ? d.u0 : d.u1));
            v.w = gm((T*)d.w);
            v.rec_t = d.record ? (a.post ? n - 1 : n) : -1;
            if (a.m == d.len - 1 && d.ck_u) {
                v.ck_u = gm((T*)d.ck_u);
                v.ck_w = gm((T*)d.ck_w);
            }
        } else if (d.phase == SEM_SLOT_FORWARD_HIST) {
            v.u_in = gm((const T*)d.seg + (long)a.m * d.seg_stride);
            v.u_out = gm((T*)d.seg + (long)(a.m + 1) * d.seg_stride);
            v.w = gm((T*)d.w);
            v.rec_t = d.record ? (a.post ? n - 1 : n) : -1;
        } else {
            v.u_in = gm((a.m == 0) ? (const T*)d.u0 : (const T*)d.seg + (long)a.m * d.seg_stride);
            v.u_out = gm((T*)d.seg + (long)(a.m + 1) * d.seg_stride);
            v.w = gm((T*)d.w);
        }
        return MODE == MODE_FWD;
    }
    // adjoint: lambda_t in ring slot (t + 1) % ring_slots
    const int it = d.n0 + d.len - 1 - a.m;
    v.u_in = gm((const T*)d.ring + (long)((it + 1) % d.ring_slots) * d.seg_stride);
    v.u_out = gm((T*)d.ring + (long)(it % d.ring_slots) * d.seg_stride);
    v.w = gm((T*)d.w);
    v.res_t = it;
    return MODE == MODE_ADJS;
}

typedef float vf4 __attribute__((ext_vector_type(4)));
typedef double vd2 __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ void load4(const GMEM T* __restrict__ p, T (&v)[4]) {
    if constexpr (std::is_same_v<T, float>) {
        const vf4 a = *reinterpret_cast<const GMEM vf4*>(p);
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
        v[3] = a.w;
    } else {
        const vd2 a = *reinterpret_cast<const GMEM vd2*>(p);
        const vd2 c = *reinterpret_cast<const GMEM vd2*>(p + 2);
        v[0] = a.x;
        v[1] = a.y;
        v[2] = c.x;
        v[3] = c.y;
    }
}

// p + k with the byte offset formed in 32 bits (one solve's padded nodal
// array stays below 4 GiB, checked when a plan is created): the compiler then
// addresses with the workgroup-uniform base in SGPRs plus a 32-bit VGPR offset
// instead of a 64-bit address per lane and row.
template <typename T>
__device__ __forceinline__ const GMEM T* at(const GMEM T* p, unsigned k) {
    return reinterpret_cast<const GMEM T*>(reinterpret_cast<const GMEM char*>(p) + k * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ GMEM T* at(GMEM T* p, unsigned k) {
    return reinterpret_cast<GMEM T*>(reinterpret_cast<GMEM char*>(p) + k * (unsigned)sizeof(T));
}

// An element's 5 x 5 block in two halves.  fetch_block issues one 16-byte load
// per row for the first four nodes and, only in lanes whose right neighbour
// element is not held by lane + 1, a load of the fifth; finish_block takes the
// fifth node of the other lanes from lane + 1, whose first node it is (a
// cross-lane read instead of a memory instruction).  finish_block must run in
// converged control flow.
// (The block's first node is at element offset `off` from the uniform base p.)
template <typename T>
__device__ __forceinline__ void fetch_block(const GMEM T* __restrict__ p, unsigned off, unsigned pitch,
                                            bool own_fifth, T (&v)[5][5]) {
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        T q[4];
        load4(at(p, off + j * pitch), q);
        v[j][0] = q[0];
        v[j][1] = q[1];
        v[j][2] = q[2];
        v[j][3] = q[3];
    }
    if (own_fifth) {
#pragma unroll
        for (int j = 0; j < 5; ++j) v[j][4] = *at(p, off + j * pitch + 4);
    }
}

template <typename T>
__device__ __forceinline__ void finish_block(bool from_right, T (&v)[5][5]) {
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const T r = __shfl_down(v[j][0], 1, 64);
        if (from_right) v[j][4] = r;
    }
}

template <typename T>
__device__ __forceinline__ void store4(GMEM T* __restrict__ p, const T (&v)[4]) {
    if constexpr (std::is_same_v<T, float>) {
        const vf4 x = {v[0], v[1], v[2], v[3]};
        *reinterpret_cast<GMEM vf4*>(p) = x;
    } else {
        const vd2 x = {v[0], v[1]}, y = {v[2], v[3]};
        *reinterpret_cast<GMEM vd2*>(p) = x;
        *reinterpret_cast<GMEM vd2*>(p + 2) = y;
    }
}

// Non-temporal variants for streams touched once per step (the increment and
// the new state): they skip cache allocation, which keeps the halo-shared
// inputs (u, c^2) resident and lets reads and writes stream.  They pay off
// only when a launch's working set is larger than the last-level cache; when
// it fits (LLC_BYTES), plain accesses let the next step find the new state and
// the increment there (step_kernel's NTS parameter).
template <typename T>
__device__ __forceinline__ void store4nt(GMEM T* __restrict__ p, const T (&v)[4]) {
    if constexpr (std::is_same_v<T, float>) {
        const vf4 x = {v[0], v[1], v[2], v[3]};
        __builtin_nontemporal_store(x, reinterpret_cast<GMEM vf4*>(p));
    } else {
        const vd2 x = {v[0], v[1]}, y = {v[2], v[3]};
        __builtin_nontemporal_store(x, reinterpret_cast<GMEM vd2*>(p));
        __builtin_nontemporal_store(y, reinterpret_cast<GMEM vd2*>(p + 2));
    }
}

template <typename T>
__device__ __forceinline__ void load4nt(const GMEM T* __restrict__ p, T (&v)[4]) {
    if constexpr (std::is_same_v<T, float>) {
        const vf4 a = __builtin_nontemporal_load(reinterpret_cast<const GMEM vf4*>(p));
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
        v[3] = a.w;
    } else {
        const vd2 a = __builtin_nontemporal_load(reinterpret_cast<const GMEM vd2*>(p));
        const vd2 c = __builtin_nontemporal_load(reinterpret_cast<const GMEM vd2*>(p + 2));
        v[0] = a.x;
        v[1] = a.y;
        v[2] = c.x;
        v[3] = c.y;
    }
}

// Element contributions r = Ke u / J (x and z parts with folded constants).
template <typename T, int P>
__device__ __forceinline__ void element_action(const ElemConst<T, P>& ec, const T (&u)[P + 1][P + 1],
                                               const T (&c2)[P + 1][P + 1], T (&r)[P + 1][P + 1]) {
#pragma clang fp contract(off)
    static_assert(P == 4, "apply5 is the 5-node operator");
    constexpr int NP = P + 1;
    // x direction, row by row
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        T d[NP], s[NP];
        apply5(ec.fD, u[j][0], u[j][1], u[j][2], u[j][3], u[j][4], d);
        apply5_scaled(ec.fDwT, d, c2[j], s);
#pragma unroll
        for (int i = 0; i < NP; ++i) r[j][i] = ec.ax[j] * s[i];
    }
    // z direction, column by column
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        T d[NP], s[NP], cc[NP];
        apply5(ec.fD, u[0][i], u[1][i], u[2][i], u[3][i], u[4][i], d);
#pragma unroll
        for (int q = 0; q < NP; ++q) cc[q] = c2[q][i];
        apply5_scaled(ec.fDwT, d, cc, s);
#pragma unroll
        for (int j = 0; j < NP; ++j) r[j][i] = fmadd(ec.az[i], s[j], r[j][i]);
    }
}

// NTS: non-temporal loads and stores of the increment and the new state (for
// launches whose working set exceeds the last-level cache).  The FP32 adjoint
// step is compiled for 5 waves per SIMD (94 VGPRs, no spills), which hides
// more memory latency; the forward step and the FP64 instances keep the
// compiler's choice (103 VGPRs, 4 waves; FP64 about 200), since at 5 waves
// they spill.
template <typename T, int P, int MODE, int EZ, bool NTS>
__global__ __launch_bounds__(tile_threads(EZ))
__attribute__((amdgpu_waves_per_eu((MODE == MODE_ADJS && sizeof(T) == 4) ? 5 : 1)))
void step_kernel(const StepArgs<T, P> a) {
#pragma clang fp contract(off)
    constexpr int TZ = EZ - 1;  // owned element rows per tile
    static_assert(P == 4, "vector loads assume 5 nodes per element edge");
    constexpr int NP = P + 1;
    constexpr bool ANY = (MODE == MODE_ANY);
    __shared__ T xr_col[EZ][EX][NP];   // element's i = P column of contributions
    __shared__ T xr_row[EZ][EX][P];    // element's j = P row (i < P)

    // XCD-contiguous numbering of (tile, solve) pairs
    const int wgid = blockIdx.x;
    const int lin = (wgid % NXCD) * a.nper + wgid / NXCD;
    if (lin >= a.total) return;  // whole workgroup leaves together
    const int tile = lin % a.ntiles;
    const long b = lin / a.ntiles;

    const int t = threadIdx.x;
    const int exl = t % EX, ezl = t / EX;
    const int tx = tile % a.ntx, tz = tile / a.ntx;
    const int ex = tx * TX - 1 + exl;
    const int ez = tz * TZ - 1 + ezl;
    const bool valid = t < EX * EZ && ex >= 0 && ex < a.nex && ez >= 0 && ez < a.nez;
    const bool owned = valid && exl >= 1 && ezl >= 1;
    const int pitch = a.pitch;
    const int base = (P * ez) * pitch + P * ex;  // offset of node (P ez, P ex) within a solve
    // direction of this workgroup's step (uniform per workgroup)
    const bool adj = ANY ? a.slots[b].phase == SEM_SLOT_ADJOINT : MODE == MODE_ADJS;
    const bool fwd = ANY ? !adj : MODE == MODE_FWD;
    const int damp = (ANY && adj) ? a.damp_adj : a.damp;
    View<T> V;
    if (!(ANY ? (adj ? make_view<T, P, MODE_ADJS>(a, b, V) : make_view<T, P, MODE_FWD>(a, b, V))
              : make_view<T, P, MODE>(a, b, V)))
        return;

    if constexpr (MODE == MODE_REC) {
        // seismogram samples only: touch the receiver elements alone
        if (V.rec_t < 0 || !owned) return;
        const int set = V.set;
        const int beg = a.rec_ptr[set * (a.ntiles + 1) + tile];
        const int end = a.rec_ptr[set * (a.ntiles + 1) + tile + 1];
        for (int k = beg; k < end; ++k) {
            const RecEntry<T, P>& e = a.recs[k];
            if (!(e.owned && e.ex == ex && e.ez == ez)) continue;
            const GMEM T* __restrict__ uin = V.u_in + base;
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                T row = T(0);
#pragma unroll
                for (int i = 0; i < NP; ++i) row = fmadd(e.lx[i], uin[j * pitch + i], row);
                acc = fmadd(e.lz[j], row, acc);
            }
            V.seis[(long)V.rec_t * a.rec_pitch + e.r] = acc;
        }
        return;
    } else {
        T u[NP][NP], r[NP][NP];
        // update-phase operands of the owned 4 x 4 block, issued right after
        // the element math (earlier they would hold registers through it) so
        // their latency overlaps the exchange
        T wv[P][P];
        T dxv[P], dzv[P];
        {
            // every lane loads a block (elements outside the domain read a
            // clamped in-range one) so that the cross-lane fifth-node reads
            // below see their neighbours' data
            const int exc = ex < 0 ? 0 : (ex >= a.nex ? a.nex - 1 : ex);
            const int ezc = ez < 0 ? 0 : (ez >= a.nez ? a.nez - 1 : ez);
            const int base_c = (P * ezc) * pitch + P * exc;
            // lane t + 1 holds the right neighbour (same tile row and wave, inside the domain)
            const bool from_right = exl < EX - 1 && (t & 63) != 63 && ex + 1 < a.nex;
            T c2[NP][NP];
            fetch_block(V.u_in, (unsigned)base_c, (unsigned)pitch, valid && !from_right, u);
            fetch_block(V.c2, (unsigned)base_c, (unsigned)pitch, valid && !from_right, c2);
            finish_block(from_right, u);
            finish_block(from_right, c2);
            // the element math runs in every lane (branch-free: fewer live
            // registers); lanes outside the domain then contribute zero
            element_action<T, P>(a.ec, u, c2, r);
            if (!valid) {
#pragma unroll
                for (int j = 0; j < NP; ++j)
#pragma unroll
                    for (int i = 0; i < NP; ++i) {
                        r[j][i] = T(0);
                        u[j][i] = T(0);
                    }
            }
            if (owned) {
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    const GMEM T* wq = at((const GMEM T*)V.w, (unsigned)(base + j * pitch));
                    if constexpr (NTS) load4nt(wq, wv[j]);
                    else load4(wq, wv[j]);
                }
                if (damp) {
                    load4(gm(a.Dx) + P * ex, dxv);
                    load4(gm(a.Dz) + P * ez, dzv);
                }
            }
        }

        if (fwd) {
            // seismogram sample of the field being read (R u at this time level)
            if (V.rec_t >= 0 && owned) {
                const int set = V.set;
                const int beg = a.rec_ptr[set * (a.ntiles + 1) + tile];
                const int end = a.rec_ptr[set * (a.ntiles + 1) + tile + 1];
                for (int k = beg; k < end; ++k) {
                    const RecEntry<T, P>& e = a.recs[k];
                    if (e.owned && e.ex == ex && e.ez == ez) {
                        T acc = T(0);
#pragma unroll
                        for (int j = 0; j < NP; ++j) {
                            T row = T(0);
#pragma unroll
                            for (int i = 0; i < NP; ++i) row = fmadd(e.lx[i], u[j][i], row);
                            acc = fmadd(e.lz[j], row, acc);
                        }
                        V.seis[(long)V.rec_t * a.rec_pitch + e.r] = acc;
                    }
                }
            }
        }
        if (adj) {
            // adjoint source R^T res, injected into the receiver element's slots
            // (entries name in-domain elements only, so lanes outside never match)
            {
                const int set = V.set;
                const int beg = a.rec_ptr[set * (a.ntiles + 1) + tile];
                const int end = a.rec_ptr[set * (a.ntiles + 1) + tile + 1];
                for (int k = beg; k < end; ++k) {
                    const RecEntry<T, P>& e = a.recs[k];
                    if (e.ex == ex && e.ez == ez) {
                        const T res = V.resid[(long)V.res_t * a.rec_pitch + e.r] * a.invJ;
#pragma unroll
                        for (int j = 0; j < NP; ++j)
#pragma unroll
                            for (int i = 0; i < NP; ++i) r[j][i] = fmadd(-(e.lz[j] * e.lx[i]), res, r[j][i]);
                    }
                }
            }
        }

        if (t < EX * EZ) {
#pragma unroll
            for (int j = 0; j < NP; ++j) xr_col[ezl][exl][j] = r[j][P];
#pragma unroll
            for (int i = 0; i < P; ++i) xr_row[ezl][exl][i] = r[P][i];
        }
        __syncthreads();
        if (!owned) return;

        // assemble the owned nodes: own contribution + left / below / below-left
#pragma unroll
        for (int j = 0; j < NP; ++j)
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                if (i == 0) r[j][i] += xr_col[ezl][exl - 1][j];
                if (j == 0) r[j][i] += (i < P) ? xr_row[ezl - 1][exl][i] : xr_col[ezl - 1][exl][P];
                if (i == 0 && j == 0) r[j][i] += xr_col[ezl - 1][exl - 1][P];
            }

        // point source: at most one element per solve owns the source node
        int s_loc = -1;
        T sterm = T(0);
        if (fwd) {
            const int src = V.src;
            const int sz = src / a.NX, sx = src - sz * a.NX;
            const int lj = sz - P * ez, li = sx - P * ex;
            const int jm = (ez == a.nez - 1) ? P : P - 1, im = (ex == a.nex - 1) ? P : P - 1;
            if (src >= 0 && lj >= 0 && lj <= jm && li >= 0 && li <= im) {
                s_loc = lj * NP + li;
                sterm = a.ec.dt2 * (*V.wav);
            }
        }

        GMEM T* __restrict__ wp = V.w + base;
        GMEM T* __restrict__ uo = V.u_out + base;
        const bool last_x = (ex == a.nex - 1), last_z = (ez == a.nez - 1);
        // dt^2 / (Wx Wz) from the GLL weights: shared element edges carry w_0 + w_P,
        // domain edges w_0 alone
        auto inv_wx = [&](int i) -> T { return (i == 0 && ex > 0) ? a.ec.iw_shared : a.ec.iw[i]; };
        auto inv_wz = [&](int j) -> T { return (j == 0 && ez > 0) ? a.ec.iw_shared : a.ec.iw[j]; };
        auto damp_at = [&](int j, int i) -> T {
            if (!damp) return T(1);
            const T dxi = (i < P) ? dxv[i] : gm(a.Dx)[P * ex + i];
            const T dzj = (j < P) ? dzv[j] : gm(a.Dz)[P * ez + j];
            return dxi * dzj;
        };
        auto update = [&](int j, int i, T wold) -> T {  // returns w^{n+1}; leaves u^{n+1} in r[j][i]
            T ws = fmadd(-((a.ec.dt2 * inv_wz(j)) * inv_wx(i)), r[j][i], wold);  // w^n + a^n
            if (fwd) {
                if (j * NP + i == s_loc) ws += sterm;
            }
            const T Dn = damp_at(j, i);
            const T wn = Dn * ws;
            r[j][i] = fmadd(Dn, u[j][i], wn);
            return wn;
        };
#pragma unroll
        for (int j = 0; j < P; ++j) {
            T wn[4], un[4];
#pragma unroll
            for (int i = 0; i < P; ++i) {
                wn[i] = update(j, i, wv[j][i]);
                un[i] = r[j][i];
            }
            const unsigned o = (unsigned)(base + j * pitch);
            if constexpr (NTS) {
                store4nt(at(V.w, o), wn);
                store4nt(at(V.u_out, o), un);
            } else {
                store4(at(V.w, o), wn);
                store4(at(V.u_out, o), un);
            }
            if (fwd) {
                if (V.ck_u) {
                    store4(at(V.ck_w, o), wn);
                    store4(at(V.ck_u, o), un);
                }
            }
        }
        // domain edges: right column (i = P) of the last element column, top row (j = P)
        if (last_x || last_z) {
#pragma unroll
            for (int j = 0; j < NP; ++j)
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    if (!((i == P && last_x && (j < P || last_z)) || (j == P && last_z && i < P))) continue;
                    const int o = j * pitch + i;
                    const T wn = update(j, i, wp[o]);
                    wp[o] = wn;
                    uo[o] = r[j][i];
                    if (fwd) {
                        if (V.ck_u) {
                            V.ck_w[base + o] = wn;
                            V.ck_u[base + o] = r[j][i];
                        }
                    }
                }
        }
    }
}

// Two steps per launch (n -> n+1 -> n+2) on tiles of EX2 x EZ2 element
// threads: the owned block O of TX2 x TZ2 elements, two halo columns on the
// left and one on the right, two halo rows below and one above.  Step one
// runs the element action on the whole block and updates the nodes of every
// element but the first column and row (O and the ring around it: the nodes
// step two's elements touch); their u^{n+1} stay in LDS (U1), and the
// increment of O's nodes too (W1).  Step two runs the element action on O plus
// its left and lower ring, reading u^{n+1} from LDS, and updates O's nodes.
// Element math, assembly order and update are those of step_kernel, so a
// two-step launch gives the same bits as two single steps.  The halo reads
// w^n of neighbouring tiles' nodes, so the increment is read from a.w and
// written to a.w_out (two buffers).  When a.ck_u is set, (u, w)^{n+2} also go
// to the checkpoint a.ck_u / a.ck_w.  u^{n+1} of O's nodes is written to
// a.u_mid when it is not null (the forward history, the adjoint ring), with
// non-temporal stores: no launch reads it back (the gradient correlation does,
// later), so it is kept out of the last-level cache that holds the state the
// next launch reads.  When a.u_copy is set (the full-history forward),
// u^{n+2} also goes there with plain stores, the next launch reads that copy,
// and u_out (the history slot) is written non-temporally too (batch launches
// only).  FP32 and FP64.
// Per-workgroup arrays and time indices of one solve in a two-step launch:
// from the uniform batch arguments, or from the slot table (FORWARD,
// FORWARD_HIST, RECOMPUTE and ADJOINT slots; the increment alternates between
// the slot's w and w2 launch by launch, and a slot whose segment ends after
// step m takes that step alone: `single`).  Returns false if this solve does
// not advance.
template <typename T>
struct View2 {
    const GMEM T* u_in;
    GMEM T* u_out;
    GMEM T* u_mid;
    GMEM T* u_copy;
    GMEM T* ck_u;
    GMEM T* ck_w;
    const GMEM T* w_in;
    GMEM T* w_out;
    const GMEM T* c2;
    GMEM T* seis;
    const GMEM T* resid;
    const GMEM T* wav;
    int src, set, rec_t, res_t;
    bool single;
};

template <typename T, int P, int MODE>
__device__ __forceinline__ bool make_view2(const StepArgs<T, P>& a, long b, View2<T>& v) {
    v.u_copy = v.ck_u = v.ck_w = nullptr;
    v.single = false;
    if (!a.slots) {
        const long soff = b * a.sb;
        const long mdl = a.model_of ? a.model_of[b] : b;
        v.u_in = gm(a.u_in + soff);
        v.u_out = gm(a.u_out + soff);
        v.u_mid = a.u_mid ? gm(a.u_mid + soff) : nullptr;
        v.u_copy = a.u_copy ? gm(a.u_copy + soff) : nullptr;
        v.ck_u = a.ck_u ? gm(a.ck_u + soff) : nullptr;
        v.ck_w = a.ck_w ? gm(a.ck_w + soff) : nullptr;
        v.w_in = gm((const T*)a.w + soff);
        v.w_out = gm(a.w_out + soff);
        v.c2 = gm(a.c2 + mdl * a.sb);
        v.seis = a.seis ? gm(a.seis + b * a.seis_bs) : nullptr;
        v.resid = a.resid ? gm(a.resid + b * a.resid_bs) : nullptr;
        v.wav = (MODE == MODE_FWD && a.src_node) ? gm(a.wavelet + b * a.wav_bs + a.n) : nullptr;
        v.src = (MODE == MODE_FWD && a.src_node) ? a.src_node[b] : -1;
        v.set = a.rec_set_of ? a.rec_set_of[b] : 0;
        v.rec_t = a.rec_t;
        v.res_t = a.res_t;
        return true;
    }
    const sem_slot& d = a.slots[b];
    if (!((a.phase_mask >> d.phase) & 1u) || a.m >= d.len) return false;
    // a table that breaks sem_slots_step2's conditions is not advanced (the
    // increment would be read and written in one buffer, or an adjoint ring
    // of 2 slots would give u_out == u_in): no race, the slot stays put
    if (!d.w2 || d.w2 == d.w || (d.phase == SEM_SLOT_ADJOINT && d.ring_slots < 3)) return false;
    v.single = (a.m + 1 == d.len);
    const bool flip = ((a.m >> 1) & 1) != 0;
    v.w_in = gm((const T*)(flip ? d.w2 : d.w));
    v.w_out = gm((T*)(flip ? d.w : d.w2));
    v.c2 = gm((const T*)d.c2);
    v.set = d.rec_set;
    v.seis = nullptr;
    v.resid = nullptr;
    v.wav = nullptr;
    v.src = -1;
    v.rec_t = v.res_t = -1;
    if (d.phase == SEM_SLOT_FORWARD_HIST) {
        if (MODE != MODE_FWD) return false;
        const int n = d.n0 + a.m;
        const GMEM T* seg = gm((const T*)d.seg);
        v.u_in = seg + (long)a.m * d.seg_stride;
        v.u_mid = (GMEM T*)(seg + (long)(a.m + 1) * d.seg_stride);
        v.u_out = (GMEM T*)(seg + (long)(a.m + 2) * d.seg_stride);  // not touched when single
        v.wav = gm((const T*)d.wavelet + n);
        v.src = d.src_node;
        if (d.record) {
            v.seis = gm((T*)d.seis);
            v.rec_t = a.post ? n - 1 : n;
        }
        return true;
    }
    if (d.phase == SEM_SLOT_FORWARD) {
        // u^n in u1 when (n + 1) / 2 is odd, else in u0 (n even here): a pair
        // writes u^{n+2} to the other buffer, a step taken alone u^{n+1}; the
        // checkpoint receives (u, w) after the segment's last pair
        if (MODE != MODE_FWD || !d.pair_parity || (d.n0 & 1) || (d.ck_u && (d.len & 1))) return false;
        const int n = d.n0 + a.m;
        const bool hi = ((n >> 1) & 1) != 0;
        v.u_in = gm((const T*)(hi ? d.u1 : d.u0));
        v.u_out = gm((T*)(hi ? d.u0 : d.u1));
        v.u_mid = v.single ? v.u_out : nullptr;
        if (a.m + 2 == d.len && d.ck_u) {
            v.ck_u = gm((T*)d.ck_u);
            v.ck_w = gm((T*)d.ck_w);
        }
        v.wav = gm((const T*)d.wavelet + n);
        v.src = d.src_node;
        if (d.record) {
            v.seis = gm((T*)d.seis);
            v.rec_t = a.post ? n - 1 : n;
        }
        return true;
    }
    if (d.phase == SEM_SLOT_RECOMPUTE) {
        // u^{n0} from the checkpoint u0, then segment slots 1..len; no samples
        if (MODE != MODE_FWD) return false;
        const int n = d.n0 + a.m;
        const GMEM T* seg = gm((const T*)d.seg);
        v.u_in = (a.m == 0) ? gm((const T*)d.u0) : seg + (long)a.m * d.seg_stride;
        v.u_mid = (GMEM T*)(seg + (long)(a.m + 1) * d.seg_stride);
        v.u_out = (GMEM T*)(seg + (long)(a.m + 2) * d.seg_stride);  // not touched when single
        v.wav = gm((const T*)d.wavelet + n);
        v.src = d.src_node;
        return true;
    }
    if (d.phase == SEM_SLOT_ADJOINT) {
        if (MODE != MODE_ADJS) return false;
        // lambda_t in ring slot (t + 1) % ring_slots (ring_slots >= 3)
        const int it = d.n0 + d.len - 1 - a.m;
        const GMEM T* ring = gm((const T*)d.ring);
        v.u_in = ring + (long)((it + 1) % d.ring_slots) * d.seg_stride;
        v.u_mid = (GMEM T*)(ring + (long)(it % d.ring_slots) * d.seg_stride);
        v.u_out = (GMEM T*)(ring + (long)((it - 1 + d.ring_slots) % d.ring_slots) * d.seg_stride);
        v.resid = gm((const T*)d.resid);
        v.res_t = it;
        return true;
    }
    return false;
}

// Two-step tile shapes (sem_plan_set_tile2): EX x EZ element threads per
// workgroup, the owned block (EX - 3) x (EZ - 3) with two halo columns / rows
// on the left / below and one on the right / above.  Owned widths of 16 and
// 32 elements keep owned FP32 rows 256- or 512-byte aligned; 20 elements
// (shape 4) divide the 40-, 120- and 200-element meshes' rows evenly.  Shape
// 0 is the default; shape 3 needs 150 KB of LDS in FP32 (one workgroup per CU)
// and has no FP64 instance (it would exceed the CU's 160 KB).
constexpr int NTILE2 = 5;
inline constexpr int TILE2_EX[NTILE2] = {19, 19, 35, 35, 23};
inline constexpr int TILE2_EZ[NTILE2] = {26, 13, 14, 28, 11};
__host__ __device__ constexpr int tile2_threads(int ex, int ez) { return (ex * ez + 63) / 64 * 64; }

template <typename T, int P, int MODE, bool NTS, int EX2, int EZ2>
__global__ __launch_bounds__(tile2_threads(EX2, EZ2)) void step2_kernel(const StepArgs<T, P> a) {
#pragma clang fp contract(off)
    static_assert(P == 4, "vector loads assume 5 nodes per element edge");
    static_assert(MODE == MODE_FWD || MODE == MODE_ADJS || MODE == MODE_ANY, "forward or adjoint steps");
    constexpr int TX2 = EX2 - 3, TZ2 = EZ2 - 3;
    constexpr int U1_W = (4 * (EX2 - 1) + 1 + 3) / 4 * 4;  // level-(n+1) nodes of the ring + O
    constexpr int U1_H = 4 * (EZ2 - 1) + 1;
    constexpr int W1_W = (4 * TX2 + 1 + 3) / 4 * 4;        // increment of O's nodes
    constexpr int W1_H = 4 * TZ2 + 1;
    constexpr int NP = P + 1;
    constexpr bool ANY = (MODE == MODE_ANY);
    __shared__ T xr_col[EZ2][EX2][NP];
    __shared__ T xr_row[EZ2][EX2][P];
    __shared__ __attribute__((aligned(16))) T U1[U1_H][U1_W];
    __shared__ __attribute__((aligned(16))) T W1[W1_H][W1_W];

    const int wgid = blockIdx.x;
    const int lin = (wgid % NXCD) * a.nper2 + wgid / NXCD;
    if (lin >= a.total2) return;  // whole workgroup leaves together
    const int tile = lin % a.ntiles2;
    const long b = lin / a.ntiles2;

    const int t = threadIdx.x;
    // This thread's element.  Step one maps thread t to element
    // (t % EX2, t / EX2) of the whole block; step two re-maps the threads onto
    // its smaller block and reassigns these variables (the lambdas below read
    // them by reference, so they see the step's element).
    int exl = t % EX2, ezl = t / EX2;
    const int tx = tile % a.ntx2, tz = tile / a.ntx2;
    int ex = tx * TX2 - 2 + exl;
    int ez = tz * TZ2 - 2 + ezl;
    bool inb = t < EX2 * EZ2;
    bool valid = inb && ex >= 0 && ex < a.nex && ez >= 0 && ez < a.nez;
    const bool n1 = valid && exl >= 1 && ezl >= 1;                                // updates level n+1
    bool own = valid && exl >= 2 && exl <= EX2 - 2 && ezl >= 2 && ezl <= EZ2 - 2; // updates level n+2
    const int pitch = a.pitch;
    int base = (P * ez) * pitch + P * ex;
    bool last_x = (ex == a.nex - 1), last_z = (ez == a.nez - 1);

    // solve b's arrays and time indices (uniform per workgroup)
    // direction of this workgroup's steps (uniform per workgroup)
    const bool adj = ANY ? a.slots[b].phase == SEM_SLOT_ADJOINT : MODE == MODE_ADJS;
    const bool fwd = ANY ? !adj : MODE == MODE_FWD;
    const int damp = (ANY && adj) ? a.damp_adj : a.damp;
    View2<T> V;
    if (!(ANY ? (adj ? make_view2<T, P, MODE_ADJS>(a, b, V) : make_view2<T, P, MODE_FWD>(a, b, V))
              : make_view2<T, P, MODE>(a, b, V)))
        return;
    const GMEM T* __restrict__ u_in = V.u_in;
    GMEM T* __restrict__ u_out = V.u_out;
    GMEM T* __restrict__ u_mid = V.u_mid;
    GMEM T* __restrict__ u_copy = V.u_copy;
    GMEM T* __restrict__ ck_u = V.ck_u;  // checkpoint of (u, w)^{n+2}
    GMEM T* __restrict__ ck_w = V.ck_w;
    const GMEM T* __restrict__ w_in = V.w_in;
    GMEM T* __restrict__ w_out = V.w_out;
    const GMEM T* __restrict__ c2p = V.c2;
    GMEM T* __restrict__ seis = V.seis;
    const GMEM T* __restrict__ resid = V.resid;
    const GMEM T* __restrict__ wav = V.wav;
    const int src = V.src;
    const int set = V.set;
    const int rbeg = a.rec_ptr2[set * (a.ntiles2 + 1) + tile];
    const int rend = a.rec_ptr2[set * (a.ntiles2 + 1) + tile + 1];

    // every lane reads an in-range block (clamped) so that the cross-lane
    // fifth-node reads see their neighbours' data
    int exc = ex < 0 ? 0 : (ex >= a.nex ? a.nex - 1 : ex);
    int ezc = ez < 0 ? 0 : (ez >= a.nez ? a.nez - 1 : ez);
    int base_c = (P * ezc) * pitch + P * exc;
    bool from_right = exl < EX2 - 1 && (t & 63) != 63 && ex + 1 < a.nex;

    auto inv_wx = [&](int i) -> T { return (i == 0 && ex > 0) ? a.ec.iw_shared : a.ec.iw[i]; };
    auto inv_wz = [&](int j) -> T { return (j == 0 && ez > 0) ? a.ec.iw_shared : a.ec.iw[j]; };
    T dxv[P], dzv[P];
    auto damp_at = [&](int j, int i) -> T {
        if (!damp) return T(1);
        const T dxi = (i < P) ? dxv[i] : gm(a.Dx)[P * ex + i];
        const T dzj = (j < P) ? dzv[j] : gm(a.Dz)[P * ez + j];
        return dxi * dzj;
    };
    // source node of this element (owned nodes incl. domain edges), else -1
    auto source_slot = [&]() -> int {
        if (src < 0) return -1;
        const int sz = src / a.NX, sx = src - sz * a.NX;
        const int lj = sz - P * ez, li = sx - P * ex;
        const int jm = last_z ? P : P - 1, im = last_x ? P : P - 1;
        return (lj >= 0 && lj <= jm && li >= 0 && li <= im) ? lj * NP + li : -1;
    };
    // assemble the element's owned nodes: own contribution + left / below / below-left
    auto assemble = [&](T (&r)[NP][NP]) {
#pragma unroll
        for (int j = 0; j < NP; ++j)
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                if (i == 0) r[j][i] += xr_col[ezl][exl - 1][j];
                if (j == 0) r[j][i] += (i < P) ? xr_row[ezl - 1][exl][i] : xr_col[ezl - 1][exl][P];
                if (i == 0 && j == 0) r[j][i] += xr_col[ezl - 1][exl - 1][P];
            }
    };
    auto publish = [&](const T (&r)[NP][NP]) {
        if (inb) {
#pragma unroll
            for (int j = 0; j < NP; ++j) xr_col[ezl][exl][j] = r[j][P];
#pragma unroll
            for (int i = 0; i < P; ++i) xr_row[ezl][exl][i] = r[P][i];
        }
    };
    // seismogram sample R u of this element's block at time index rt
    auto sample = [&](const T (&u)[NP][NP], int rt) {
        if (rt < 0 || !seis) return;
        for (int k = rbeg; k < rend; ++k) {
            const RecEntry<T, P>& e = a.recs2[k];
            if (e.owned && e.ex == ex && e.ez == ez) {
                T acc = T(0);
#pragma unroll
                for (int j = 0; j < NP; ++j) {
                    T row = T(0);
#pragma unroll
                    for (int i = 0; i < NP; ++i) row = fmadd(e.lx[i], u[j][i], row);
                    acc = fmadd(e.lz[j], row, acc);
                }
                seis[(long)rt * a.rec_pitch + e.r] = acc;
            }
        }
    };
    // adjoint source R^T res at time index rt into the element's contributions
    auto inject = [&](T (&r)[NP][NP], int rt) {
        for (int k = rbeg; k < rend; ++k) {
            const RecEntry<T, P>& e = a.recs2[k];
            if (e.ex == ex && e.ez == ez) {
                const T res = resid[(long)rt * a.rec_pitch + e.r] * a.invJ;
#pragma unroll
                for (int j = 0; j < NP; ++j)
#pragma unroll
                    for (int i = 0; i < NP; ++i) r[j][i] = fmadd(-(e.lz[j] * e.lx[i]), res, r[j][i]);
            }
        }
    };

    // ---- step one: level n -> n+1 on the whole block ----
    {
        T u[NP][NP], r[NP][NP];
        T wv[P][P];
        {
            T c2[NP][NP];
            fetch_block(u_in, (unsigned)base_c, (unsigned)pitch, valid && !from_right, u);
            fetch_block(c2p, (unsigned)base_c, (unsigned)pitch, valid && !from_right, c2);
            finish_block(from_right, u);
            finish_block(from_right, c2);
            element_action<T, P>(a.ec, u, c2, r);
        }
        if (!valid) {
#pragma unroll
            for (int j = 0; j < NP; ++j)
#pragma unroll
                for (int i = 0; i < NP; ++i) r[j][i] = T(0);
        }
        if (fwd) {
            if (own) sample(u, V.rec_t);
        }
        if (n1) {
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const GMEM T* wq = at(w_in, (unsigned)(base + j * pitch));
                if constexpr (NTS) load4nt(wq, wv[j]);
                else load4(wq, wv[j]);
            }
            if (damp) {
                load4(gm(a.Dx) + P * ex, dxv);
                load4(gm(a.Dz) + P * ez, dzv);
            }
        }
        if (adj) {
            if (valid) inject(r, V.res_t);
        }
        publish(r);
        __syncthreads();
        if (n1) {
            assemble(r);
            const int s_loc = source_slot();
            const T sterm = (s_loc >= 0) ? a.ec.dt2 * wav[0] : T(0);
            auto update = [&](int j, int i, T wold, T uold) -> T {  // w^{n+1}; u^{n+1} left in r[j][i]
                T ws = fmadd(-((a.ec.dt2 * inv_wz(j)) * inv_wx(i)), r[j][i], wold);  // w^n + a^n
                if (fwd) {
                    if (j * NP + i == s_loc) ws += sterm;
                }
                const T Dn = damp_at(j, i);
                const T wn = Dn * ws;
                r[j][i] = fmadd(Dn, uold, wn);
                return wn;
            };
            const int X0 = P * (exl - 1), Z0 = P * (ezl - 1);     // in U1
            const int X0w = P * (exl - 2), Z0w = P * (ezl - 2);   // in W1 (own only)
#pragma unroll
            for (int j = 0; j < P; ++j) {
                T wn[4], un[4];
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    wn[i] = update(j, i, wv[j][i], u[j][i]);
                    un[i] = r[j][i];
                }
#pragma unroll
                for (int i = 0; i < P; ++i) U1[Z0 + j][X0 + i] = un[i];
                if (own) {
#pragma unroll
                    for (int i = 0; i < P; ++i) W1[Z0w + j][X0w + i] = wn[i];
                    if (u_mid) store4nt(at(u_mid, (unsigned)(base + j * pitch)), un);
                    if (V.single) store4(at(w_out, (unsigned)(base + j * pitch)), wn);
                }
            }
            if (last_x || last_z) {
#pragma unroll
                for (int j = 0; j < NP; ++j)
#pragma unroll
                    for (int i = 0; i < NP; ++i) {
                        if (!((i == P && last_x && (j < P || last_z)) || (j == P && last_z && i < P))) continue;
                        const int o = j * pitch + i;
                        const T wn = update(j, i, w_in[base + o], u[j][i]);
                        U1[Z0 + j][X0 + i] = r[j][i];
                        if (own) {
                            W1[Z0w + j][X0w + i] = wn;
                            if (u_mid) __builtin_nontemporal_store(r[j][i], &u_mid[base + o]);
                            if (V.single) w_out[base + o] = wn;
                        }
                    }
            }
        }
        __syncthreads();
    }
    // a slot whose segment ends after this step: u^{n+1} went to u_mid, w^{n+1}
    // to w_out (uniform per workgroup)
    if (V.single) return;

    // ---- step two: level n+1 -> n+2 on O and its left / lower ring ----
    // The threads are re-mapped onto the step-two block of (EX2 - 2) x
    // (EZ2 - 2) elements, element (xl, zl) in thread xl + (EX2 - 2) zl, so that
    // the threads past it form whole waves that skip the element math (one of
    // the four waves of the 19 x 13 tile).  The lanes past the block in the
    // last partly used wave take a clamped in-range block and publish nothing.
    {
        constexpr int XW = EX2 - 2, ZW = EZ2 - 2, NS2 = XW * ZW;
        const bool in2 = t < NS2;
        const int xl = t % XW, zl = in2 ? t / XW : ZW - 1;
        exl = xl + 1;
        ezl = zl + 1;
        ex = tx * TX2 - 1 + xl;
        ez = tz * TZ2 - 1 + zl;
        inb = in2;
        valid = in2 && ex >= 0 && ex < a.nex && ez >= 0 && ez < a.nez;
        own = valid && xl >= 1 && zl >= 1;
        base = (P * ez) * pitch + P * ex;
        last_x = (ex == a.nex - 1);
        last_z = (ez == a.nez - 1);
        exc = ex < 0 ? 0 : (ex >= a.nex ? a.nex - 1 : ex);
        ezc = ez < 0 ? 0 : (ez >= a.nez ? a.nez - 1 : ez);
        base_c = (P * ezc) * pitch + P * exc;
        from_right = xl < XW - 1 && (t & 63) != 63 && ex + 1 < a.nex;
        const bool wave_on = (t & ~63) < NS2;  // uniform per wave

        T u[NP][NP], r[NP][NP];
        if (wave_on) {
            // u^{n+1} block from LDS, c^2 again from memory (L2)
#pragma unroll
            for (int j = 0; j < NP; ++j)
#pragma unroll
                for (int i = 0; i < NP; ++i) u[j][i] = U1[P * zl + j][P * xl + i];
            {
                T c2[NP][NP];
                fetch_block(c2p, (unsigned)base_c, (unsigned)pitch, valid && !from_right, c2);
                finish_block(from_right, c2);
                element_action<T, P>(a.ec, u, c2, r);
            }
            if (!valid) {
#pragma unroll
                for (int j = 0; j < NP; ++j)
#pragma unroll
                    for (int i = 0; i < NP; ++i) r[j][i] = T(0);
            }
            if (fwd) {
                if (own) sample(u, V.rec_t + 1);
            }
            if (own && damp) {  // sponge factors of this step's element
                load4(gm(a.Dx) + P * ex, dxv);
                load4(gm(a.Dz) + P * ez, dzv);
            }
            if (adj) {
                if (valid) inject(r, V.res_t - 1);
            }
            publish(r);
        }
        __syncthreads();
        if (!own) return;
        assemble(r);
        const int s_loc = source_slot();
        const T sterm = (s_loc >= 0) ? a.ec.dt2 * wav[1] : T(0);
        const int X0w = P * (exl - 2), Z0w = P * (ezl - 2);
        auto update = [&](int j, int i, T wold, T uold) -> T {
            T ws = fmadd(-((a.ec.dt2 * inv_wz(j)) * inv_wx(i)), r[j][i], wold);  // w^n + a^n
            if (fwd) {
                if (j * NP + i == s_loc) ws += sterm;
            }
            const T Dn = damp_at(j, i);
            const T wn = Dn * ws;
            r[j][i] = fmadd(Dn, uold, wn);
            return wn;
        };
#pragma unroll
        for (int j = 0; j < P; ++j) {
            T wn[4], un[4];
#pragma unroll
            for (int i = 0; i < P; ++i) {
                wn[i] = update(j, i, W1[Z0w + j][X0w + i], u[j][i]);
                un[i] = r[j][i];
            }
            const unsigned o = (unsigned)(base + j * pitch);
            if constexpr (NTS) {
                store4nt(at(w_out, o), wn);
                store4nt(at(u_out, o), un);
                if (u_copy) store4nt(at(u_copy, o), un);
            } else {
                store4(at(w_out, o), wn);
                if (u_copy) {
                    store4nt(at(u_out, o), un);
                    store4(at(u_copy, o), un);
                } else {
                    store4(at(u_out, o), un);
                }
            }
            if (ck_u) {
                store4(at(ck_w, o), wn);
                store4(at(ck_u, o), un);
            }
        }
        if (last_x || last_z) {
#pragma unroll
            for (int j = 0; j < NP; ++j)
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    if (!((i == P && last_x && (j < P || last_z)) || (j == P && last_z && i < P))) continue;
                    const int o = j * pitch + i;
                    const T wn = update(j, i, W1[Z0w + j][X0w + i], u[j][i]);
                    w_out[base + o] = wn;
                    u_out[base + o] = r[j][i];
                    if (u_copy) u_copy[base + o] = r[j][i];
                    if (ck_u) {
                        ck_w[base + o] = wn;
                        ck_u[base + o] = r[j][i];
                    }
                }
        }
    }
}

// Element-local gradient correlation of (uf, l) added node by node into FP64
// running sums: g(j, i) = gx(j, i) (D_x uf)(D_x l) + gz(j, i) (D_z uf)(D_z l)
// at node (j, i), with no 5 x 5 temporary held across the element.
template <typename T, int P>
__device__ __forceinline__ void element_corr_add(const ElemConst<T, P>& ec, const T (&l)[P + 1][P + 1],
                                                 const T (&uf)[P + 1][P + 1], double (&acc)[P + 1][P + 1]) {
    static_assert(P == 4, "apply5 is the 5-node operator");
    constexpr int NP = P + 1;
    T gxp[NP][NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        T dl[NP], du[NP];
        apply5(ec.fD, l[j][0], l[j][1], l[j][2], l[j][3], l[j][4], dl);
        apply5(ec.fD, uf[j][0], uf[j][1], uf[j][2], uf[j][3], uf[j][4], du);
#pragma unroll
        for (int i = 0; i < NP; ++i) gxp[j][i] = ec.gx[j][i] * (du[i] * dl[i]);
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        T dl[NP], du[NP];
        apply5(ec.fD, l[0][i], l[1][i], l[2][i], l[3][i], l[4][i], dl);
        apply5(ec.fD, uf[0][i], uf[1][i], uf[2][i], uf[3][i], uf[4][i], du);
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            T g = gxp[j][i];
            g += ec.gz[j][i] * (du[j] * dl[j]);
            acc[j][i] += (double)g;
        }
    }
}

// Correlation workgroups: one wave of CORR_W consecutive elements in
// row-major element order (rows wrap inside a wave), no halo.
constexpr int CORR_W = 64;

// Gradient accumulator of one solve: `sb` doubles in the padded nodal layout
// (each element's 16 owned nodes, node (4 ez + j, 4 ex + i), i, j < 4),
// followed by 9 element planes of `gplane` = nez * nexp doubles (element
// (ex, ez) at ez * nexp + ex): planes 0..4 hold the element's contributions to
// its right-column nodes (i = 4, j = 0..4), planes 5..8 those to its top-row
// nodes (j = 4, i = 0..3).  grad_finalize_kernel adds the (up to 4) element
// contributions of each node in a fixed order.
template <typename T, int P>
struct CorrArgs {
    ElemConst<T, P> ec;
    int nex, nez, pitch;
    int ctiles, total, nper;       // correlation workgroups per solve; grid
    long sb;                       // elements per nodal array of one solve
    long gsb, gplane;              // accumulator doubles per solve, per element plane
    int nexp;                      // element-plane row length
    const T* uf;      // u^{c0+1} of solve 0; u^{it+1} at uf + (it - c0) * ustep
    long ustep;
    const T* ring;    // adjoint states: lambda_it in slot (it + 1) % ring_slots
    int ring_slots;
    long rstep;       // elements between ring slots
    int c0, clen;     // chunk of adjoint steps it = c0 + clen - 1 .. c0
    double* G;
    // slot mode: per-solve buffers from the slot table; chunk `chunk` of
    // `chunk_len` steps of each ADJOINT slot's current segment
    const sem_slot* slots;
    int chunk, chunk_len;
};

// Gradient correlation of one chunk of adjoint steps, run after the chunk's
// adjoint steps: for it = c0 + clen - 1 down to c0 each element adds its 25
// nodal contributions corr(u^{it+1}, lambda_it) to FP64 running sums that
// start from the accumulator and are stored back at the end.  Every
// contribution has a sum of its own (assembly happens once, in
// grad_finalize_kernel), so each is the same sequence of FP64 additions
// however the steps are chunked (checkpoint spacing, full history, slots),
// and no step needs an exchange between threads: no LDS, no barrier, no halo
// elements.  Without a software pipeline the kernel holds 146 VGPRs (3 waves
// per SIMD), which hides the load latency as well as prefetching the next
// step's blocks at 2 waves per SIMD does.
template <typename T, int P>
__global__ __launch_bounds__(CORR_W) void corr_kernel(const CorrArgs<T, P> a) {
    static_assert(P == 4, "vector loads assume 5 nodes per element edge");
    constexpr int NP = P + 1;
    const int wgid = blockIdx.x;
    const int lin = (wgid % NXCD) * a.nper + wgid / NXCD;
    if (lin >= a.total) return;  // whole workgroup leaves together
    const int tile = lin % a.ctiles;
    const long b = lin / a.ctiles;
    const int t = threadIdx.x;
    // lanes past the last element work on it (clamped) and store nothing
    const int ne = a.nex * a.nez;
    const int e = tile * CORR_W + t;
    const bool valid = e < ne;
    const int ex = (valid ? e : ne - 1) % a.nex;
    const int ez = (valid ? e : ne - 1) / a.nex;
    const int pitch = a.pitch;
    // this solve's chunk: forward states u^{c0+1..}, adjoint ring, accumulator
    const T* ufb;
    const T* ringb;
    double* Gb;
    long ustep, rstep;
    int c0, clen, ring_slots;
    if (!a.slots) {
        const long soff = b * a.sb;
        ufb = a.uf + soff;
        ringb = a.ring + soff;
        Gb = a.G + b * a.gsb;
        ustep = a.ustep;
        rstep = a.rstep;
        c0 = a.c0;
        clen = a.clen;
        ring_slots = a.ring_slots;
    } else {
        const sem_slot& d = a.slots[b];
        const int m_lo = a.chunk * a.chunk_len;
        if (d.phase != SEM_SLOT_ADJOINT || m_lo >= d.len) return;  // uniform per workgroup
        const int m_hi = m_lo + a.chunk_len < d.len ? m_lo + a.chunk_len : d.len;
        // segment steps m_lo .. m_hi-1 are adjoint steps it = n0+len-1-m
        c0 = d.n0 + d.len - m_hi;
        clen = m_hi - m_lo;
        ustep = rstep = d.seg_stride;
        ufb = (const T*)d.seg + (long)(c0 + 1 - d.n0) * d.seg_stride;
        ringb = (const T*)d.ring;
        Gb = d.G;
        ring_slots = d.ring_slots;
    }
    // every lane loads a block so that the cross-lane fifth-node reads see
    // their neighbours' data; lane + 1 holds the right neighbour unless the
    // row ends there
    const int base = (P * ez) * pitch + P * ex;
    const bool from_right = t < CORR_W - 1 && ex + 1 < a.nex;
    GMEM double* __restrict__ Gp = gm(Gb) + base;
    GMEM double* __restrict__ Ep = gm(Gb) + a.sb + (long)ez * a.nexp + ex;

    double acc[NP][NP];
    if (valid) {
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const vd2 lo = __builtin_nontemporal_load(reinterpret_cast<const GMEM vd2*>(Gp + j * pitch));
            const vd2 hi = __builtin_nontemporal_load(reinterpret_cast<const GMEM vd2*>(Gp + j * pitch + 2));
            acc[j][0] = lo.x;
            acc[j][1] = lo.y;
            acc[j][2] = hi.x;
            acc[j][3] = hi.y;
        }
#pragma unroll
        for (int j = 0; j < NP; ++j) acc[j][P] = __builtin_nontemporal_load(Ep + j * a.gplane);
#pragma unroll
        for (int i = 0; i < P; ++i) acc[P][i] = __builtin_nontemporal_load(Ep + (NP + i) * a.gplane);
    } else {
#pragma unroll
        for (int j = 0; j < NP; ++j)
#pragma unroll
            for (int i = 0; i < NP; ++i) acc[j][i] = 0.0;
    }
    T lc[NP][NP], uc[NP][NP];
    auto fetch = [&](int s, T (&l)[NP][NP], T (&u)[NP][NP]) {
        const int it = c0 + s;
        const GMEM T* __restrict__ up = gm(ufb + (long)s * ustep);
        const GMEM T* __restrict__ lp = gm(ringb + (long)((it + 1) % ring_slots) * rstep);
        fetch_block(lp, (unsigned)base, (unsigned)pitch, !from_right, l);
        fetch_block(up, (unsigned)base, (unsigned)pitch, !from_right, u);
    };
    for (int s = clen - 1; s >= 0; --s) {
        fetch(s, lc, uc);
        finish_block(from_right, lc);
        finish_block(from_right, uc);
        element_corr_add<T, P>(a.ec, lc, uc, acc);
    }
    if (valid) {
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const vd2 lo = {acc[j][0], acc[j][1]}, hi = {acc[j][2], acc[j][3]};
            __builtin_nontemporal_store(lo, reinterpret_cast<GMEM vd2*>(Gp + j * pitch));
            __builtin_nontemporal_store(hi, reinterpret_cast<GMEM vd2*>(Gp + j * pitch + 2));
        }
#pragma unroll
        for (int j = 0; j < NP; ++j) __builtin_nontemporal_store(acc[j][P], Ep + j * a.gplane);
#pragma unroll
        for (int i = 0; i < P; ++i) __builtin_nontemporal_store(acc[P][i], Ep + (NP + i) * a.gplane);
    }
}

// Copy [nmodel][N] (row length NX) into the padded [nmodel][NZ][pitch] layout.
template <typename T>
__global__ void pad_rows_kernel(const T* src, T* dst, int NX, int NZ, int pitch, long sb) {
    const long m = blockIdx.y;
    const long N = (long)NX * NZ;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        const long jz = k / NX, ix = k - jz * NX;
        dst[m * sb + jz * pitch + ix] = src[m * N + k];
    }
}

// r = seis - obs (first nrec columns of each row), misfit[b] = 0.5 * sum r^2.
// The observations are double and the difference is formed in double (the
// reference forms d_syn - y_obs in FP64, core.py:1224-1227): near a fit, where
// |r| << |seis|, an FP32 difference would lose the residual's low digits.  The
// adjoint source res keeps the state precision; the misfit sums the double r.
// Rows outside the misfit window [w0, w1) or off its decimation grid give r = 0.
// nrec_of: per-solve receiver counts on the device, or null to use nrec.
// One block of RESID_THREADS per solve: each wave walks rows (wave, wave +
// nwaves, ...), its lanes the columns; per-thread FP64 sums are combined in a
// fixed tree, so the misfit is deterministic.
constexpr int RESID_THREADS = 1024;
// The body for one solve (seis/obs/res at its first row); red: RESID_THREADS
// doubles of LDS.  Called by the whole block.
template <typename T>
__device__ __forceinline__ void residual_rows(const T* seis, const double* obs, T* res, double* misfit, int nrec,
                                              int nt, int pitch, int w0, int w1, int decim, double* red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = RESID_THREADS / 64;
    double acc = 0.0;
    for (int row = wave; row < nt; row += nwaves) {
        const bool in = row >= w0 && row < w1 && (row - w0) % decim == 0;
        const long o = (long)row * pitch;
        for (int c = lane; c < pitch; c += 64) {
            const double rv = (c < nrec && in) ? (double)seis[o + c] - obs[o + c] : 0.0;
            res[o + c] = (T)rv;
            acc += rv * rv;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = RESID_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *misfit = 0.5 * red[0];
}

template <typename T>
__global__ __launch_bounds__(RESID_THREADS) void residual_kernel(const T* seis, const double* obs, T* res,
                                                                 double* misfit, const int* nrec_of, int nrec,
                                                                 int nt, int pitch, long bs, int w0, int w1,
                                                                 int decim) {
    const long b = blockIdx.x;
    if (nrec_of) nrec = nrec_of[b];
    __shared__ double red[RESID_THREADS];
    residual_rows<T>(seis + b * bs, obs + b * bs, res + b * bs, misfit + b, nrec, nt, pitch, w0, w1, decim, red);
}

// Wavefield snapshot: box [iz0, iz0+nzr) x [ix0, ix0+nxr) of every solve's
// padded state into out[b][nzr][nxr] (out_bs elements between solves).
template <typename T>
__global__ void snapshot_kernel(const T* state, long sb, int pitch, int ix0, int nxr, int iz0, int nzr, T* out,
                                long out_bs) {
    const long b = blockIdx.y;
    const long n = (long)nxr * nzr;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const long jz = k / nxr, ix = k - jz * nxr;
        out[b * out_bs + k] = state[b * sb + (iz0 + jz) * pitch + ix0 + ix];
    }
}

// Regular-grid background onto the nodes (coordinates clipped into the grid):
// cell i = largest index with g[i] <= x, limited to n-2; bilinear weights, or
// the nearer grid value with ties to the lower index.
__device__ inline int grid_cell(const double* g, int n, double x, double& f) {
    int lo = 0, hi = n - 1;  // g[lo] <= x, invariant after clipping
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (g[mid] <= x) lo = mid; else hi = mid;
    }
    if (lo > n - 2) lo = n - 2;
    f = (x - g[lo]) / (g[lo + 1] - g[lo]);
    return lo;
}

__global__ void grid_to_nodes_kernel(const double* xg, int nxg, const double* zg, int nzg, const double* vp,
                                     int method, const double* x1d, const double* z1d, int NX, int NZ,
                                     double* out) {
    const long N = (long)NX * NZ;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        const double x = fmin(fmax(x1d[k % NX], xg[0]), xg[nxg - 1]);
        const double z = fmin(fmax(z1d[k / NX], zg[0]), zg[nzg - 1]);
        double fx, fz;
        const int i = grid_cell(xg, nxg, x, fx);
        const int j = grid_cell(zg, nzg, z, fz);
        const double* r0 = vp + (long)j * nxg;
        const double* r1 = r0 + nxg;
        if (method == SEM_INTERP_NEAREST) {
            const int ii = fx <= 0.5 ? i : i + 1, jj = fz <= 0.5 ? j : j + 1;
            out[k] = vp[(long)jj * nxg + ii];
        } else {
            out[k] = (1.0 - fz) * ((1.0 - fx) * r0[i] + fx * r0[i + 1]) + fz * ((1.0 - fx) * r1[i] + fx * r1[i + 1]);
        }
    }
}

// flag[0] = 1 if any of the n values is NaN or infinite (plain stores of the
// same value: no atomics needed).
template <typename T>
__global__ void finite_check_kernel(const T* p, long n, int* flag) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        if (!isfinite(p[k])) flag[0] = 1;
}

template <typename T>
__global__ void fill_kernel(T* p, T v, long n) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) p[k] = v;
}

// Assembled sensitivity of node (ix, jz) from the gradient accumulator (layout
// at CorrArgs): the owning element's own sum, then the left, lower and
// lower-left elements' sums for the node, in that order.
__device__ __forceinline__ double assemble_node(const double* Gb, long sb, long gplane, int nexp, int nex, int nez,
                                                int pitch, int ix, int jz) {
    const double* E = Gb + sb;
    const int ex = ix >> 2, i = ix & 3, ez = jz >> 2, j = jz & 3;
    const bool inx = ex < nex, inz = ez < nez;
    double s = (inx && inz) ? Gb[(long)jz * pitch + ix] : 0.0;
    if (i == 0 && ex > 0 && inz) s += E[j * gplane + (long)ez * nexp + (ex - 1)];
    if (j == 0 && ez > 0 && inx) s += E[(5 + i) * gplane + (long)(ez - 1) * nexp + ex];
    if (i == 0 && j == 0 && ex > 0 && ez > 0) s += E[4 * gplane + (long)(ez - 1) * nexp + (ex - 1)];
    return s;
}

// grad_c2 = scale * G ; grad_c = 2 sqrt(c2) grad_c2   (G assembled from the
// accumulator, c2 padded, outputs [b][N])
template <typename T>
__global__ void grad_finalize_kernel(const double* G, long gsb, long sb, long gplane, int nexp, int nex, int nez,
                                     int NX, int NZ, int pitch, const T* c2, const int* model_of, double scale,
                                     double* grad_c, double* grad_c2) {
    const long b = blockIdx.y;
    const long m = model_of ? model_of[b] : b;
    const long N = (long)NX * NZ;
    const double* Gb = G + b * gsb;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        const int jz = (int)(k / NX), ix = (int)(k - (long)jz * NX);
        const double g2 = scale * assemble_node(Gb, sb, gplane, nexp, nex, nez, pitch, ix, jz);
        if (grad_c2) grad_c2[b * N + k] = g2;
        if (grad_c) grad_c[b * N + k] = 2.0 * sqrt((double)c2[m * sb + (long)jz * pitch + ix]) * g2;
    }
}

// c[m, jz, ix] = base + sum_p sum_q Bz[jz, p] Bx[ix, q] coef[m, p, q]
// using the <= 4 non-zero cubic basis values per node and axis.
template <typename T>
__global__ void bspline_map_kernel(const double* coef, int nbz, int nbx, const double* base,
                                   const int* bz0, const double* bzw, const int* bx0, const double* bxw,
                                   int NX, int NZ, double* c, T* c2) {
    const long m = blockIdx.y;
    const long N = (long)NX * NZ;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
        const int jz = (int)(k / NX), ix = (int)(k % NX);
        const int p0 = bz0[jz], q0 = bx0[ix];
        const double* C = coef + m * nbz * nbx;
        double v = base ? base[k] : 0.0;
        for (int pp = 0; pp < 4; ++pp) {
            const int p = p0 + pp;
            if (p >= nbz) break;
            double row = 0.0;
            for (int qq = 0; qq < 4; ++qq) {
                const int q = q0 + qq;
                if (q >= nbx) break;
                row += bxw[ix * 4 + qq] * C[p * nbx + q];
            }
            v += bzw[jz * 4 + pp] * row;
        }
        if (c) c[m * N + k] = v;
        c2[m * N + k] = (T)(v * v);
    }
}

// stage 1 of g_coef = Bz^T g Bx:  t[m, jz, q] = sum_ix g[m, jz, ix] Bx[ix, q]
// one workgroup per (model, row): the row is read once, coalesced, into LDS.
__global__ __launch_bounds__(256) void bspline_proj_x_kernel(const double* g, int NX, int NZ, int nbx,
                                                             const int* bx0, const double* bxw, double* tmp) {
    extern __shared__ double acc_sh[];  // [256][nbx]
    const long m = blockIdx.y;
    const int jz = blockIdx.x;
    const double* gr = g + m * (long)NX * NZ + (long)jz * NX;
    double* mine = acc_sh + threadIdx.x * nbx;
    for (int q = 0; q < nbx; ++q) mine[q] = 0.0;
    for (int ix = threadIdx.x; ix < NX; ix += blockDim.x) {
        const double v = gr[ix];
        const int q0 = bx0[ix];
        for (int o = 0; o < 4; ++o)
            if (q0 + o < nbx) mine[q0 + o] += v * bxw[ix * 4 + o];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nbx; q += blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < (int)blockDim.x; ++k) s += acc_sh[k * nbx + q];
        tmp[(m * NZ + jz) * nbx + q] = s;
    }
}

// stage 2: g_coef[m, p, q] = sum_jz Bz[jz, p] t[m, jz, q]
__global__ void bspline_proj_z_kernel(const double* tmp, int NZ, int nbz, int nbx, const int* plo, const int* phi,
                                      const int* bz0, const double* bzw, double* gcoef) {
    const long m = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbz * nbx) return;
    const int p = k / nbx, q = k % nbx;
    double s = 0.0;
    for (int jz = plo[p]; jz < phi[p]; ++jz) {
        const int o = p - bz0[jz];
        if (o >= 0 && o < 4) s += bzw[jz * 4 + o] * tmp[m * (long)NZ * nbx + (long)jz * nbx + q];
    }
    gcoef[m * nbz * nbx + k] = s;
}

}  // namespace sem
