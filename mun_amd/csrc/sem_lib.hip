// Host side of libmun_sem.so: plan set-up (FP64 tables), receiver sets,
// workspace layout and the forward / forward+adjoint orchestration.
// The C ABI is declared in include/sem_capi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/sem_capi.h"
#include "sem_curve.h"
#include "sem_kernels.h"
#include "sem_loop.h"

using namespace sem;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(SEM_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Function-try-block tail of every C entry point: no C++ exception (host
// allocation, thread start) leaves the C ABI.
#define ABI_CATCH                                                                  \
    catch (const std::bad_alloc&) {                                                \
        return fail(SEM_E_NOMEM, "host allocation failed");                        \
    }                                                                              \
    catch (const std::exception& e_) {                                             \
        return fail(SEM_E_DEVICE, std::string("host error: ") + e_.what());        \
    }                                                                              \
    catch (...) {                                                                  \
        return fail(SEM_E_DEVICE, "host error");                                   \
    }

#define LAUNCH_CHECK()                                                                      \
    do {                                                                                    \
        hipError_t e_ = hipGetLastError();                                                  \
        if (e_ != hipSuccess)                                                               \
            return fail(SEM_E_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------- 1-D GLL algebra
// P_n(x), P_n'(x), P_n''(x) by the three-term recurrence (interior points).
void legendre(int n, double x, double& p, double& dp, double& d2p) {
    double p0 = 1.0, p1 = x;
    if (n == 0) {
        p = 1.0;
        dp = d2p = 0.0;
        return;
    }
    for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    p = p1;
    dp = n * (p0 - x * p1) / (1.0 - x * x);
    d2p = (2.0 * x * dp - n * (n + 1.0) * p) / (1.0 - x * x);
}

// GLL nodes: +-1 and the roots of P_p', Newton from Chebyshev-Lobatto guesses
// (the scheme of Bspline_Coupling/SEM/sem_waveform/utils.py:5-21).
std::vector<double> gll_nodes(int p) {
    std::vector<double> x(p + 1);
    for (int k = 0; k <= p; ++k) x[k] = -std::cos(M_PI * k / p);
    for (int k = 1; k < p; ++k) {
        double xi = x[k];
        for (int it = 0; it < 100; ++it) {
            double P, dP, d2P;
            legendre(p, xi, P, dP, d2P);
            double step = dP / d2P;
            xi -= step;
            if (std::fabs(step) < 1e-16) break;
        }
        x[k] = xi;
    }
    x[0] = -1.0;
    x[p] = 1.0;
    return x;
}

std::vector<double> gll_weights(const std::vector<double>& x) {
    int p = (int)x.size() - 1;
    std::vector<double> w(p + 1);
    for (int i = 0; i <= p; ++i) {
        double P0 = 1.0, P1 = x[i];
        for (int k = 2; k <= p; ++k) {
            double P2 = ((2.0 * k - 1.0) * x[i] * P1 - (k - 1.0) * P0) / k;
            P0 = P1;
            P1 = P2;
        }
        w[i] = 2.0 / (p * (p + 1.0) * P1 * P1);
    }
    return w;
}

// D[i][j] = l_j'(x_i) from barycentric weights, negative-sum diagonal
// (utils.py:39-61).
std::vector<double> diff_matrix(const std::vector<double>& x) {
    int n = (int)x.size();
    std::vector<double> bw(n, 1.0), D(n * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k)
            if (k != j) bw[j] /= (x[j] - x[k]);
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j)
            if (i != j) {
                D[i * n + j] = bw[j] / (bw[i] * (x[i] - x[j]));
                s += D[i * n + j];
            }
        D[i * n + i] = -s;
    }
    return D;
}

// Lagrange polynomial with +eps in each denominator (utils.py:31-37).
double lagrange(double x, int i, const std::vector<double>& nodes, double eps) {
    double r = 1.0;
    for (int j = 0; j < (int)nodes.size(); ++j)
        if (j != i) r *= (x - nodes[j]) / (nodes[i] - nodes[j] + eps);
    return r;
}

// Clamped uniform cubic B-spline on [a, b] with n control values: for each
// coordinate, the first non-zero basis index and up to 4 values.
void bspline_tables(const std::vector<double>& coord, double a, double b, int n, std::vector<int>& first,
                    std::vector<double>& vals) {
    const int k = 3;
    std::vector<double> t;
    for (int i = 0; i < k; ++i) t.push_back(a);
    for (int i = 0; i <= n - k; ++i) t.push_back(a + (b - a) * i / (n - k));
    for (int i = 0; i < k; ++i) t.push_back(b);
    const int m = (int)coord.size();
    first.assign(m, 0);
    vals.assign(4 * m, 0.0);
    for (int c = 0; c < m; ++c) {
        double x = std::min(std::max(coord[c], a), b);
        // knot span s with t[s] <= x < t[s+1], right end in the last span
        int s = k;
        while (s < n - 1 && x >= t[s + 1]) ++s;
        // Cox-de Boor triangle for degree k at span s
        double N[4] = {1.0, 0.0, 0.0, 0.0}, left[4], right[4];
        for (int d = 1; d <= k; ++d) {
            left[d] = x - t[s + 1 - d];
            right[d] = t[s + d] - x;
            double saved = 0.0;
            for (int r = 0; r < d; ++r) {
                double tmp = N[r] / (right[r + 1] + left[d - r]);
                N[r] = saved + right[r + 1] * tmp;
                saved = left[d - r] * tmp;
            }
            N[d] = saved;
        }
        first[c] = s - k;
        for (int r = 0; r < 4; ++r) vals[4 * c + r] = N[r];
    }
}

// ---------------------------------------------------------------- device buffers
struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int upload(const void* src, size_t bytes) {
        if (bytes > n) {
            if (p) (void)hipFree(p);
            p = nullptr;
            n = 0;
            hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 16));
            if (e != hipSuccess) return fail(SEM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
            n = bytes;
        }
        if (bytes) {
            hipError_t e = hipMemcpy(p, src, bytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) return fail(SEM_E_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
        }
        return SEM_OK;
    }
    // grow to at least `bytes` (contents are not preserved); returns 1 if reallocated
    int reserve(size_t bytes, bool& moved) {
        moved = false;
        if (bytes <= n) return SEM_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return fail(SEM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        n = bytes;
        moved = true;
        return SEM_OK;
    }
    int write_at(size_t off, const void* src, size_t bytes) {
        if (!bytes) return SEM_OK;
        hipError_t e = hipMemcpy((char*)p + off, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(SEM_E_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
        return SEM_OK;
    }
};

struct RecHost {
    int ex, ez, r, owned;
    double lx[5], lz[5];
};

struct BsplineTables {
    DevBuf bz0, bzw, bx0, bxw, plo, phi, qlo, qhi;
};

}  // namespace

struct sem_plan {
    sem_desc d;
    int P = 4, NX = 0, NZ = 0;
    long N = 0, sb = 0;
    int pitch = 0;
    double dx = 0, dz = 0, J = 0;
    int ntx = 0, ntz = 0, ntiles = 0;
    int ez = EZ_DEFAULT;  // tile height in element rows incl. halo (15, 7 or 3)
    int tz = EZ_DEFAULT - 1;
    // gradient correlation: ctiles workgroups of CORR_W elements per solve;
    // accumulator of gsb doubles per solve = sb nodal + 9 element planes of
    // gplane doubles
    int ctiles = 0, nexp = 0;
    long gplane = 0, gsb = 0;
    std::vector<double> xi, wq, Dm, x1d, z1d;
    std::vector<double> Dxv, Dzv;
    DevBuf d_Dx, d_Dz, d_x1d, d_z1d;
    // receiver sets
    std::vector<std::vector<RecHost>> sets;  // per set, entries across tiles (ordered by tile)
    std::vector<int> set_size;
    DevBuf d_recs, d_rec_ptr;
    bool rec_dirty = true;
    size_t rec_up_entries = 0, rec_up_ptrs = 0, rec_up_sets = 0;  // prefix already on the device
    // two-step tiling (step2_kernel, tile shape `tile2` of TILE2_EX/TILE2_EZ)
    // and its receiver table
    int tile2 = 0, tx2 = TILE2_EX[0] - 3, tz2 = TILE2_EZ[0] - 3;
    int ntx2 = 0, ntz2 = 0, ntiles2 = 0;
    DevBuf d_recs2, d_rec_ptr2;
    size_t rec2_up_entries = 0, rec2_up_ptrs = 0, rec2_up_sets = 0;
    int fusion = 2;  // time steps per step-kernel launch where the batch path allows two (sem_plan_set_step_fusion)
    int corr_max = 64;  // adjoint steps per gradient-correlation chunk, at most (sem_plan_set_corr_chunk)
    std::map<std::pair<int, int>, BsplineTables*> bsp;
    // schedule of sem_forward_adjoint (sem_plan_set_schedule): groups of at
    // most sched_group solves (0: the whole batch), sched_lanes in flight
    int sched_group = 0, sched_lanes = 1;
    int ckpt_desc = 0;  // checkpoint_every as created (sem_plan_schedule_auto's fallback spacing)
    int streaming = SEM_STREAM_AUTO;  // state-stream memory hints (sem_plan_set_streaming)
    // lane streams and events, created on first use on device lane_dev
    int lane_dev = -1;
    std::vector<hipStream_t> lane_st;
    std::vector<hipEvent_t> lane_ev;
    hipEvent_t fork_ev = nullptr;
    ~sem_plan() {
        for (auto& kv : bsp) delete kv.second;
        for (hipStream_t s : lane_st) (void)hipStreamDestroy(s);
        for (hipEvent_t e : lane_ev) (void)hipEventDestroy(e);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
    }
};

namespace {

template <typename T>
int upload_vec(DevBuf& buf, const std::vector<double>& v) {
    std::vector<T> tmp(v.begin(), v.end());
    tmp.resize((v.size() + 7) / 8 * 8, T(1));  // vector loads read up to 3 past the end
    return buf.upload(tmp.data(), tmp.size() * sizeof(T));
}

template <typename T>
int upload_tables(sem_plan* pl) {
    int rc;
    if ((rc = upload_vec<T>(pl->d_Dx, pl->Dxv))) return rc;
    if ((rc = upload_vec<T>(pl->d_Dz, pl->Dzv))) return rc;
    return SEM_OK;
}

// Receiver entries grouped by tile for one tiling: tiles of tx_own x tz_own
// owned elements whose element block reaches lo elements to the left / below
// and hi to the right / above; `owned` marks the tile that owns the element.
template <typename T>
void build_rec_table(const sem_plan* pl, int tx_own, int tz_own, int ntx, int ntz, int lo, int hi, size_t first_set,
                     size_t base, std::vector<RecEntry<T, 4>>& flat, std::vector<int>& ptr) {
    const int ntiles = ntx * ntz;
    for (size_t si = first_set; si < pl->sets.size(); ++si) {
        const auto& s = pl->sets[si];
        std::vector<std::vector<const RecHost*>> by_tile(ntiles);
        for (auto& e : s) {
            for (int tz = std::max(0, e.ez / tz_own - 2); tz <= std::min(ntz - 1, e.ez / tz_own + 2); ++tz)
                for (int tx = std::max(0, e.ex / tx_own - 2); tx <= std::min(ntx - 1, e.ex / tx_own + 2); ++tx) {
                    if (e.ex < tx * tx_own - lo || e.ex > tx * tx_own + tx_own - 1 + hi) continue;
                    if (e.ez < tz * tz_own - lo || e.ez > tz * tz_own + tz_own - 1 + hi) continue;
                    by_tile[tz * ntx + tx].push_back(&e);
                }
        }
        for (int t = 0; t < ntiles; ++t) {
            ptr.push_back((int)(base + flat.size()));
            const int tx = t % ntx, tz = t / ntx;
            for (const RecHost* e : by_tile[t]) {
                RecEntry<T, 4> r;
                r.ex = e->ex;
                r.ez = e->ez;
                r.r = e->r;
                r.owned = (e->ex / tx_own == tx && e->ez / tz_own == tz) ? 1 : 0;
                for (int q = 0; q < 5; ++q) {
                    r.lx[q] = (T)e->lx[q];
                    r.lz[q] = (T)e->lz[q];
                }
                flat.push_back(r);
            }
        }
        ptr.push_back((int)(base + flat.size()));
    }
}

// Sets are only ever appended, so a table's device prefix stays valid: build
// and upload only the sets added since the last upload, in place after it
// (kernels in flight keep reading unchanged bytes), so that adding a set
// costs its own entries however many came before.  Growth past the reserved
// capacity reallocates, and then every set is built and uploaded again.
template <typename T>
int update_rec_table(const sem_plan* pl, int tx_own, int tz_own, int ntx, int ntz, int lo, int hi, DevBuf& d_recs,
                     DevBuf& d_ptr, size_t& up_entries, size_t& up_ptrs, size_t& up_sets) {
    const size_t eb = sizeof(RecEntry<T, 4>);
    std::vector<RecEntry<T, 4>> flat;
    std::vector<int> ptr;
    build_rec_table<T>(pl, tx_own, tz_own, ntx, ntz, lo, hi, up_sets, up_entries, flat, ptr);
    int rc;
    bool m1 = false, m2 = false;
    if ((rc = d_recs.reserve(std::max<size_t>((up_entries + flat.size()) * eb, 16), m1))) return rc;
    if ((rc = d_ptr.reserve((up_ptrs + ptr.size()) * sizeof(int), m2))) return rc;
    if (m1 || m2) {
        // the old buffers held the prefix
        up_entries = up_ptrs = up_sets = 0;
        flat.clear();
        ptr.clear();
        build_rec_table<T>(pl, tx_own, tz_own, ntx, ntz, lo, hi, 0, 0, flat, ptr);
    }
    if ((rc = d_recs.write_at(up_entries * eb, flat.data(), flat.size() * eb))) return rc;
    if ((rc = d_ptr.write_at(up_ptrs * sizeof(int), ptr.data(), ptr.size() * sizeof(int)))) return rc;
    up_entries += flat.size();
    up_ptrs += ptr.size();
    up_sets = pl->sets.size();
    return SEM_OK;
}

template <typename T>
int upload_receivers(sem_plan* pl) {
    // single-step tiles: one halo column / row on the left and below
    int rc = update_rec_table<T>(pl, TX, pl->tz, pl->ntx, pl->ntz, 1, 0, pl->d_recs, pl->d_rec_ptr,
                                 pl->rec_up_entries, pl->rec_up_ptrs, pl->rec_up_sets);
    if (rc) return rc;
    // two-step tiles: two halo columns / rows on the left and below, one on the right and above
    rc = update_rec_table<T>(pl, pl->tx2, pl->tz2, pl->ntx2, pl->ntz2, 2, 1, pl->d_recs2, pl->d_rec_ptr2,
                             pl->rec2_up_entries, pl->rec2_up_ptrs, pl->rec2_up_sets);
    if (rc) return rc;
    pl->rec_dirty = false;
    return SEM_OK;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Batch-level workspace: padded squared velocities of every model, the
// per-solve receiver-set ids and receiver counts, a flag word.
struct BatchLayout {
    size_t C2P, NREC, RSET, FLAG, total;
};

BatchLayout batch_layout(const sem_plan* pl, int B) {
    const size_t Ts = pl->d.precision == SEM_FP64 ? 8 : 4;
    BatchLayout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o += align256(bytes);
        return at;
    };
    L.C2P = take((size_t)B * pl->sb * Ts);
    L.NREC = take((size_t)B * 4);
    L.RSET = take((size_t)B * 4);
    L.FLAG = take(4);
    L.total = o;
    return L;
}

// Workspace of one group of solves (a lane's region; the whole batch when it
// is not split): states, checkpoints or the full forward history, the adjoint
// ring, residuals and the gradient accumulator.
struct Layout {
    bool full_history = false;
    int S = 0, nseg = 0;
    int C = 0;  // adjoint steps per gradient-correlation chunk
    size_t U0, U1, W, W2, CKU, CKW, SEG, RING, OM, RES, G, total;
};

// Correlation chunk: at most `cmax` adjoint steps (the plan's, default 64),
// dividing a segment of S steps into equal chunks (S = 100 -> 2 x 50).  The
// adjoint states of one chunk are kept in a ring of C + 1 slots; the FP64
// accumulator is read and written once per chunk (64 rather than 32 steps: 1 %
// faster on the lane schedule for a ring of 65 states per solve).
int corr_chunk(int S, int cmax) {
    const int nchunk = (S + cmax - 1) / cmax;
    return (S + nchunk - 1) / nchunk;
}

// Every nodal array is [solves][sb] in the padded layout.  ckpt: steps per
// checkpoint, <= 0 or >= nt for the full forward history.
Layout layout(const sem_plan* pl, int B, int pitch, bool adj, int ckpt) {
    const size_t Ts = pl->d.precision == SEM_FP64 ? 8 : 4;
    const size_t state = (size_t)B * pl->sb * Ts;
    Layout L{};
    const int nt = pl->d.nt;
    L.full_history = adj && (ckpt <= 0 || ckpt >= nt);
    L.S = L.full_history ? nt : ckpt;
    L.nseg = adj ? (L.full_history ? 1 : (nt + L.S - 1) / L.S) : 0;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o += align256(bytes);
        return at;
    };
    L.U0 = take(state);
    L.U1 = take(state);
    L.W = take(state);
    // the second increment buffer of two-step launches in a forward-only or a
    // checkpointed forward; under the full history the second buffer of the
    // state copies the forward's two-step launches read (with U1)
    L.W2 = take(state);
    if (adj) {
        if (!L.full_history) {
            L.CKU = take(state * L.nseg);
            L.CKW = take(state * L.nseg);
            L.SEG = take(state * (size_t)(L.S + 1));
        } else {
            L.SEG = take(state * (size_t)(nt + 1));
        }
        L.C = corr_chunk(L.S, pl->corr_max);
        L.RING = take(state * (size_t)(L.C + 1));
        L.OM = take(state);
        L.RES = take((size_t)B * nt * pitch * Ts);
        L.G = take((size_t)B * pl->gsb * 8);
    }
    L.total = o;
    return L;
}

// How sem_forward_adjoint runs a batch of B solves: ngroups groups of
// consecutive solves whose sizes differ by at most one, `lanes` of them in
// flight at once (each lane on a stream of its own with a workspace region of
// its own), every group with the history policy `ckpt`.
struct Sched {
    int group = 1, lanes = 1, ngroups = 1, ckpt = 0;
    int first(int i, int B) const { return i * (B / ngroups) + std::min(i, B % ngroups); }
    int size(int i, int B) const { return B / ngroups + (i < B % ngroups ? 1 : 0); }
};

Sched schedule(const sem_plan* pl, int B, bool adj) {
    Sched s;
    s.ckpt = pl->d.checkpoint_every;
    s.group = B;
    if (!adj || pl->sched_group <= 0 || pl->sched_group >= B) return s;
    s.ngroups = (B + pl->sched_group - 1) / pl->sched_group;
    s.lanes = std::max(1, std::min(pl->sched_lanes, s.ngroups));
    // as many groups per lane on every lane (no lane idles at the end)
    s.ngroups = std::min(B, (s.ngroups + s.lanes - 1) / s.lanes * s.lanes);
    s.group = (B + s.ngroups - 1) / s.ngroups;  // the largest group
    return s;
}

size_t workspace_total(const sem_plan* pl, int B, int pitch, bool adj) {
    const Sched s = schedule(pl, B, adj);
    return batch_layout(pl, B).total + (size_t)s.lanes * layout(pl, s.group, pitch, adj, s.ckpt).total;
}

// Schedules with full-history groups in lanes need this many step workgroups
// in flight (1.5 times the 1024 that 256 CUs hold at once with the step
// kernels' 4 waves per SIMD); below it the batch runs whole, with checkpoints.
constexpr long MIN_INFLIGHT_TILES = 1536;
// Lanes of groups in flight: 3 lanes and the caller's stream use the 4
// hardware queues HIP gives a process by default; a fourth lane shares a
// queue and was slower (801^2: 3 x 6 solves best; 161^2: 3 x 86 solves 1772
// solves/s, 4 x 64 1261).
constexpr int MAX_LANES = 3;

// Non-temporal state streams when the solves in flight touch more than the
// last-level cache per step (u^n, u^{n+1}, w, c^2 of each).
bool use_nt(const sem_plan* pl, long inflight) {
    if (pl->streaming != SEM_STREAM_AUTO) return pl->streaming == SEM_STREAM_NONTEMPORAL;
    const size_t Ts = pl->d.precision == SEM_FP64 ? 8 : 4;
    return (size_t)inflight * 4 * pl->sb * Ts > LLC_BYTES;
}

template <typename T>
int fill(T* p, T v, long n, hipStream_t st) {
    if (n <= 0) return SEM_OK;
    int blocks = (int)std::min<long>((n + 255) / 256, 4096);
    fill_kernel<T><<<blocks, 256, 0, st>>>(p, v, n);
    LAUNCH_CHECK();
    return SEM_OK;
}

// Step arguments of the solves of `bt`: c2p holds the padded squared
// velocities indexed by bt->model_of (or by solve), rset the device
// receiver-set id of each solve.
template <typename T>
StepArgs<T, 4> base_args(const sem_plan* pl, const sem_batch* bt, const void* c2p, const int* rset) {
    StepArgs<T, 4> a;
    std::memset(&a, 0, sizeof(a));
    const int P = 4;
    const double cx = 4.0 / (pl->dx * pl->dx), cz = 4.0 / (pl->dz * pl->dz);
    for (int i = 0; i <= P; ++i) {
        a.ec.ax[i] = (T)(cx * pl->wq[i]);
        a.ec.az[i] = (T)(cz * pl->wq[i]);
        for (int k = 0; k <= P; ++k) {
            a.ec.D[i][k] = (T)pl->Dm[i * (P + 1) + k];
            a.ec.Dw[i][k] = (T)(pl->Dm[i * (P + 1) + k] * pl->wq[i]);
            a.ec.gx[i][k] = (T)(cx * pl->wq[i] * pl->wq[k]);
            a.ec.gz[i][k] = (T)(cz * pl->wq[i] * pl->wq[k]);
        }
    }
    a.nex = pl->d.nelem_x;
    a.nez = pl->d.nelem_z;
    a.NX = pl->NX;
    a.NZ = pl->NZ;
    a.pitch = pl->pitch;
    a.ntx = pl->ntx;
    a.ntiles = pl->ntiles;
    a.total = pl->ntiles * bt->batch;
    a.nper = (a.total + NXCD - 1) / NXCD;
    a.sb = pl->sb;
    a.c2 = (const T*)c2p;
    a.model_of = bt->model_of;
    // folded D and Dw^T for apply5 (M[a][b] = -M[4-a][4-b])
    auto fold = [&](auto M, Fold5<T>& f) {
        for (int r = 0; r < 2; ++r) {
            for (int b = 0; b < 2; ++b) {
                f.s[r][b] = (T)(0.5 * (M(r, b) - M(r, 4 - b)));
                f.t[r][b] = (T)(0.5 * (M(r, b) + M(r, 4 - b)));
            }
            f.t[r][2] = (T)M(r, 2);
        }
        f.m[0] = (T)M(2, 0);
        f.m[1] = (T)M(2, 1);
    };
    fold([&](int r, int c) { return pl->Dm[r * (P + 1) + c]; }, a.ec.fD);
    fold([&](int r, int c) { return pl->Dm[c * (P + 1) + r] * pl->wq[c]; }, a.ec.fDwT);
    for (int i = 0; i <= P; ++i) a.ec.iw[i] = (T)(1.0 / pl->wq[i]);
    a.ec.iw_shared = (T)(1.0 / (pl->wq[0] + pl->wq[P]));
    a.ec.dt2 = (T)(pl->d.dt * pl->d.dt);
    a.Dx = (const T*)pl->d_Dx.p;
    a.Dz = (const T*)pl->d_Dz.p;
    a.damp = 1;
    a.src_node = bt->src_node;
    a.wavelet = (const T*)bt->wavelet;
    a.wav_bs = bt->wavelet_stride;
    a.rec_set_of = rset;
    a.rec_ptr = (const int*)pl->d_rec_ptr.p;
    a.recs = (const RecEntry<T, 4>*)pl->d_recs.p;
    a.rec_ptr2 = (const int*)pl->d_rec_ptr2.p;
    a.recs2 = (const RecEntry<T, 4>*)pl->d_recs2.p;
    a.ntx2 = pl->ntx2;
    a.ntiles2 = pl->ntiles2;
    a.total2 = pl->ntiles2 * bt->batch;
    a.nper2 = (a.total2 + NXCD - 1) / NXCD;
    a.seis = (T*)bt->seis;
    a.seis_bs = (long)pl->d.nt * bt->rec_pitch;
    a.rec_pitch = bt->rec_pitch;
    a.rec_t = -1;
    a.invJ = (T)(1.0 / pl->J);
    a.res_t = -1;
    return a;
}

template <typename T, int MODE, bool NTS>
int launch_step_nt(const sem_plan* pl, const StepArgs<T, 4>& a, hipStream_t st) {
    const unsigned grid = (unsigned)a.nper * NXCD;
    switch (pl->ez) {
        case 15: step_kernel<T, 4, MODE, 15, NTS><<<grid, tile_threads(15), 0, st>>>(a); break;
        case 7: step_kernel<T, 4, MODE, 7, NTS><<<grid, tile_threads(7), 0, st>>>(a); break;
        case 3: step_kernel<T, 4, MODE, 3, NTS><<<grid, tile_threads(3), 0, st>>>(a); break;
        default: return fail(SEM_E_UNSUPPORTED, "tile height without a kernel instance");
    }
    LAUNCH_CHECK();
    return SEM_OK;
}

template <typename T, int MODE>
int launch_step(const sem_plan* pl, const StepArgs<T, 4>& a, hipStream_t st, bool nts) {
    if constexpr (MODE == MODE_REC) return launch_step_nt<T, MODE, true>(pl, a, st);  // no state streams
    return nts ? launch_step_nt<T, MODE, true>(pl, a, st) : launch_step_nt<T, MODE, false>(pl, a, st);
}

// Two steps (n, n+1) per launch: u_in -> u_mid -> u_out, increment a.w -> a.w_out,
// on the plan's two-step tile shape.
template <typename T, int MODE, bool NTS, int S>
void launch_step2_shape(const StepArgs<T, 4>& a, unsigned grid, hipStream_t st) {
    constexpr int ex = TILE2_EX[S], ez = TILE2_EZ[S];
    step2_kernel<T, 4, MODE, NTS, ex, ez><<<grid, tile2_threads(ex, ez), 0, st>>>(a);
}

template <typename T, int MODE, bool NTS>
int launch_step2_nt(const sem_plan* pl, const StepArgs<T, 4>& a, hipStream_t st) {
    const unsigned grid = (unsigned)a.nper2 * NXCD;
    switch (pl->tile2) {
        case 0: launch_step2_shape<T, MODE, NTS, 0>(a, grid, st); break;
        case 1: launch_step2_shape<T, MODE, NTS, 1>(a, grid, st); break;
        case 2: launch_step2_shape<T, MODE, NTS, 2>(a, grid, st); break;
        case 3:
            if constexpr (sizeof(T) == 4) {
                launch_step2_shape<T, MODE, NTS, 3>(a, grid, st);
                break;
            } else {
                return fail(SEM_E_UNSUPPORTED, "two-step tile shape 3 has no FP64 instance");
            }
        case 4: launch_step2_shape<T, MODE, NTS, 4>(a, grid, st); break;
        default: return fail(SEM_E_UNSUPPORTED, "two-step tile shape without a kernel instance");
    }
    LAUNCH_CHECK();
    return SEM_OK;
}

template <typename T, int MODE>
int launch_step2(const sem_plan* pl, const StepArgs<T, 4>& a, hipStream_t st, bool nts) {
    if (!a.rec_ptr2 || (!a.slots && (!a.w_out || a.w_out == a.w || a.u_out == a.u_in || a.u_mid == a.u_in)))
        return fail(SEM_E_INVALID, "two-step launch with aliased or missing buffers");
    return nts ? launch_step2_nt<T, MODE, true>(pl, a, st) : launch_step2_nt<T, MODE, false>(pl, a, st);
}

template <typename T>
CorrArgs<T, 4> corr_args(const sem_plan* pl, const StepArgs<T, 4>& a, int B) {
    CorrArgs<T, 4> c;
    std::memset(&c, 0, sizeof(c));
    c.ec = a.ec;
    c.nex = a.nex;
    c.nez = a.nez;
    c.pitch = a.pitch;
    c.ctiles = pl->ctiles;
    c.total = pl->ctiles * B;
    c.nper = (c.total + NXCD - 1) / NXCD;
    c.sb = a.sb;
    c.gsb = pl->gsb;
    c.gplane = pl->gplane;
    c.nexp = pl->nexp;
    return c;
}

template <typename T>
int launch_corr(const sem_plan* pl, const CorrArgs<T, 4>& c, hipStream_t st) {
    (void)pl;
    const unsigned grid = (unsigned)c.nper * NXCD;
    corr_kernel<T, 4><<<grid, CORR_W, 0, st>>>(c);
    LAUNCH_CHECK();
    return SEM_OK;
}

int check_batch(sem_plan* pl, const sem_batch* bt, size_t ws_bytes, size_t need) {
    if (!bt || bt->batch <= 0) return fail(SEM_E_INVALID, "batch must be >= 1");
    if ((long)std::max(pl->ntiles, pl->ctiles) * bt->batch >= (1L << 31) - 64)
        return fail(SEM_E_INVALID, "batch too large for this mesh");
    if (!bt->c2 || !bt->wavelet || !bt->src_node || !bt->rec_set || (!bt->seis && bt->rec_pitch > 0))
        return fail(SEM_E_INVALID, "sem_batch: null pointer in c2/wavelet/src_node/rec_set/seis");
    if (bt->rec_pitch < 0) return fail(SEM_E_INVALID, "rec_pitch must be >= 0");
    if (bt->nmodel < 1 || bt->nmodel > bt->batch) return fail(SEM_E_INVALID, "nmodel must be in [1, batch]");
    if (!bt->model_of && bt->nmodel != bt->batch) return fail(SEM_E_INVALID, "model_of is NULL but nmodel != batch");
    if (ws_bytes < need)
        return fail(SEM_E_NOMEM, "workspace too small: need " + std::to_string(need) + " bytes");
    for (int b = 0; b < bt->batch; ++b) {
        int s = bt->rec_set[b];
        if (s < 0 || s >= (int)pl->sets.size())
            return fail(SEM_E_INVALID, "rec_set[" + std::to_string(b) + "] is not a registered receiver set");
        if (pl->set_size[s] > bt->rec_pitch) return fail(SEM_E_INVALID, "rec_pitch smaller than a receiver set");
    }
    return SEM_OK;
}

template <typename T>
int stage_batch(sem_plan* pl, const sem_batch* bt, char* ws, const BatchLayout& L, hipStream_t st) {
    if (pl->rec_dirty) {
        int rc = upload_receivers<T>(pl);
        if (rc) return rc;
    }
    std::vector<int> nrec(bt->batch);
    for (int b = 0; b < bt->batch; ++b) nrec[b] = pl->set_size[bt->rec_set[b]];
    HIP_TRY(hipMemcpyAsync(ws + L.RSET, bt->rec_set, sizeof(int) * bt->batch, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(ws + L.NREC, nrec.data(), sizeof(int) * bt->batch, hipMemcpyHostToDevice, st));
    // squared velocities into the padded layout
    dim3 grid((unsigned)std::min<long>((pl->N + 255) / 256, 2048), bt->nmodel);
    pad_rows_kernel<T><<<grid, 256, 0, st>>>((const T*)bt->c2, (T*)(ws + L.C2P), pl->NX, pl->NZ, pl->pitch, pl->sb);
    LAUNCH_CHECK();
    // a non-finite model is rejected before any time step (the reference's CFL
    // check raises on it, utils.py:64-93)
    int* flag = (int*)(ws + L.FLAG);
    fill_kernel<int><<<1, 64, 0, st>>>(flag, 0, 1);
    LAUNCH_CHECK();
    const long nc2 = (long)bt->nmodel * pl->N;
    finite_check_kernel<T><<<(unsigned)std::min<long>((nc2 + 255) / 256, 4096), 256, 0, st>>>((const T*)bt->c2, nc2,
                                                                                            flag);
    LAUNCH_CHECK();
    int bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    // the pageable buffers above must stay valid until the copies land
    HIP_TRY(hipStreamSynchronize(st));
    if (bad) return fail(SEM_E_NONFINITE, "squared velocity contains NaN or Inf");
    return SEM_OK;
}

// Wavefield snapshots of a node box every `every` steps (null: none).
struct SnapSpec {
    int every, ix0, nxr, iz0, nzr;
    void* out;
};

// The solves [g0, g0 + n) of a batch and the buffers they run in.
struct Group {
    int g0 = 0, n = 0;
    sem_batch sub{};             // the group's solves (batch fields offset to solve g0)
    const void* c2p = nullptr;   // padded c^2 as indexed by sub.model_of (or by solve)
    const int* rset = nullptr;   // device receiver-set ids of the group's solves
    const int* nrec = nullptr;   // device receiver counts of the group's solves
    char* ws = nullptr;          // the group's workspace region (Layout offsets)
    bool nts = true;             // non-temporal state streams
};

template <typename T>
Group make_group(const sem_plan* pl, const sem_batch* bt, char* ws, const BatchLayout& BL, int g0, int n,
                 char* region, bool nts) {
    Group g;
    g.g0 = g0;
    g.n = n;
    g.sub = *bt;
    g.sub.batch = n;
    g.sub.src_node = bt->src_node + g0;
    g.sub.model_of = bt->model_of ? bt->model_of + g0 : nullptr;
    g.sub.wavelet = (const T*)bt->wavelet + (long)g0 * bt->wavelet_stride;
    g.sub.rec_set = bt->rec_set + g0;
    g.sub.seis = bt->seis ? (void*)((T*)bt->seis + (long)g0 * pl->d.nt * bt->rec_pitch) : nullptr;
    g.c2p = bt->model_of ? (const void*)(ws + BL.C2P) : (const void*)((const T*)(ws + BL.C2P) + (long)g0 * pl->sb);
    g.rset = (const int*)(ws + BL.RSET) + g0;
    g.nrec = (const int*)(ws + BL.NREC) + g0;
    g.ws = region;
    g.nts = nts;
    return g;
}

// The state copies of the full-history forward's two-step launches: each
// launch also writes its new state (plain stores) to one buffer of a pair,
// which the next launch reads, and writes the history slot itself
// non-temporally, so that the write-once history stays out of the last-level
// cache (the same for the adjoint's ring measured no faster).  `cur` is null
// until a launch has written a copy.
template <typename T>
struct CopyPair {
    T* cur;
    T* a;
    T* b;
    void next_launch(StepArgs<T, 4>& s) {
        if (cur) s.u_in = cur;
        s.u_copy = (cur == a) ? b : a;
        cur = s.u_copy;
    }
};

// Forward time loop of a group.  Checkpoint mode: ping-pong U0/U1 with (u, w)
// mirrored into checkpoint k after step k*S-1.  Full-history mode: u^n in SEG
// slot n.
template <typename T>
int forward_loop(sem_plan* pl, const Group& g, const Layout& L, bool adj, hipStream_t st,
                 const SnapSpec* snap = nullptr) {
    const sem_batch* bt = &g.sub;
    char* ws = g.ws;
    const int B = bt->batch, nt = pl->d.nt;
    const long state = (long)B * pl->sb;
    StepArgs<T, 4> a = base_args<T>(pl, bt, g.c2p, g.rset);
    T* U0 = (T*)(ws + L.U0);
    T* U1 = (T*)(ws + L.U1);
    T* W = (T*)(ws + L.W);
    T* SEG = (T*)(ws + L.SEG);
    int rc;
    if ((rc = fill<T>(W, T(0), state, st))) return rc;
    if (adj && L.full_history) {
        if ((rc = fill<T>(SEG, T(0), state, st))) return rc;
    } else {
        if ((rc = fill<T>(U0, T(0), state, st))) return rc;
    }
    if (adj && !L.full_history) {
        if ((rc = fill<T>((T*)(ws + L.CKU), T(0), state, st))) return rc;
        if ((rc = fill<T>((T*)(ws + L.CKW), T(0), state, st))) return rc;
    }
    const bool post = pl->d.record_mode == SEM_RECORD_POST;
    a.w = W;
    const T* last = nullptr;
    // two steps per launch under the full history (the increment alternates
    // between W and U0, which the full history leaves unused) and in a
    // forward-only pass (u between U0 and U1, the increment between W and W2);
    // not with checkpoints or snapshots, which take every step's state
    const bool hist = adj && L.full_history;
    const bool two = pl->fusion == 2;
    T* W2 = hist ? U0 : (T*)(ws + L.W2);
    T* ucur = U0;  // u^n outside the full history (ping-pong)
    T* unext = U1;
    // full history: u^{n+2} also goes to a ping-pong pair (U1, L.W2) that the
    // next launch reads, so that every history slot is written non-temporally
    CopyPair<T> cp{nullptr, U1, (T*)(ws + L.W2)};
    // checkpointed forward: checkpoint k holds (u, w) after step k*S-1; a pair
    // of steps may end on a checkpoint but not straddle one
    const bool ckpt = adj && !L.full_history;
    // snapshot of u^{n+1} after step n when n % every == 0: a pair of steps
    // may end on one (its output is the state) but not hold one in its middle
    auto snapshot = [&](int n, const T* u) -> int {
        const long per = (long)snap->nxr * snap->nzr;
        const long nsnap = (nt + snap->every - 1) / snap->every;
        dim3 grid((unsigned)std::min<long>((per + 255) / 256, 1024), B);
        snapshot_kernel<T><<<grid, 256, 0, st>>>(u, pl->sb, pl->pitch, snap->ix0, snap->nxr, snap->iz0, snap->nzr,
                                                 (T*)snap->out + (n / snap->every) * per, nsnap * per);
        LAUNCH_CHECK();
        return SEM_OK;
    };
    for (int n = 0; n < nt; ++n) {
        if (two && n + 1 < nt && !(ckpt && (n + 1) % L.S == 0) && !(snap && n % snap->every == 0)) {
            a.u_in = hist ? SEG + (long)n * state : ucur;
            a.u_mid = hist ? SEG + (long)(n + 1) * state : nullptr;
            a.u_out = hist ? SEG + (long)(n + 2) * state : unext;
            a.u_copy = nullptr;
            if (hist) cp.next_launch(a);
            a.w_out = (a.w == W) ? W2 : W;
            a.n = n;
            a.rec_t = post ? n - 1 : n;
            a.ck_u = a.ck_w = nullptr;
            if (ckpt && (n + 2) % L.S == 0 && n + 2 < nt) {
                const long k = (n + 2) / L.S;
                a.ck_u = (T*)(ws + L.CKU) + k * state;
                a.ck_w = (T*)(ws + L.CKW) + k * state;
            }
            if ((rc = launch_step2<T, MODE_FWD>(pl, a, st, g.nts))) return rc;
            a.w = a.w_out;
            last = a.u_out;
            std::swap(ucur, unext);
            ++n;
            if (snap && n % snap->every == 0 && (rc = snapshot(n, a.u_out))) return rc;
            continue;
        }
        if (hist) {
            a.u_in = SEG + (long)n * state;
            a.u_out = SEG + (long)(n + 1) * state;
            a.u_copy = nullptr;
        } else {
            a.u_in = ucur;
            a.u_out = unext;
            std::swap(ucur, unext);
        }
        a.n = n;
        a.rec_t = post ? n - 1 : n;
        a.ck_u = a.ck_w = nullptr;
        if (adj && !L.full_history && (n + 1) % L.S == 0 && n + 1 < nt) {
            const long k = (n + 1) / L.S;
            a.ck_u = (T*)(ws + L.CKU) + k * state;
            a.ck_w = (T*)(ws + L.CKW) + k * state;
        }
        if ((rc = launch_step<T, MODE_FWD>(pl, a, st, g.nts))) return rc;
        last = a.u_out;
        if (snap && n % snap->every == 0 && (rc = snapshot(n, a.u_out))) return rc;
    }
    if (post) {
        a.u_in = last;
        a.rec_t = nt - 1;
        if ((rc = launch_step<T, MODE_REC>(pl, a, st, g.nts))) return rc;
    }
    return SEM_OK;
}

template <typename T>
int forward_impl(sem_plan* pl, const sem_batch* bt, void* wsv, size_t ws_bytes, hipStream_t st,
                 const SnapSpec* snap = nullptr) {
    const int B = bt ? bt->batch : 1;
    const BatchLayout BL = batch_layout(pl, B);
    const Layout L = layout(pl, B, bt ? bt->rec_pitch : 0, false, 0);
    int rc;
    if ((rc = check_batch(pl, bt, ws_bytes, BL.total + L.total))) return rc;
    char* ws = (char*)wsv;
    if ((rc = stage_batch<T>(pl, bt, ws, BL, st))) return rc;
    if ((rc = fill<T>((T*)bt->seis, T(0), (long)bt->batch * pl->d.nt * bt->rec_pitch, st))) return rc;
    const Group g = make_group<T>(pl, bt, ws, BL, 0, B, ws + BL.total, use_nt(pl, B));
    return forward_loop<T>(pl, g, L, false, st, snap);
}

// Forward, misfit, adjoint and gradient of one group, in order on one stream.
template <typename T>
int run_group(sem_plan* pl, const Group& g, const Layout& L, const void* obs, double inv_sigma2, double* misfit,
              double* grad_c, double* grad_c2, hipStream_t st) {
    const sem_batch* bt = &g.sub;
    char* ws = g.ws;
    const int B = bt->batch, nt = pl->d.nt;
    const long state = (long)B * pl->sb;
    // without a gradient output only the misfit is wanted: no history, no adjoint
    const bool adjoint = grad_c || grad_c2;
    int rc;
    if ((rc = forward_loop<T>(pl, g, L, adjoint, st))) return rc;

    // residual and misfit
    T* RES = (T*)(ws + L.RES);
    residual_kernel<T><<<B, RESID_THREADS, 0, st>>>((const T*)bt->seis, (const double*)obs + (long)g.g0 * nt * bt->rec_pitch, RES,
                                          misfit + g.g0, g.nrec, 0, nt, bt->rec_pitch, (long)nt * bt->rec_pitch,
                                          pl->d.win_begin, pl->d.win_end, pl->d.win_decim);
    LAUNCH_CHECK();
    if (!adjoint) return SEM_OK;

    // adjoint sweep: per segment, recompute u into SEG, then chunks of up to C
    // adjoint steps (states in a ring indexed by time) each followed by the
    // gradient correlation of the chunk
    T* RING = (T*)(ws + L.RING);
    T* OM = (T*)(ws + L.OM);
    T* OM2 = (T*)(ws + L.U1);  // second increment buffer of two-step launches
    double* G = (double*)(ws + L.G);
    const int ring_slots = L.C + 1;
    auto ring_slot = [&](int t) { return RING + (long)((t + 1) % ring_slots) * state; };  // lambda_t
    // adjoint steps t_lo <= it < t_hi: lambda vanishes after the last windowed
    // sample; the reference adjoint stops at the window start
    const int t_hi = pl->d.win_end;
    const int t_lo = pl->d.adjoint_mode == SEM_ADJ_REFERENCE ? pl->d.win_begin : 0;
    if ((rc = fill<T>(ring_slot(t_hi - 1), T(0), state, st))) return rc;
    if ((rc = fill<T>(OM, T(0), state, st))) return rc;
    if ((rc = fill<double>(G, 0.0, (long)B * pl->gsb, st))) return rc;

    StepArgs<T, 4> f = base_args<T>(pl, bt, g.c2p, g.rset);  // recompute steps
    f.seis = nullptr;  // record nothing (a two-step launch samples rec_t and rec_t + 1)
    StepArgs<T, 4> a = base_args<T>(pl, bt, g.c2p, g.rset);  // adjoint steps
    a.w = OM;
    a.damp = pl->d.adjoint_mode == SEM_ADJ_EXACT ? 1 : 0;
    a.src_node = nullptr;
    a.resid = RES;
    a.resid_bs = (long)nt * bt->rec_pitch;
    CorrArgs<T, 4> c = corr_args<T>(pl, a, B);
    c.ustep = state;
    c.ring = RING;
    c.ring_slots = ring_slots;
    c.rstep = state;
    c.G = G;
    T* SEG = (T*)(ws + L.SEG);
    for (int k = L.nseg - 1; k >= 0; --k) {
        const int s0 = k * L.S;
        const int end = std::min(std::min(L.S, nt - s0) + s0, t_hi);  // last adjoint step + 1 in this segment
        const int beg = std::max(s0, t_lo);
        if (end <= beg) continue;
        const int len = end - s0;
        if (!L.full_history) {
            // recompute u^{s0+1..s0+len} from checkpoint k into SEG[1..len]
            const T* slot0 = (T*)(ws + L.CKU) + (long)k * state;
            T* ckw = (T*)(ws + L.CKW) + (long)k * state;  // updated in place (its last use)
            f.w = ckw;
            for (int m = 0; m < len; ++m) {
                f.u_in = (m == 0) ? slot0 : SEG + (long)m * state;
                f.n = s0 + m;
                if (pl->fusion == 2 && m + 1 < len) {
                    // two steps per launch; the increment alternates between the
                    // checkpoint's copy and U0 (idle while the adjoint runs)
                    f.u_mid = SEG + (long)(m + 1) * state;
                    f.u_out = SEG + (long)(m + 2) * state;
                    f.w_out = (f.w == ckw) ? (T*)(ws + L.U0) : ckw;
                    if ((rc = launch_step2<T, MODE_FWD>(pl, f, st, g.nts))) return rc;
                    f.w = f.w_out;
                    ++m;
                    continue;
                }
                f.u_out = SEG + (long)(m + 1) * state;
                if ((rc = launch_step<T, MODE_FWD>(pl, f, st, g.nts))) return rc;
            }
        }
        const int useg0 = L.full_history ? 0 : s0;  // SEG slot j holds u^{useg0 + j}
        for (int ce = end; ce > beg;) {
            const int clen = std::min(L.C, ce - beg);
            const int c0 = ce - clen;
            for (int it = ce - 1; it >= c0; --it) {
                a.u_in = ring_slot(it);
                a.res_t = it;
                if (pl->fusion == 2 && it - 1 >= c0 && ring_slots >= 3) {
                    // steps it and it - 1 in one launch (the increment alternates
                    // between OM and U1, idle while the adjoint runs)
                    a.u_mid = ring_slot(it - 1);
                    a.u_out = ring_slot(it - 2);
                    a.w_out = (a.w == OM) ? OM2 : OM;
                    if ((rc = launch_step2<T, MODE_ADJS>(pl, a, st, g.nts))) return rc;
                    a.w = a.w_out;
                    --it;
                    continue;
                }
                a.u_out = ring_slot(it - 1);
                if ((rc = launch_step<T, MODE_ADJS>(pl, a, st, g.nts))) return rc;
            }
            c.uf = SEG + (long)(c0 + 1 - useg0) * state;  // u^{c0+1}
            c.c0 = c0;
            c.clen = clen;
            if ((rc = launch_corr<T>(pl, c, st))) return rc;
            ce = c0;
        }
    }
    if (grad_c || grad_c2) {
        const long N = pl->N;
        dim3 grid((unsigned)std::min<long>((N + 255) / 256, 2048), B);
        grad_finalize_kernel<T><<<grid, 256, 0, st>>>(G, pl->gsb, pl->sb, pl->gplane, pl->nexp, pl->d.nelem_x,
                                                      pl->d.nelem_z, pl->NX, pl->NZ, pl->pitch, (const T*)g.c2p,
                                                      bt->model_of, pl->J * inv_sigma2,
                                                      grad_c ? grad_c + (long)g.g0 * N : nullptr,
                                                      grad_c2 ? grad_c2 + (long)g.g0 * N : nullptr);
        LAUNCH_CHECK();
    }
    return SEM_OK;
}

// The plan's lane streams and events on device `dev` (created on first use).
int ensure_lanes(sem_plan* pl, int lanes, int dev) {
    if (pl->lane_dev != dev) {
        if (!pl->lane_st.empty()) return fail(SEM_E_INVALID, "a plan's lanes serve the device they were created on");
        pl->lane_dev = dev;
    }
    if (!pl->fork_ev) HIP_TRY(hipEventCreateWithFlags(&pl->fork_ev, hipEventDisableTiming));
    while ((int)pl->lane_st.size() < lanes) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        pl->lane_st.push_back(s);
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        pl->lane_ev.push_back(e);
    }
    return SEM_OK;
}

template <typename T>
int forward_adjoint_impl(sem_plan* pl, const sem_batch* bt, const void* obs, double inv_sigma2, double* misfit,
                         double* grad_c, double* grad_c2, void* wsv, size_t ws_bytes, hipStream_t st) {
    const int Bt = bt ? bt->batch : 1;
    const Sched s = schedule(pl, Bt, true);
    const BatchLayout BL = batch_layout(pl, Bt);
    const Layout L = layout(pl, s.group, bt ? bt->rec_pitch : 0, true, s.ckpt);
    int rc;
    if ((rc = check_batch(pl, bt, ws_bytes, BL.total + (size_t)s.lanes * L.total))) return rc;
    if ((!obs && bt->rec_pitch > 0) || !misfit) return fail(SEM_E_INVALID, "obs and misfit must be non-null");
    char* ws = (char*)wsv;
    if ((rc = stage_batch<T>(pl, bt, ws, BL, st))) return rc;
    if ((rc = fill<T>((T*)bt->seis, T(0), (long)Bt * pl->d.nt * bt->rec_pitch, st))) return rc;
    const bool nts = use_nt(pl, (long)s.lanes * s.group);
    auto group = [&](int i, int lane) {
        return make_group<T>(pl, bt, ws, BL, s.first(i, Bt), s.size(i, Bt), ws + BL.total + (size_t)lane * L.total,
                             nts);
    };
    if (s.lanes == 1) {
        for (int i = 0; i < s.ngroups; ++i)
            if ((rc = run_group<T>(pl, group(i, 0), L, obs, inv_sigma2, misfit, grad_c, grad_c2, st))) return rc;
        return SEM_OK;
    }
    // lanes: groups i = k, k + lanes, ... run on lane k's stream, issued by a
    // host thread per lane; the lanes start after the staging on `st`, and
    // `st` waits for every lane
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if ((rc = ensure_lanes(pl, s.lanes, dev))) return rc;
    HIP_TRY(hipEventRecord(pl->fork_ev, st));
    for (int k = 0; k < s.lanes; ++k) HIP_TRY(hipStreamWaitEvent(pl->lane_st[k], pl->fork_ev, 0));
    std::vector<int> rcs(s.lanes, SEM_OK);
    std::vector<std::string> msgs(s.lanes);
    auto lane_work = [&](int k) {
        if (hipSetDevice(dev) != hipSuccess) {
            rcs[k] = SEM_E_DEVICE;
            msgs[k] = "hipSetDevice failed in a lane thread";
            return;
        }
        for (int i = k; i < s.ngroups; i += s.lanes) {
            const int r = run_group<T>(pl, group(i, k), L, obs, inv_sigma2, misfit, grad_c, grad_c2,
                                       pl->lane_st[k]);
            if (r) {
                rcs[k] = r;
                msgs[k] = g_err;
                return;
            }
        }
    };
    // lane 0 runs on the calling thread; a lane whose thread cannot be
    // started runs there too, after it (no exception leaves the C ABI)
    std::vector<std::thread> th;
    std::vector<int> inline_lanes{0};
    for (int k = 1; k < s.lanes; ++k) {
        try {
            th.emplace_back(lane_work, k);
        } catch (...) {
            inline_lanes.push_back(k);
        }
    }
    for (int k : inline_lanes) lane_work(k);
    for (auto& t : th) t.join();
    for (int k = 0; k < s.lanes; ++k) {
        HIP_TRY(hipEventRecord(pl->lane_ev[k], pl->lane_st[k]));
        HIP_TRY(hipStreamWaitEvent(st, pl->lane_ev[k], 0));
    }
    for (int k = 0; k < s.lanes; ++k)
        if (rcs[k]) return fail(rcs[k], msgs[k]);
    return SEM_OK;
}

BsplineTables* get_bspline(sem_plan* pl, int nbz, int nbx, int& rc) {
    auto key = std::make_pair(nbz, nbx);
    auto it = pl->bsp.find(key);
    if (it != pl->bsp.end()) return it->second;
    auto* t = new BsplineTables;
    std::vector<int> bz0, bx0;
    std::vector<double> bzw, bxw;
    bspline_tables(pl->z1d, pl->d.zmin, pl->d.zmax, nbz, bz0, bzw);
    bspline_tables(pl->x1d, pl->d.xmin, pl->d.xmax, nbx, bx0, bxw);
    // support ranges [lo, hi) per basis function
    auto ranges = [](const std::vector<int>& first, int n, std::vector<int>& lo, std::vector<int>& hi) {
        lo.assign(n, (int)first.size());
        hi.assign(n, 0);
        for (int c = 0; c < (int)first.size(); ++c)
            for (int r = 0; r < 4; ++r) {
                int q = first[c] + r;
                if (q < n) {
                    lo[q] = std::min(lo[q], c);
                    hi[q] = std::max(hi[q], c + 1);
                }
            }
        for (int q = 0; q < n; ++q)
            if (hi[q] < lo[q]) lo[q] = hi[q] = 0;
    };
    std::vector<int> plo, phi, qlo, qhi;
    ranges(bz0, nbz, plo, phi);
    ranges(bx0, nbx, qlo, qhi);
    rc = SEM_OK;
    if (!rc) rc = t->bz0.upload(bz0.data(), bz0.size() * 4);
    if (!rc) rc = t->bzw.upload(bzw.data(), bzw.size() * 8);
    if (!rc) rc = t->bx0.upload(bx0.data(), bx0.size() * 4);
    if (!rc) rc = t->bxw.upload(bxw.data(), bxw.size() * 8);
    if (!rc) rc = t->plo.upload(plo.data(), plo.size() * 4);
    if (!rc) rc = t->phi.upload(phi.data(), phi.size() * 4);
    if (!rc) rc = t->qlo.upload(qlo.data(), qlo.size() * 4);
    if (!rc) rc = t->qhi.upload(qhi.data(), qhi.size() * 4);
    if (rc) {
        delete t;
        return nullptr;
    }
    pl->bsp[key] = t;
    return t;
}

}  // namespace

// =============================================================== C ABI
extern "C" {

const char* sem_last_error(void) { return g_err.c_str(); }

int sem_stream_create(void** stream) try {
    if (!stream) return fail(SEM_E_INVALID, "sem_stream_create: bad argument");
    *stream = nullptr;
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return SEM_OK;
} ABI_CATCH

int sem_stream_destroy(void* stream) try {
    if (!stream) return SEM_OK;
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return SEM_OK;
} ABI_CATCH

int sem_plan_create(const sem_desc* d, sem_plan** out) try {
    if (!d || !out) return fail(SEM_E_INVALID, "null argument");
    *out = nullptr;
    if (d->order != 4) return fail(SEM_E_UNSUPPORTED, "only polynomial order 4 has kernels");
    if (d->nelem_x < 1 || d->nelem_z < 1) return fail(SEM_E_INVALID, "nelem_x/nelem_z must be >= 1");
    // kernels address one solve's padded arrays with 32-bit node and byte offsets
    const long elem_bytes = d->precision == SEM_FP64 ? 8 : 4;
    if (d->nelem_x > (1 << 20) || d->nelem_z > (1 << 20) ||
        ((long)d->nelem_z * 4 + 1) * (((long)d->nelem_x * 4 + 1 + 31) / 32 * 32) * elem_bytes >= (1L << 32) - 256)
        return fail(SEM_E_INVALID, "mesh too large: one padded nodal array must stay below 4 GiB");
    if (!(d->xmax > d->xmin) || !(d->zmax > d->zmin)) return fail(SEM_E_INVALID, "empty domain");
    if (!(d->dt > 0) || d->nt < 1) return fail(SEM_E_INVALID, "dt must be > 0 and nt >= 1");
    if (d->precision != SEM_FP32 && d->precision != SEM_FP64) return fail(SEM_E_INVALID, "bad precision");
    if (d->adjoint_mode != SEM_ADJ_REFERENCE && d->adjoint_mode != SEM_ADJ_EXACT)
        return fail(SEM_E_INVALID, "bad adjoint_mode");
    if (d->record_mode != SEM_RECORD_POST && d->record_mode != SEM_RECORD_PRE)
        return fail(SEM_E_INVALID, "bad record_mode");
    if (d->sponge_sides && !(d->sponge_width > 0)) return fail(SEM_E_INVALID, "sponge_width must be > 0");
    if (d->tile_rows != 0 && d->tile_rows != 15 && d->tile_rows != 7 && d->tile_rows != 3)
        return fail(SEM_E_UNSUPPORTED, "tile_rows must be 0 (default), 15, 7 or 3");
    const int wend = d->win_end > 0 ? d->win_end : d->nt;
    if (d->win_begin < 0 || wend > d->nt || d->win_begin >= wend)
        return fail(SEM_E_INVALID, "misfit window must satisfy 0 <= win_begin < win_end <= nt");
    auto* pl = new sem_plan;
    pl->d = *d;
    pl->d.win_end = wend;
    pl->d.win_decim = d->win_decim > 1 ? d->win_decim : 1;
    pl->ckpt_desc = d->checkpoint_every;
    const int P = 4;
    pl->NX = d->nelem_x * P + 1;
    pl->NZ = d->nelem_z * P + 1;
    pl->N = (long)pl->NX * pl->NZ;
    pl->pitch = (pl->NX + 31) / 32 * 32;
    pl->sb = ((long)pl->NZ * pl->pitch + 63) / 64 * 64;
    pl->dx = (d->xmax - d->xmin) / d->nelem_x;
    pl->dz = (d->zmax - d->zmin) / d->nelem_z;
    pl->J = pl->dx * pl->dz * 0.25;
    pl->ez = d->tile_rows > 0 ? d->tile_rows : EZ_DEFAULT;
    pl->tz = pl->ez - 1;
    pl->ntx = (d->nelem_x + TX - 1) / TX;
    pl->ntz = (d->nelem_z + pl->tz - 1) / pl->tz;
    pl->ntiles = pl->ntx * pl->ntz;
    pl->ntx2 = (d->nelem_x + pl->tx2 - 1) / pl->tx2;
    pl->ntz2 = (d->nelem_z + pl->tz2 - 1) / pl->tz2;
    pl->ntiles2 = pl->ntx2 * pl->ntz2;
    pl->ctiles = (int)(((long)d->nelem_x * d->nelem_z + CORR_W - 1) / CORR_W);
    pl->nexp = (d->nelem_x + 31) / 32 * 32;
    pl->gplane = (long)d->nelem_z * pl->nexp;
    pl->gsb = (pl->sb + 9 * pl->gplane + 63) / 64 * 64;
    pl->xi = gll_nodes(P);
    pl->wq = gll_weights(pl->xi);
    pl->Dm = diff_matrix(pl->xi);
    // node coordinates (mesh.py:22-30 arithmetic)
    auto axis = [&](double a, double b, int ne, std::vector<double>& out1) {
        out1.assign(ne * P + 1, 0.0);
        for (int e = 0; e < ne; ++e) {
            double x0 = a + (b - a) * e / ne, x1 = a + (b - a) * (e + 1) / ne;
            if (e == ne - 1) x1 = b;
            for (int i = 0; i <= P; ++i) out1[e * P + i] = 0.5 * ((1 - pl->xi[i]) * x0 + (1 + pl->xi[i]) * x1);
        }
    };
    axis(d->xmin, d->xmax, d->nelem_x, pl->x1d);
    axis(d->zmin, d->zmax, d->nelem_z, pl->z1d);
    // the lumped mass M = J Wx Wz (operators.py:65-84) is evaluated in the kernels from the GLL weights
    // sponge: cubic ramp, d0 = -(m+1) c_max ln R / (2 delta)  (boundary.py:26-52)
    const double R = d->sponge_R > 0 ? d->sponge_R : 1e-8;
    const double del = d->sponge_width;
    const double d0 = del > 0 ? -4.0 * d->sponge_cmax * std::log(R) / (2.0 * del) : 0.0;
    auto ramp = [&](double depth) { return d0 * std::pow(depth / del, 3); };
    pl->Dxv.assign(pl->NX, 1.0);
    pl->Dzv.assign(pl->NZ, 1.0);
    for (int i = 0; i < pl->NX; ++i) {
        double x = pl->x1d[i], s = 0.0;
        if ((d->sponge_sides & SEM_SIDE_XMIN) && x < d->xmin + del) s = ramp(d->xmin + del - x);
        if ((d->sponge_sides & SEM_SIDE_XMAX) && x > d->xmax - del) s = ramp(x - (d->xmax - del));
        pl->Dxv[i] = std::exp(-s * d->dt);
    }
    for (int j = 0; j < pl->NZ; ++j) {
        double z = pl->z1d[j], s = 0.0;
        if ((d->sponge_sides & SEM_SIDE_ZMIN) && z < d->zmin + del) s = ramp(d->zmin + del - z);
        if ((d->sponge_sides & SEM_SIDE_ZMAX) && z > d->zmax - del) s = ramp(z - (d->zmax - del));
        pl->Dzv[j] = std::exp(-s * d->dt);
    }
    int rc = d->precision == SEM_FP64 ? upload_tables<double>(pl) : upload_tables<float>(pl);
    if (!rc) rc = pl->d_x1d.upload(pl->x1d.data(), sizeof(double) * pl->NX);
    if (!rc) rc = pl->d_z1d.upload(pl->z1d.data(), sizeof(double) * pl->NZ);
    if (rc) {
        delete pl;
        return rc;
    }
    *out = pl;
    return SEM_OK;
} ABI_CATCH

int sem_plan_destroy(sem_plan* plan) try {
    delete plan;
    return SEM_OK;
} ABI_CATCH

int sem_plan_dims(const sem_plan* pl, int* NX, int* NZ, long* N) try {
    if (!pl) return fail(SEM_E_INVALID, "null plan");
    if (NX) *NX = pl->NX;
    if (NZ) *NZ = pl->NZ;
    if (N) *N = pl->N;
    return SEM_OK;
} ABI_CATCH

int sem_plan_axes(const sem_plan* pl, double* x1d, double* z1d) try {
    if (!pl) return fail(SEM_E_INVALID, "null plan");
    if (x1d) std::memcpy(x1d, pl->x1d.data(), sizeof(double) * pl->NX);
    if (z1d) std::memcpy(z1d, pl->z1d.data(), sizeof(double) * pl->NZ);
    return SEM_OK;
} ABI_CATCH

int sem_receiver_set_add(sem_plan* pl, int nrec, const double* xz, int* set_id) try {
    if (!pl || nrec < 0 || (nrec > 0 && !xz) || !set_id) return fail(SEM_E_INVALID, "bad receiver set");
    std::vector<RecHost> s;
    const auto& d = pl->d;
    for (int r = 0; r < nrec; ++r) {
        const double rx = xz[2 * r], rz = xz[2 * r + 1];
        if (!std::isfinite(rx) || !std::isfinite(rz)) return fail(SEM_E_INVALID, "non-finite receiver position");
        // element by truncation, clamped (receivers.py:32-35)
        int ex = (int)((rx - d.xmin) / pl->dx);
        int ez = (int)((rz - d.zmin) / pl->dz);
        ex = std::max(0, std::min(ex, d.nelem_x - 1));
        ez = std::max(0, std::min(ez, d.nelem_z - 1));
        const double xi = 2.0 * (rx - (d.xmin + ex * pl->dx)) / pl->dx - 1.0;
        const double eta = 2.0 * (rz - (d.zmin + ez * pl->dz)) / pl->dz - 1.0;
        RecHost e;
        e.ex = ex;
        e.ez = ez;
        e.r = r;
        e.owned = 0;
        for (int q = 0; q <= 4; ++q) {
            e.lx[q] = lagrange(xi, q, pl->xi, d.lagrange_eps);
            e.lz[q] = lagrange(eta, q, pl->xi, d.lagrange_eps);
        }
        s.push_back(e);
    }
    pl->sets.push_back(std::move(s));
    pl->set_size.push_back(nrec);
    pl->rec_dirty = true;
    *set_id = (int)pl->sets.size() - 1;
    return SEM_OK;
} ABI_CATCH

int sem_receiver_set_size(const sem_plan* pl, int set_id, int* nrec) try {
    if (!pl || !nrec || set_id < 0 || set_id >= (int)pl->sets.size()) return fail(SEM_E_INVALID, "bad set id");
    *nrec = pl->set_size[set_id];
    return SEM_OK;
} ABI_CATCH

int sem_source_node(const sem_plan* pl, double x, double z, int* node) try {
    if (!pl || !node) return fail(SEM_E_INVALID, "null argument");
    // nearest node; on a tensor grid the squared distance separates per axis
    auto nearest = [](const std::vector<double>& v, double q) {
        int best = 0;
        double bd = std::fabs(v[0] - q);
        for (int i = 1; i < (int)v.size(); ++i) {
            double dd = std::fabs(v[i] - q);
            if (dd < bd) {
                bd = dd;
                best = i;
            }
        }
        return best;
    };
    *node = nearest(pl->z1d, z) * pl->NX + nearest(pl->x1d, x);
    return SEM_OK;
} ABI_CATCH

int sem_workspace_bytes(const sem_plan* pl, int batch, int rec_pitch, int with_adjoint, size_t* bytes) try {
    if (!pl || !bytes || batch < 1 || rec_pitch < 0) return fail(SEM_E_INVALID, "bad argument");
    *bytes = workspace_total(pl, batch, rec_pitch, with_adjoint != 0);
    return SEM_OK;
} ABI_CATCH

int sem_plan_set_schedule(sem_plan* pl, int checkpoint_every, int group, int lanes) try {
    if (!pl || lanes < 0) return fail(SEM_E_INVALID, "sem_plan_set_schedule: bad argument");
    pl->d.checkpoint_every = checkpoint_every;
    pl->sched_group = group > 0 ? group : 0;
    pl->sched_lanes = lanes > 0 ? lanes : 1;
    return SEM_OK;
} ABI_CATCH

int sem_plan_set_streaming(sem_plan* pl, int mode) try {
    if (!pl || mode < SEM_STREAM_AUTO || mode > SEM_STREAM_NONTEMPORAL)
        return fail(SEM_E_INVALID, "sem_plan_set_streaming: bad argument");
    pl->streaming = mode;
    return SEM_OK;
} ABI_CATCH

int sem_plan_set_step_fusion(sem_plan* pl, int steps) try {
    if (!pl || steps < 1 || steps > 2) return fail(SEM_E_INVALID, "sem_plan_set_step_fusion: steps must be 1 or 2");
    pl->fusion = steps;
    return SEM_OK;
} ABI_CATCH

int sem_plan_set_corr_chunk(sem_plan* pl, int steps) try {
    if (!pl || steps < 1 || steps > 256) return fail(SEM_E_INVALID, "sem_plan_set_corr_chunk: steps must be in [1, 256]");
    pl->corr_max = steps;
    return SEM_OK;
} ABI_CATCH

int sem_plan_get_corr_chunk(const sem_plan* pl, int* steps) try {
    if (!pl || !steps) return fail(SEM_E_INVALID, "sem_plan_get_corr_chunk: bad argument");
    *steps = pl->corr_max;
    return SEM_OK;
} ABI_CATCH

int sem_plan_set_tile2(sem_plan* pl, int shape) try {
    if (!pl || shape < 0 || shape >= NTILE2) return fail(SEM_E_INVALID, "sem_plan_set_tile2: bad argument");
    if (shape == 3 && pl->d.precision == SEM_FP64)
        return fail(SEM_E_UNSUPPORTED, "two-step tile shape 3 has no FP64 instance");
    if (shape == pl->tile2) return SEM_OK;
    pl->tile2 = shape;
    pl->tx2 = TILE2_EX[shape] - 3;
    pl->tz2 = TILE2_EZ[shape] - 3;
    pl->ntx2 = (pl->d.nelem_x + pl->tx2 - 1) / pl->tx2;
    pl->ntz2 = (pl->d.nelem_z + pl->tz2 - 1) / pl->tz2;
    pl->ntiles2 = pl->ntx2 * pl->ntz2;
    // the two-step receiver table is rebuilt for the new tiles on the next use
    pl->rec2_up_entries = pl->rec2_up_ptrs = pl->rec2_up_sets = 0;
    pl->rec_dirty = true;
    return SEM_OK;
} ABI_CATCH

int sem_plan_get_tile2(const sem_plan* pl, int* shape, int* ex, int* ez) try {
    if (!pl) return fail(SEM_E_INVALID, "sem_plan_get_tile2: null plan");
    if (shape) *shape = pl->tile2;
    if (ex) *ex = TILE2_EX[pl->tile2];
    if (ez) *ez = TILE2_EZ[pl->tile2];
    return SEM_OK;
} ABI_CATCH

int sem_plan_get_step_fusion(const sem_plan* pl, int* steps) try {
    if (!pl || !steps) return fail(SEM_E_INVALID, "sem_plan_get_step_fusion: bad argument");
    *steps = pl->fusion;
    return SEM_OK;
} ABI_CATCH

int sem_plan_get_schedule(const sem_plan* pl, int* checkpoint_every, int* group, int* lanes) try {
    if (!pl) return fail(SEM_E_INVALID, "null plan");
    if (checkpoint_every) *checkpoint_every = pl->d.checkpoint_every;
    if (group) *group = pl->sched_group;
    if (lanes) *lanes = pl->sched_lanes;
    return SEM_OK;
} ABI_CATCH

int sem_plan_schedule_auto(sem_plan* pl, int batch, int rec_pitch, size_t budget_bytes) try {
    if (!pl || batch < 1 || rec_pitch < 0) return fail(SEM_E_INVALID, "sem_plan_schedule_auto: bad argument");
    const size_t bl = batch_layout(pl, batch).total;
    if (budget_bytes <= bl) return fail(SEM_E_NOMEM, "budget below the batch-level workspace");
    const size_t avail = budget_bytes - bl;
    const int nt = pl->d.nt;
    auto bytes = [&](int G, int ck) { return layout(pl, G, rec_pitch, true, ck).total; };
    // 1. the full forward history (no recompute pass): the whole batch, one
    // group per lane, if it fits in memory and its step working set in 3/4 of
    // the last-level cache; else lanes of groups as large as both limits
    // allow, provided the solves in flight keep the GPU busy
    const size_t Ts = pl->d.precision == SEM_FP64 ? 8 : 4;
    const long f_llc = std::max<long>(1, (long)(LLC_BYTES * 3 / 4 / (4 * pl->sb * Ts)));
    const long F = std::min<long>(std::min<long>(batch, f_llc), (long)(avail / bytes(1, 0)));
    if (F >= batch && bytes(batch, 0) <= avail) {
        // one group per lane: the lanes overlap each other's launch ramps and
        // tails (small launches are dispatch-bound)
        const int lanes = std::min(MAX_LANES, batch);
        const int G = (batch + lanes - 1) / lanes;
        if (lanes > 1 && (size_t)lanes * bytes(G, 0) <= avail) return sem_plan_set_schedule(pl, 0, G, lanes);
        return sem_plan_set_schedule(pl, 0, 0, 1);
    }
    if (F >= 1 && F * pl->ntiles >= MIN_INFLIGHT_TILES) {
        const int lanes = (int)std::min<long>(MAX_LANES, F);
        int G = (int)(F / lanes);
        while (G >= 1 && (size_t)lanes * bytes(G, 0) > avail) --G;
        if (G >= 1) return sem_plan_set_schedule(pl, 0, G, lanes);
    }
    // 2. checkpoints every S steps, as many solves at once as fit, one lane
    const int S = pl->ckpt_desc > 0 && pl->ckpt_desc < nt ? pl->ckpt_desc : std::min(100, nt);
    if (bytes(batch, S) <= avail) return sem_plan_set_schedule(pl, S, 0, 1);
    int G = (int)std::min<long>(batch, (long)(avail / bytes(1, S)));
    while (G >= 1 && bytes(G, S) > avail) --G;
    if (G < 1) return fail(SEM_E_NOMEM, "budget below one solve's workspace");
    return sem_plan_set_schedule(pl, S, G, 1);
} ABI_CATCH

int sem_forward(sem_plan* pl, const sem_batch* b, void* ws, size_t ws_bytes, void* stream) try {
    if (!pl) return fail(SEM_E_INVALID, "null plan");
    hipStream_t st = (hipStream_t)stream;
    return pl->d.precision == SEM_FP64 ? forward_impl<double>(pl, b, ws, ws_bytes, st)
                                       : forward_impl<float>(pl, b, ws, ws_bytes, st);
} ABI_CATCH

int sem_forward_adjoint(sem_plan* pl, const sem_batch* b, const double* obs, double inv_sigma2, double* misfit,
                        double* grad_c, double* grad_c2, void* ws, size_t ws_bytes, void* stream) try {
    if (!pl) return fail(SEM_E_INVALID, "null plan");
    hipStream_t st = (hipStream_t)stream;
    return pl->d.precision == SEM_FP64
               ? forward_adjoint_impl<double>(pl, b, obs, inv_sigma2, misfit, grad_c, grad_c2, ws, ws_bytes, st)
               : forward_adjoint_impl<float>(pl, b, obs, inv_sigma2, misfit, grad_c, grad_c2, ws, ws_bytes, st);
} ABI_CATCH

int sem_bspline_map(sem_plan* pl, int nmodel, int nbz, int nbx, const double* coef, const double* base, double* c,
                    void* c2, void* stream) try {
    if (!pl || nmodel < 1 || nmodel > 65535 || nbz < 4 || nbx < 4 || !coef || !c2)
        return fail(SEM_E_INVALID, "sem_bspline_map: bad argument (need >= 4 nodes per axis)");
    int rc = SEM_OK;
    BsplineTables* t = get_bspline(pl, nbz, nbx, rc);
    if (!t) return rc;
    hipStream_t st = (hipStream_t)stream;
    dim3 grid((unsigned)std::min<long>((pl->N + 255) / 256, 4096), nmodel);
    if (pl->d.precision == SEM_FP64)
        bspline_map_kernel<double><<<grid, 256, 0, st>>>(coef, nbz, nbx, base, (int*)t->bz0.p, (double*)t->bzw.p,
                                                         (int*)t->bx0.p, (double*)t->bxw.p, pl->NX, pl->NZ, c,
                                                         (double*)c2);
    else
        bspline_map_kernel<float><<<grid, 256, 0, st>>>(coef, nbz, nbx, base, (int*)t->bz0.p, (double*)t->bzw.p,
                                                        (int*)t->bx0.p, (double*)t->bxw.p, pl->NX, pl->NZ, c,
                                                        (float*)c2);
    LAUNCH_CHECK();
    return SEM_OK;
} ABI_CATCH

}  // extern "C"

// ---------------------------------------------------------------- allocation-free projection
// Scratch of sem_bspline_project_noalloc: the x stage's partial sums
// t[m][jz][q] for as many models as fit, the models processed in chunks of
// that size.  A file-scope device array: every GPU the library runs on holds
// its own copy until the process ends, and every call on that GPU shares it
// (calls that overlap on one GPU race).  Each call writes every cell it then
// reads, so nothing needs clearing.
constexpr size_t SCRATCH_DOUBLES = size_t(1) << 20;  // 8 MiB
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[SCRATCH_MAX_DEVICES] = {};

// The current device's address of `scratch`, looked up once per device (with
// that device current) and kept: host code never passes the array's name.
int scratch_on_current_device(double** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return fail(SEM_E_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SEM_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = (double*)p;
    }
    *out = g_scratch_dev[dev];
    return SEM_OK;
}

}  // namespace

extern "C" {

int sem_bspline_project_noalloc(sem_plan* pl, int nmodel, int nbz, int nbx, const double* g_c, double* g_coef,
                                void* stream) try {
    if (!pl || nmodel < 1 || nmodel > 65535 || nbz < 4 || nbx < 4 || !g_c || !g_coef)
        return fail(SEM_E_INVALID, "sem_bspline_project_noalloc: bad argument");
    if (nbx > 32) return fail(SEM_E_UNSUPPORTED, "sem_bspline_project_noalloc: at most 32 nodes along x");
    // the basis tables are built (and allocated) by the first sem_bspline_map
    // or sem_bspline_project call with these node counts; this call allocates nothing
    auto it = pl->bsp.find(std::make_pair(nbz, nbx));
    if (it == pl->bsp.end())
        return fail(SEM_E_INVALID, "sem_bspline_project_noalloc: no basis tables for " + std::to_string(nbz) + " x " +
                                       std::to_string(nbx) + " nodes yet (call sem_bspline_map or sem_bspline_project "
                                       "with these node counts once first)");
    const size_t per_model = (size_t)pl->NZ * nbx;
    if (per_model > SCRATCH_DOUBLES)
        return fail(SEM_E_NOMEM, "sem_bspline_project_noalloc: one model needs " + std::to_string(per_model * 8) +
                                     " bytes of scratch, more than its " + std::to_string(SCRATCH_DOUBLES * 8));
    double* tmp = nullptr;
    int rc = scratch_on_current_device(&tmp);
    if (rc) return rc;
    const BsplineTables* t = it->second;
    hipStream_t st = (hipStream_t)stream;
    const int chunk = (int)std::min<size_t>((size_t)nmodel, SCRATCH_DOUBLES / per_model);
    for (int m0 = 0; m0 < nmodel; m0 += chunk) {
        const int n = std::min(chunk, nmodel - m0);
        bspline_proj_x_kernel<<<dim3(pl->NZ, n), 256, 256 * nbx * sizeof(double), st>>>(
            g_c + (long)m0 * pl->N, pl->NX, pl->NZ, nbx, (int*)t->bx0.p, (double*)t->bxw.p, tmp);
        LAUNCH_CHECK();
        dim3 g2((unsigned)((nbz * nbx + 255) / 256), n);
        bspline_proj_z_kernel<<<g2, 256, 0, st>>>(tmp, pl->NZ, nbz, nbx, (int*)t->plo.p, (int*)t->phi.p,
                                                  (int*)t->bz0.p, (double*)t->bzw.p, g_coef + (long)m0 * nbz * nbx);
        LAUNCH_CHECK();
    }
    return SEM_OK;
} ABI_CATCH

int sem_bspline_project(sem_plan* pl, int nmodel, int nbz, int nbx, const double* g_c, double* g_coef,
                        void* stream) try {
    if (!pl || nmodel < 1 || nmodel > 65535 || nbz < 4 || nbx < 4 || !g_c || !g_coef)
        return fail(SEM_E_INVALID, "sem_bspline_project: bad argument");
    if (nbx > 32) return fail(SEM_E_UNSUPPORTED, "sem_bspline_project: at most 32 nodes along x");
    int rc = SEM_OK;
    BsplineTables* t = get_bspline(pl, nbz, nbx, rc);
    if (!t) return rc;
    hipStream_t st = (hipStream_t)stream;
    double* tmp = nullptr;
    const size_t bytes = (size_t)nmodel * pl->NZ * nbx * sizeof(double);
    HIP_TRY(hipMallocAsync((void**)&tmp, bytes, st));
    bspline_proj_x_kernel<<<dim3(pl->NZ, nmodel), 256, 256 * nbx * sizeof(double), st>>>(
        g_c, pl->NX, pl->NZ, nbx, (int*)t->bx0.p, (double*)t->bxw.p, tmp);
    LAUNCH_CHECK();
    dim3 g2((unsigned)((nbz * nbx + 255) / 256), nmodel);
    bspline_proj_z_kernel<<<g2, 256, 0, st>>>(tmp, pl->NZ, nbz, nbx, (int*)t->plo.p, (int*)t->phi.p,
                                              (int*)t->bz0.p, (double*)t->bzw.p, g_coef);
    LAUNCH_CHECK();
    HIP_TRY(hipFreeAsync(tmp, st));
    return SEM_OK;
} ABI_CATCH

static int curve_params(const sem_plan* pl, const sem_curve_desc* cd, int nmodel, CurveParams& p) {
    if (!pl || !cd || nmodel < 1 || nmodel > 65535) return fail(SEM_E_INVALID, "sem_curve: bad argument");
    if (cd->kind == SEM_CURVE_CLAMPED6) {
        if (cd->nctrl != 6) return fail(SEM_E_INVALID, "sem_curve: the clamped curve has 6 control points");
    } else if (cd->kind == SEM_CURVE_PERIODIC) {
        if (cd->nctrl < 4 || cd->nctrl > SEM_CURVE_MAX_CTRL)
            return fail(SEM_E_INVALID, "sem_curve: periodic curve needs 4..32 control points");
    } else {
        return fail(SEM_E_INVALID, "sem_curve: unknown curve kind");
    }
    if (cd->samples < 2 || cd->samples > CURVE_MAX_SAMPLES || cd->newton_steps < 0)
        return fail(SEM_E_INVALID, "sem_curve: samples must be in [2, 2048], newton_steps >= 0");
    if (cd->blend != SEM_BLEND_SMOOTH && cd->blend != SEM_BLEND_REPLACE)
        return fail(SEM_E_INVALID, "sem_curve: unknown blend");
    p.kind = cd->kind;
    p.K = cd->nctrl;
    p.S = cd->samples;
    p.newton = cd->newton_steps;
    p.blend = cd->blend;
    p.v_in = cd->v_inside;
    p.v_out = cd->v_outside;
    p.tau = cd->tau;
    p.bg = cd->background;
    return SEM_OK;
}

int sem_curve_velocity_ex(sem_plan* pl, const sem_curve_desc* cd, int nmodel, const double* ctrl, double* c,
                          void* c2, double* aux, void* stream) try {
    CurveParams p;
    int rc = curve_params(pl, cd, nmodel, p);
    if (rc) return rc;
    if (!ctrl || !c2 || !aux) return fail(SEM_E_INVALID, "sem_curve_velocity: null array");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid((unsigned)std::min<long>((pl->N + 255) / 256, 1024), nmodel);
    if (pl->d.precision == SEM_FP64)
        curve_velocity_kernel<double><<<grid, 256, 0, st>>>(p, ctrl, (double*)pl->d_x1d.p, (double*)pl->d_z1d.p,
                                                            pl->NX, pl->NZ, c, (double*)c2, aux);
    else
        curve_velocity_kernel<float><<<grid, 256, 0, st>>>(p, ctrl, (double*)pl->d_x1d.p, (double*)pl->d_z1d.p,
                                                           pl->NX, pl->NZ, c, (float*)c2, aux);
    LAUNCH_CHECK();
    return SEM_OK;
} ABI_CATCH

int sem_curve_project_ex(sem_plan* pl, const sem_curve_desc* cd, int nmodel, const double* dL_dv,
                         const double* aux, double* grad_ctrl, void* stream) try {
    CurveParams p;
    int rc = curve_params(pl, cd, nmodel, p);
    if (rc) return rc;
    if (!dL_dv || !aux || !grad_ctrl) return fail(SEM_E_INVALID, "sem_curve_project: null array");
    hipStream_t st = (hipStream_t)stream;
    const int nblk = (int)std::min<long>((pl->N + CURVE_PROJ_THREADS - 1) / CURVE_PROJ_THREADS, 1024);
    const int w2 = 2 * (p.K + 3);
    double* partial = nullptr;
    HIP_TRY(hipMallocAsync((void**)&partial, sizeof(double) * w2 * nblk * nmodel, st));
    curve_project_kernel<<<dim3(nblk, nmodel), CURVE_PROJ_THREADS, 0, st>>>(p, dL_dv, aux, pl->N, partial);
    LAUNCH_CHECK();
    curve_project_finish_kernel<<<nmodel, 64, 0, st>>>(p, partial, nblk, grad_ctrl);
    LAUNCH_CHECK();
    HIP_TRY(hipFreeAsync(partial, st));
    return SEM_OK;
} ABI_CATCH

int sem_curve_velocity(sem_plan* pl, int nmodel, const double* ctrl, double v_in, double v_out, double tau,
                       int samples, int newton_steps, double* c, void* c2, double* aux, void* stream) try {
    sem_curve_desc cd{SEM_CURVE_CLAMPED6, 6, samples, newton_steps, v_in, v_out, tau, nullptr, SEM_BLEND_SMOOTH};
    return sem_curve_velocity_ex(pl, &cd, nmodel, ctrl, c, c2, aux, stream);
} ABI_CATCH

int sem_curve_project(sem_plan* pl, int nmodel, const double* dL_dv, const double* aux, double v_in, double v_out,
                      double tau, double* grad_ctrl, void* stream) try {
    sem_curve_desc cd{SEM_CURVE_CLAMPED6, 6, 2, 0, v_in, v_out, tau, nullptr, SEM_BLEND_SMOOTH};
    return sem_curve_project_ex(pl, &cd, nmodel, dL_dv, aux, grad_ctrl, stream);
} ABI_CATCH

int sem_forward_snapshots(sem_plan* pl, const sem_batch* b, int every, int ix0, int ix1, int iz0, int iz1,
                          void* snapshots, void* ws, size_t ws_bytes, void* stream) try {
    if (!pl || !snapshots) return fail(SEM_E_INVALID, "sem_forward_snapshots: null argument");
    if (every < 1) return fail(SEM_E_INVALID, "sem_forward_snapshots: every must be >= 1");
    if (ix0 < 0 || ix1 > pl->NX || ix0 >= ix1 || iz0 < 0 || iz1 > pl->NZ || iz0 >= iz1)
        return fail(SEM_E_INVALID, "sem_forward_snapshots: node box outside the grid or empty");
    SnapSpec sp{every, ix0, ix1 - ix0, iz0, iz1 - iz0, snapshots};
    hipStream_t st = (hipStream_t)stream;
    return pl->d.precision == SEM_FP64 ? forward_impl<double>(pl, b, ws, ws_bytes, st, &sp)
                                       : forward_impl<float>(pl, b, ws, ws_bytes, st, &sp);
} ABI_CATCH

int sem_grid_to_nodes(sem_plan* pl, int nxg, int nzg, const double* xg, const double* zg, const double* vp,
                      int method, double* out, void* stream) try {
    if (!pl || !xg || !zg || !vp || !out) return fail(SEM_E_INVALID, "sem_grid_to_nodes: null argument");
    if (nxg < 2 || nzg < 2) return fail(SEM_E_INVALID, "sem_grid_to_nodes: the grid needs >= 2 points per axis");
    if (method != SEM_INTERP_LINEAR && method != SEM_INTERP_NEAREST)
        return fail(SEM_E_INVALID, "sem_grid_to_nodes: unknown method");
    hipStream_t st = (hipStream_t)stream;
    const unsigned blocks = (unsigned)std::min<long>((pl->N + 255) / 256, 2048);
    grid_to_nodes_kernel<<<blocks, 256, 0, st>>>(xg, nxg, zg, nzg, vp, method, (double*)pl->d_x1d.p,
                                                (double*)pl->d_z1d.p, pl->NX, pl->NZ, out);
    LAUNCH_CHECK();
    return SEM_OK;
} ABI_CATCH

int sem_plan_layout(const sem_plan* pl, int* pitch, long* elems) try {
    if (!pl) return fail(SEM_E_INVALID, "null plan");
    if (pitch) *pitch = pl->pitch;
    if (elems) *elems = pl->sb;
    return SEM_OK;
} ABI_CATCH

int sem_plan_gradient_elems(const sem_plan* pl, long* elems) try {
    if (!pl || !elems) return fail(SEM_E_INVALID, "sem_plan_gradient_elems: bad argument");
    *elems = pl->gsb;
    return SEM_OK;
} ABI_CATCH

}  // extern "C"

static int prepare_receivers(sem_plan* pl) {
    if (!pl->rec_dirty) return SEM_OK;
    return pl->d.precision == SEM_FP64 ? upload_receivers<double>(pl) : upload_receivers<float>(pl);
}

template <typename T>
static int slots_corr(sem_plan* pl, const sem_slot* slots, int nslots, int chunk, int chunk_len, hipStream_t st) {
    sem_batch dummy;
    std::memset(&dummy, 0, sizeof(dummy));
    dummy.batch = nslots;
    const StepArgs<T, 4> a = base_args<T>(pl, &dummy, nullptr, nullptr);
    CorrArgs<T, 4> c = corr_args<T>(pl, a, nslots);
    c.slots = slots;
    c.chunk = chunk;
    c.chunk_len = chunk_len;
    return launch_corr<T>(pl, c, st);
}

template <typename T>
static int slots_launch(sem_plan* pl, const sem_slot* slots, int nslots, int m, unsigned mask, int rec_pitch,
                        bool final_record, hipStream_t st) {
    int rc = prepare_receivers(pl);
    if (rc) return rc;
    sem_batch dummy;
    std::memset(&dummy, 0, sizeof(dummy));
    dummy.batch = nslots;
    dummy.rec_pitch = rec_pitch;
    StepArgs<T, 4> a = base_args<T>(pl, &dummy, nullptr, nullptr);
    a.slots = slots;
    a.m = m;
    a.phase_mask = mask;
    a.post = pl->d.record_mode == SEM_RECORD_POST ? 1 : 0;
    a.rec_set_of = nullptr;
    a.c2 = nullptr;
    if (final_record) return launch_step<T, MODE_REC>(pl, a, st, true);
    const bool nts = use_nt(pl, nslots);
    const unsigned fwd = (1u << SEM_SLOT_FORWARD) | (1u << SEM_SLOT_RECOMPUTE) | (1u << SEM_SLOT_FORWARD_HIST);
    const unsigned adj = 1u << SEM_SLOT_ADJOINT;
    if ((mask & fwd) && (mask & adj)) {
        // one launch for both directions: a second launch per step costs a
        // kernel boundary (dispatch, drain, cache release) even where none of
        // its workgroups has work; the per-workgroup branch costs nothing
        a.damp = 1;
        a.damp_adj = pl->d.adjoint_mode == SEM_ADJ_EXACT ? 1 : 0;
        return launch_step<T, MODE_ANY>(pl, a, st, nts);
    }
    if (mask & fwd) {
        a.damp = 1;
        if ((rc = launch_step<T, MODE_FWD>(pl, a, st, nts))) return rc;
    }
    if (mask & adj) {
        a.damp = pl->d.adjoint_mode == SEM_ADJ_EXACT ? 1 : 0;
        if ((rc = launch_step<T, MODE_ADJS>(pl, a, st, nts))) return rc;
    }
    return SEM_OK;
}

template <typename T>
static int slots_launch2(sem_plan* pl, const sem_slot* slots, int nslots, int m, unsigned mask, int rec_pitch,
                         hipStream_t st) {
    int rc = prepare_receivers(pl);
    if (rc) return rc;
    sem_batch dummy;
    std::memset(&dummy, 0, sizeof(dummy));
    dummy.batch = nslots;
    dummy.rec_pitch = rec_pitch;
    StepArgs<T, 4> a = base_args<T>(pl, &dummy, nullptr, nullptr);
    a.slots = slots;
    a.m = m;
    a.post = pl->d.record_mode == SEM_RECORD_POST ? 1 : 0;
    a.rec_set_of = nullptr;
    a.c2 = nullptr;
    const bool nts = use_nt(pl, nslots);
    const unsigned fwd = (1u << SEM_SLOT_FORWARD) | (1u << SEM_SLOT_FORWARD_HIST) | (1u << SEM_SLOT_RECOMPUTE);
    const unsigned adj = 1u << SEM_SLOT_ADJOINT;
    if ((mask & fwd) && (mask & adj)) {
        a.phase_mask = mask & (fwd | adj);
        a.damp = 1;
        a.damp_adj = pl->d.adjoint_mode == SEM_ADJ_EXACT ? 1 : 0;
        return launch_step2<T, MODE_ANY>(pl, a, st, nts);
    }
    if (mask & fwd) {
        a.phase_mask = mask & fwd;
        a.damp = 1;
        if ((rc = launch_step2<T, MODE_FWD>(pl, a, st, nts))) return rc;
    }
    if (mask & adj) {
        a.phase_mask = adj;
        a.damp = pl->d.adjoint_mode == SEM_ADJ_EXACT ? 1 : 0;
        if ((rc = launch_step2<T, MODE_ADJS>(pl, a, st, nts))) return rc;
    }
    return SEM_OK;
}

extern "C" {

int sem_slots_step2(sem_plan* pl, const sem_slot* slots, int nslots, int m, unsigned mask, int rec_pitch,
                    void* stream) try {
    if (!pl || !slots || nslots < 1 || m < 0 || (m & 1) || rec_pitch < 0)
        return fail(SEM_E_INVALID, "sem_slots_step2: bad argument (m must be even)");
    if ((long)std::max(std::max(pl->ntiles, pl->ntiles2), pl->ctiles) * nslots >= (1L << 31) - 64)
        return fail(SEM_E_INVALID, "too many slots for this mesh");
    hipStream_t st = (hipStream_t)stream;
    return pl->d.precision == SEM_FP64 ? slots_launch2<double>(pl, slots, nslots, m, mask, rec_pitch, st)
                                       : slots_launch2<float>(pl, slots, nslots, m, mask, rec_pitch, st);
} ABI_CATCH

int sem_slots_step(sem_plan* pl, const sem_slot* slots, int nslots, int m, unsigned mask, int rec_pitch,
                   void* stream) try {
    if (!pl || !slots || nslots < 1 || m < 0 || rec_pitch < 0) return fail(SEM_E_INVALID, "sem_slots_step: bad argument");
    if ((long)std::max(pl->ntiles, pl->ctiles) * nslots >= (1L << 31) - 64)
        return fail(SEM_E_INVALID, "too many slots for this mesh");
    hipStream_t st = (hipStream_t)stream;
    return pl->d.precision == SEM_FP64 ? slots_launch<double>(pl, slots, nslots, m, mask, rec_pitch, false, st)
                                       : slots_launch<float>(pl, slots, nslots, m, mask, rec_pitch, false, st);
} ABI_CATCH

int sem_slots_correlate(sem_plan* pl, const sem_slot* slots, int nslots, int chunk, int chunk_len, void* stream) try {
    if (!pl || !slots || nslots < 1 || chunk < 0 || chunk_len < 1)
        return fail(SEM_E_INVALID, "sem_slots_correlate: bad argument");
    if ((long)std::max(pl->ntiles, pl->ctiles) * nslots >= (1L << 31) - 64)
        return fail(SEM_E_INVALID, "too many slots for this mesh");
    hipStream_t st = (hipStream_t)stream;
    return pl->d.precision == SEM_FP64 ? slots_corr<double>(pl, slots, nslots, chunk, chunk_len, st)
                                       : slots_corr<float>(pl, slots, nslots, chunk, chunk_len, st);
} ABI_CATCH

int sem_slots_record_final(sem_plan* pl, const sem_slot* slots, int nslots, unsigned mask, int rec_pitch,
                           void* stream) try {
    if (!pl || !slots || nslots < 1) return fail(SEM_E_INVALID, "sem_slots_record_final: bad argument");
    hipStream_t st = (hipStream_t)stream;
    return pl->d.precision == SEM_FP64 ? slots_launch<double>(pl, slots, nslots, 0, mask, rec_pitch, true, st)
                                       : slots_launch<float>(pl, slots, nslots, 0, mask, rec_pitch, true, st);
} ABI_CATCH

int sem_receivers_reserve(sem_plan* pl, long max_entries, int max_sets) try {
    if (!pl || max_entries < 0 || max_sets < 0) return fail(SEM_E_INVALID, "sem_receivers_reserve: bad argument");
    const size_t eb = pl->d.precision == SEM_FP64 ? sizeof(RecEntry<double, 4>) : sizeof(RecEntry<float, 4>);
    bool m1 = false, m2 = false, m3 = false, m4 = false;
    int rc = pl->d_recs.reserve(std::max<size_t>(max_entries * eb, 16), m1);
    if (!rc) rc = pl->d_rec_ptr.reserve((size_t)max_sets * (pl->ntiles + 1) * sizeof(int), m2);
    // the two-step tiles' table: an element lies in at most two tiles along
    // each axis there too, so the same entry bound holds
    if (!rc) rc = pl->d_recs2.reserve(std::max<size_t>(max_entries * eb, 16), m3);
    if (!rc) rc = pl->d_rec_ptr2.reserve((size_t)max_sets * (pl->ntiles2 + 1) * sizeof(int), m4);
    if (rc) return rc;
    if (m1 || m2) {
        pl->rec_up_entries = pl->rec_up_ptrs = pl->rec_up_sets = 0;
        pl->rec_dirty = true;
    }
    if (m3 || m4) {
        pl->rec2_up_entries = pl->rec2_up_ptrs = pl->rec2_up_sets = 0;
        pl->rec_dirty = true;
    }
    return prepare_receivers(pl);
} ABI_CATCH

int sem_receivers_sync(sem_plan* pl) try {
    if (!pl) return fail(SEM_E_INVALID, "null plan");
    return prepare_receivers(pl);
} ABI_CATCH

int sem_pad_nodal(sem_plan* pl, const void* src, void* dst, int nmodel, void* stream) try {
    if (!pl || !src || !dst || nmodel < 1 || nmodel > 65535) return fail(SEM_E_INVALID, "sem_pad_nodal: bad argument");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid((unsigned)std::min<long>((pl->N + 255) / 256, 2048), nmodel);
    if (pl->d.precision == SEM_FP64)
        pad_rows_kernel<double><<<grid, 256, 0, st>>>((const double*)src, (double*)dst, pl->NX, pl->NZ, pl->pitch, pl->sb);
    else
        pad_rows_kernel<float><<<grid, 256, 0, st>>>((const float*)src, (float*)dst, pl->NX, pl->NZ, pl->pitch, pl->sb);
    LAUNCH_CHECK();
    return SEM_OK;
} ABI_CATCH

int sem_residual(sem_plan* pl, const void* seis, const double* obs, void* res, double* misfit, int nrec, int nt,
                 int rec_pitch, void* stream) try {
    if (!pl || !seis || !obs || !res || !misfit || nrec < 0 || nrec > rec_pitch || nt < 1)
        return fail(SEM_E_INVALID, "sem_residual: bad argument");
    hipStream_t st = (hipStream_t)stream;
    if (pl->d.precision == SEM_FP64)
        residual_kernel<double><<<1, RESID_THREADS, 0, st>>>((const double*)seis, (const double*)obs, (double*)res, misfit,
                                                   nullptr, nrec, nt, rec_pitch, (long)nt * rec_pitch, 0, nt, 1);
    else
        residual_kernel<float><<<1, RESID_THREADS, 0, st>>>((const float*)seis, (const double*)obs, (float*)res, misfit,
                                                  nullptr, nrec, nt, rec_pitch, (long)nt * rec_pitch, 0, nt, 1);
    LAUNCH_CHECK();
    return SEM_OK;
} ABI_CATCH

int sem_gradient_finalize(sem_plan* pl, const double* G, const void* c2p, double scale, double* grad_c,
                          double* grad_c2, void* stream) try {
    if (!pl || !G || !c2p) return fail(SEM_E_INVALID, "sem_gradient_finalize: bad argument");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid((unsigned)std::min<long>((pl->N + 255) / 256, 2048), 1);
    if (pl->d.precision == SEM_FP64)
        grad_finalize_kernel<double><<<grid, 256, 0, st>>>(G, pl->gsb, pl->sb, pl->gplane, pl->nexp, pl->d.nelem_x,
                                                           pl->d.nelem_z, pl->NX, pl->NZ, pl->pitch,
                                                           (const double*)c2p, nullptr, pl->J * scale, grad_c, grad_c2);
    else
        grad_finalize_kernel<float><<<grid, 256, 0, st>>>(G, pl->gsb, pl->sb, pl->gplane, pl->nexp, pl->d.nelem_x,
                                                          pl->d.nelem_z, pl->NX, pl->NZ, pl->pitch,
                                                          (const float*)c2p, nullptr, pl->J * scale, grad_c, grad_c2);
    LAUNCH_CHECK();
    return SEM_OK;
} ABI_CATCH

}  // extern "C"

// ---- device-resident request loop ------------------------------------------

struct sem_loop {
    sem_plan* pl = nullptr;
    int dev = -1, rec_pitch = 0;
    LoopParams p{};
    void* dmem = nullptr;
    LoopHost* host = nullptr;  // host address of the mapped block
    long long pushed = 0, rounds = 0;
    ~sem_loop() {
        if (dmem) (void)hipFree(dmem);
        if (host) (void)hipHostFree(host);
    }
};

namespace {

// Adjoint steps per correlation chunk of a loop: at most 50, the segment split
// into near-equal chunks, rounded up to an even length when shorter than the
// segment (two-step launches never straddle an even chunk boundary); the ring
// then holds chunk + 1 states per request.
int loop_chunk(int S) {
    const int n = (S + 49) / 50;
    const int c = (S + n - 1) / n;
    return (c % 2 && c < S) ? c + 1 : c;
}

template <typename T>
int loop_round(sem_loop* lp, hipStream_t st) {
    sem_plan* pl = lp->pl;
    const LoopParams& p = lp->p;
    const int C = p.nslots;
    const unsigned FWD = 1u << SEM_SLOT_FORWARD, RC = 1u << SEM_SLOT_RECOMPUTE, ADJ = 1u << SEM_SLOT_ADJOINT,
                   FH = 1u << SEM_SLOT_FORWARD_HIST;
    int rc;
    loop_begin_kernel<<<1, LOOP_MAX_SLOTS, 0, st>>>(p);
    LAUNCH_CHECK();
    for (int m = 0; m < p.S; ++m) {
        if (p.two_step) {
            if (m % 2 == 0 && (rc = slots_launch2<T>(pl, p.table, C, m, FWD | FH | RC | ADJ, lp->rec_pitch, st)))
                return rc;
        } else if ((rc = slots_launch<T>(pl, p.table, C, m, FWD | RC | FH | ADJ, lp->rec_pitch, false, st))) {
            return rc;
        }
        if ((m + 1) % p.chunk == 0 || m == p.S - 1)
            if ((rc = slots_corr<T>(pl, p.table, C, m / p.chunk, p.chunk, st))) return rc;
    }
    if ((rc = slots_launch<T>(pl, p.ftable, C, 0, FWD | FH, lp->rec_pitch, true, st))) return rc;
    loop_residual_kernel<T><<<C, RESID_THREADS, 0, st>>>(p, lp->rec_pitch);
    LAUNCH_CHECK();
    dim3 grid((unsigned)std::min<long>((pl->N + 255) / 256, 2048), C);
    loop_finalize_kernel<T><<<grid, 256, 0, st>>>(p, pl->sb, pl->gplane, pl->nexp, pl->d.nelem_x, pl->d.nelem_z,
                                                  pl->NX, pl->NZ, pl->pitch, pl->J);
    LAUNCH_CHECK();
    loop_advance_kernel<<<1, 64, 0, st>>>(p);
    LAUNCH_CHECK();
    return SEM_OK;
}

int loop_device_check(const sem_loop* lp) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != lp->dev) return fail(SEM_E_INVALID, "a loop runs on the device it was created on");
    return SEM_OK;
}

}  // namespace

extern "C" {

int sem_loop_create(sem_plan* pl, int nslots, int segment, int two_step, int rec_pitch, int queue_capacity,
                    sem_loop** out) try {
    if (!pl || !out || nslots < 1 || nslots > LOOP_MAX_SLOTS || segment < 1 || rec_pitch < 1 || queue_capacity < 1)
        return fail(SEM_E_INVALID, "sem_loop_create: bad argument");
    *out = nullptr;
    const int chunk = loop_chunk(segment);
    if (two_step && (segment % 2 || chunk % 2))
        return fail(SEM_E_INVALID, "sem_loop_create: two steps per launch need an even segment and correlation chunk");
    if ((long)std::max(std::max(pl->ntiles, pl->ntiles2), pl->ctiles) * nslots >= (1L << 31) - 64)
        return fail(SEM_E_INVALID, "too many slots for this mesh");
    std::unique_ptr<sem_loop> lp(new sem_loop);
    lp->pl = pl;
    lp->rec_pitch = rec_pitch;
    HIP_TRY(hipGetDevice(&lp->dev));
    LoopParams& p = lp->p;
    p.nslots = nslots;
    p.S = segment;
    p.two_step = two_step ? 1 : 0;
    p.chunk = chunk;
    p.qcap = queue_capacity;
    p.elem = pl->d.precision == SEM_FP64 ? 8 : 4;
    p.E = pl->sb;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t bq = up(sizeof(sem_request) * queue_capacity), bs = up(sizeof(LoopSlot) * nslots),
                 bt = up(sizeof(sem_slot) * nslots), bc = up(sizeof(LoopCtl));
    const size_t total = bq + bs + 2 * bt + bc;
    hipError_t e = hipMalloc(&lp->dmem, total);
    if (e != hipSuccess) {
        lp->dmem = nullptr;
        return fail(SEM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    HIP_TRY(hipMemset(lp->dmem, 0, total));
    char* b = (char*)lp->dmem;
    p.q = (sem_request*)b;
    p.st = (LoopSlot*)(b + bq);
    p.table = (sem_slot*)(b + bq + bs);
    p.ftable = (sem_slot*)(b + bq + bs + bt);
    p.ctl = (LoopCtl*)(b + bq + bs + 2 * bt);
    void* h = nullptr;
    e = hipHostMalloc(&h, sizeof(LoopHost), hipHostMallocMapped | hipHostMallocCoherent);
    if (e != hipSuccess) return fail(SEM_E_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    lp->host = (LoopHost*)h;
    std::memset(h, 0, sizeof(LoopHost));
    void* dh = nullptr;
    HIP_TRY(hipHostGetDevicePointer(&dh, h, 0));
    p.host = (LoopHost*)dh;
    HIP_TRY(hipDeviceSynchronize());  // the cleared state is in place before any stream uses it
    *out = lp.release();
    return SEM_OK;
} ABI_CATCH

int sem_loop_destroy(sem_loop* lp) try {
    delete lp;
    return SEM_OK;
} ABI_CATCH

int sem_loop_chunk(const sem_loop* lp, int* chunk) try {
    if (!lp || !chunk) return fail(SEM_E_INVALID, "sem_loop_chunk: bad argument");
    *chunk = lp->p.chunk;
    return SEM_OK;
} ABI_CATCH

int sem_loop_push(sem_loop* lp, const sem_request* r, void* stream) try {
    if (!lp || !r) return fail(SEM_E_INVALID, "sem_loop_push: bad argument");
    const sem_plan* pl = lp->pl;
    const LoopParams& p = lp->p;
    if (r->nt < 1 || r->nt > pl->d.nt || r->nrec < 0 || r->nrec > lp->rec_pitch || r->src_node < 0 ||
        r->src_node >= pl->sb || r->rec_set < 0 || r->rec_set >= (int)pl->sets.size())
        return fail(SEM_E_INVALID, "sem_loop_push: bad request (nt, nrec, source node or receiver set)");
    bool ok = r->c2 && r->wavelet && r->seis && r->resid && r->obs && r->misfit && r->G && r->grad_c && r->w && r->om &&
              r->ring;
    if (p.two_step) ok = ok && r->w2 && r->om2 && r->w2 != r->w && r->om2 != r->om;
    if (r->full_history)
        ok = ok && r->hist;
    else
        ok = ok && r->u0 && r->u1 && r->cku && r->ckw && r->seg;
    if (!ok) return fail(SEM_E_INVALID, "sem_loop_push: a buffer the request needs is missing");
    int rc = loop_device_check(lp);
    if (rc) return rc;
    const long long nd = __atomic_load_n(&lp->host->ndone, __ATOMIC_ACQUIRE);
    if (lp->pushed - nd >= p.qcap)
        return fail(SEM_E_INVALID, "sem_loop_push: queue full (requests pushed and not completed)");
    loop_push_kernel<<<1, 64, 0, (hipStream_t)stream>>>(p, *r);
    LAUNCH_CHECK();
    ++lp->pushed;
    return SEM_OK;
} ABI_CATCH

int sem_loop_round(sem_loop* lp, void* stream) try {
    if (!lp) return fail(SEM_E_INVALID, "sem_loop_round: bad argument");
    int rc = loop_device_check(lp);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    rc = lp->pl->d.precision == SEM_FP64 ? loop_round<double>(lp, st) : loop_round<float>(lp, st);
    if (!rc) ++lp->rounds;
    return rc;
} ABI_CATCH

int sem_loop_poll(sem_loop* lp, long long since, long long* ids, long long* rounds, int max, long long* total) try {
    if (!lp || !total || max < 0 || (max > 0 && !ids)) return fail(SEM_E_INVALID, "sem_loop_poll: bad argument");
    const long long nd = __atomic_load_n(&lp->host->ndone, __ATOMIC_ACQUIRE);
    if (since < 0 || since > nd) return fail(SEM_E_INVALID, "sem_loop_poll: `since` beyond the completions");
    const long long n = std::min<long long>(max, nd - since);
    if (nd - since > LOOP_DONE_CAP && n > 0)
        return fail(SEM_E_INVALID, "sem_loop_poll: completion records overwritten (poll more often)");
    for (long long i = 0; i < n; ++i) {
        const long long k = (since + i) % LOOP_DONE_CAP;
        ids[i] = lp->host->ids[k];
        if (rounds) rounds[i] = lp->host->id_round[k];
    }
    *total = nd;
    return SEM_OK;
} ABI_CATCH

int sem_loop_counts(const sem_loop* lp, long long* pushed, long long* rounds_enqueued, long long* rounds_done,
                    int* active) try {
    if (!lp) return fail(SEM_E_INVALID, "sem_loop_counts: bad argument");
    if (pushed) *pushed = lp->pushed;
    if (rounds_enqueued) *rounds_enqueued = lp->rounds;
    if (rounds_done) *rounds_done = __atomic_load_n(&lp->host->round, __ATOMIC_ACQUIRE);
    if (active) *active = __atomic_load_n(&lp->host->active, __ATOMIC_ACQUIRE);
    return SEM_OK;
} ABI_CATCH

}  // extern "C"
