/*
 * sem_capi.h -- C ABI of the MI355X SEM forward/adjoint library (libmun_sem.so).
 *
 * 2-D acoustic spectral-element solver on a structured mesh of order-p GLL
 * elements (p = 4), the one Bspline_Coupling/SEM/sem_waveform/core.py
 * implements on the CPU.  The reference is pure Python and has no C ABI; each
 * entry point below states which reference call it replaces.
 *
 * Conventions
 *   - Plain C types only.  Every function returns an int status: SEM_OK (0) or
 *     a negative SEM_E* code; sem_last_error() gives a thread-local message.
 *   - Bulk arrays are DEVICE pointers owned by the caller (e.g. torch tensor
 *     data_ptr()).  The plan owns only its constant tables.  Scratch memory is
 *     a caller-provided workspace sized by sem_workspace_bytes().
 *   - State precision is fixed per plan (SEM_FP32 or SEM_FP64); arrays typed
 *     "real" below are float for SEM_FP32 plans and double for SEM_FP64 plans.
 *     Gradients and misfits are always double.
 *   - Work is enqueued asynchronously on the given stream (NULL = default).
 *     A plan may be used by one stream at a time; distinct plans are
 *     independent.  Receiver sets must be added before the first solve.
 *   - Node numbering is lexicographic on the (NZ x NX) GLL grid, x fastest:
 *     node = jz * NX + ix, NX = nelem_x*p + 1, NZ = nelem_z*p + 1.
 */
#ifndef MUN_SEM_CAPI_H
#define MUN_SEM_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEM_OK 0
#define SEM_E_INVALID (-1)     /* bad argument / descriptor                     */
#define SEM_E_UNSUPPORTED (-2) /* e.g. polynomial order without a kernel        */
#define SEM_E_NOMEM (-3)       /* workspace too small / allocation failed       */
#define SEM_E_DEVICE (-4)      /* HIP runtime error                             */
#define SEM_E_NONFINITE (-5)   /* NaN/Inf in an input model                     */

#define SEM_FP32 0
#define SEM_FP64 1

/* adjoint recurrence: the reference's (sponge ignored in the adjoint,
 * core.py:1252-1263) or the exact discrete adjoint of the damped forward. */
#define SEM_ADJ_REFERENCE 0
#define SEM_ADJ_EXACT 1

/* seismogram sampling: R u^{n+1} after the update (run_forward_and_adjoint,
 * core.py:1211,1218) or R u^n before it (run(), core.py:633-639). */
#define SEM_RECORD_POST 0
#define SEM_RECORD_PRE 1

/* absorbing (sponge) sides, bitmask; the reference damps all four
 * (boundary.py:38-48). */
#define SEM_SIDE_XMIN 1u
#define SEM_SIDE_XMAX 2u
#define SEM_SIDE_ZMIN 4u
#define SEM_SIDE_ZMAX 8u

typedef struct sem_plan sem_plan;

typedef struct {
    int nelem_x, nelem_z;      /* elements per axis                                */
    int order;                 /* polynomial order p (kernels: 4)                  */
    double xmin, xmax, zmin, zmax;
    double dt;                 /* time step [s], used as given (no CFL shrink)    */
    int nt;                    /* number of time steps                             */
    double sponge_width;       /* [m]   (config method.pml_thickness)              */
    double sponge_cmax;        /* [m/s] c_max in d0 (boundary.py:30)               */
    double sponge_R;           /* reflection target R (boundary.py:5, 1e-8)        */
    unsigned sponge_sides;     /* SEM_SIDE_* mask                                  */
    int precision;             /* SEM_FP32 | SEM_FP64                              */
    int adjoint_mode;          /* SEM_ADJ_REFERENCE | SEM_ADJ_EXACT                */
    int record_mode;           /* SEM_RECORD_POST | SEM_RECORD_PRE                 */
    int checkpoint_every;      /* steps per wavefield checkpoint; <=0 or >=nt:
                                  keep the full forward history instead           */
    double lagrange_eps;       /* +eps in Lagrange denominators (utils.py:36), 1e-12 */
    int tile_rows;             /* step-kernel tile height in element rows incl. the
                                  halo row: 15, 7 or 3 (0 = 15); see sem_autotune  */
    /* Misfit window (CO2_plume/one_stage/single_cubic/sem_waveform/core.py:1433-1520):
     * only samples t in [win_begin, win_end) with (t - win_begin) % win_decim == 0
     * enter the misfit and the adjoint source.  win_end = 0 means nt, win_decim
     * <= 1 means every sample.  SEM_ADJ_REFERENCE then runs the adjoint over the
     * window only, t = win_end-1 .. win_begin, as the reference does (its
     * gradient leaves out t < win_begin); SEM_ADJ_EXACT runs it back to t = 0. */
    int win_begin, win_end, win_decim;
} sem_desc;

/* One batch of independent solves on a plan's mesh.  Solve b uses model
 * model_of[b] (or b), source node src_node[b], wavelet row b*wavelet_stride and
 * receiver set rec_set[b].  Seismograms / observations are [batch][nt][rec_pitch]
 * with rec_pitch >= the largest receiver count of the sets used. */
typedef struct {
    int batch;
    int nmodel;                /* rows of c2                                       */
    const void* c2;            /* device real [nmodel][N]  squared velocity        */
    const int* model_of;       /* device int [batch] or NULL (identity)            */
    const void* wavelet;       /* device real [.][nt]  source time function        */
    long wavelet_stride;       /* elements between solves' wavelets (0 = shared)   */
    const int* src_node;       /* device int [batch]  injection node ids           */
    const int* rec_set;        /* HOST int [batch]  receiver set id per solve      */
    int rec_pitch;             /* row length of seis/obs                           */
    void* seis;                /* device real [batch][nt][rec_pitch]  output       */
} sem_batch;

/* Replaces SEMSimulation.__init__'s mesh/operator/sponge set-up
 * (core.py:151-325, mesh.py:4-43, operators.py:6-42, boundary.py:5-58). */
int sem_plan_create(const sem_desc* desc, sem_plan** out);
int sem_plan_destroy(sem_plan* plan);
/* NX, NZ, N of the plan's grid. */
int sem_plan_dims(const sem_plan* plan, int* NX, int* NZ, long* N);
/* 1-D node coordinates (host double [NX], [NZ]). */
int sem_plan_axes(const sem_plan* plan, double* x1d, double* z1d);

/* Register receivers at arbitrary positions (host double [nrec][2], x then z);
 * replaces setup_receivers / _setup_receivers_from_positions
 * (receivers.py:6-80, CO2_plume/2D_model/inversion/sem_waveform/core.py:90-172). */
int sem_receiver_set_add(sem_plan* plan, int nrec, const double* xz, int* set_id);
int sem_receiver_set_size(const sem_plan* plan, int set_id, int* nrec);

/* Nearest GLL node to (x, z); replaces find_closest_node (mesh.py:45-54). */
int sem_source_node(const sem_plan* plan, double x, double z, int* node);

/* Workspace bytes for sem_forward (with_adjoint = 0) or sem_forward_adjoint
 * (with_adjoint = 1) at the given batch size and receiver pitch, under the
 * plan's current schedule. */
int sem_workspace_bytes(const sem_plan* plan, int batch, int rec_pitch, int with_adjoint, size_t* bytes);

/* Schedule of sem_forward_adjoint (no reference counterpart: the reference
 * solves one model at a time, core.py:1067-1371).  A batch of more than
 * `group` solves runs as groups of consecutive solves (sizes differing by at
 * most one), `lanes` groups at a time on streams the plan owns, each lane with
 * a workspace region of its own; the call still enqueues on the caller's
 * stream (the lanes start after it and the stream waits for them).
 * checkpoint_every as in sem_desc (<= 0: full forward history, no recompute).
 * group <= 0 runs the whole batch at once (the default).  Results are
 * bitwise identical under every schedule; the workspace size changes. */
int sem_plan_set_schedule(sem_plan* plan, int checkpoint_every, int group, int lanes);
int sem_plan_get_schedule(const sem_plan* plan, int* checkpoint_every, int* group, int* lanes);
/* Choose and set a schedule for `batch` solves whose workspace fits in
 * budget_bytes: the full forward history for the whole batch, split into one
 * group per lane (up to 3 lanes), if it fits and the solves' per-step working
 * set (u^n, u^{n+1}, w, c^2) fits in 3/4 of the GPU's 256 MB last-level
 * cache; else 3 lanes of full-history groups as large as both limits allow
 * (when they keep the GPU busy); else checkpoints (the plan's
 * checkpoint_every, or 100 steps) for as many solves at once as fit.
 * SEM_E_NOMEM if not even one solve fits. */
int sem_plan_schedule_auto(sem_plan* plan, int batch, int rec_pitch, size_t budget_bytes);

/* Memory hints of the step kernels' once-per-step streams (the increment and
 * the new state): SEM_STREAM_AUTO (default) uses plain loads and stores when
 * the solves in flight touch at most the 256 MB last-level cache per step,
 * non-temporal ones otherwise; the other two force one kind (for measurement).
 * Results do not depend on it.  No reference counterpart. */
#define SEM_STREAM_AUTO 0
#define SEM_STREAM_PLAIN 1
#define SEM_STREAM_NONTEMPORAL 2
int sem_plan_set_streaming(sem_plan* plan, int mode);

/* Time steps per step-kernel launch in sem_forward and sem_forward_adjoint:
 * 2 (default; a two-step kernel keeps the intermediate field in LDS, so a
 * pair of steps reads the state once, and stores it with non-temporal
 * writes), or 1.  Two-step launches cover the forward pass (full history,
 * checkpointed, forward-only or with snapshots), the recompute and the
 * adjoint; sem_slots_step takes one step per launch (sem_slots_step2 two),
 * and so do the odd last step of a pass and a step pair that would straddle
 * a checkpoint or hold a snapshot in its middle.  Results are
 * bitwise the same for 1 and 2.  No reference counterpart (the
 * reference loops over single steps, core.py:1197-1275). */
int sem_plan_set_step_fusion(sem_plan* plan, int steps);
int sem_plan_get_step_fusion(const sem_plan* plan, int* steps);

/* Tile shape of the two-step kernel: shape 0 (default) 19 x 26 element
 * threads per workgroup (16 x 23 owned), 1: 19 x 13 (16 x 10), 2: 35 x 14
 * (32 x 11), 3: 35 x 28 (32 x 25; FP32 plans only), 4: 23 x 11 (20 x 8).
 * Results are bitwise the
 * same for every shape.  The two-step receiver table is rebuilt on the next
 * use: set the shape before sem_receivers_reserve and not while solves of the
 * plan are in flight.  No reference counterpart (chosen by the autotuner). */
/* Adjoint steps per gradient-correlation chunk of sem_forward_adjoint, at
 * most (default 64; 1 .. 256): a segment is split into equal chunks of at
 * most this many steps, whose adjoint states are kept in a ring of chunk + 1
 * states per solve, and the FP64 accumulator is read and written once per
 * chunk.  Results are bitwise the same for every value; the workspace size
 * changes.  No reference counterpart (chosen by the autotuner). */
int sem_plan_set_corr_chunk(sem_plan* plan, int steps);
int sem_plan_get_corr_chunk(const sem_plan* plan, int* steps);

#define SEM_TILE2_SHAPES 5
int sem_plan_set_tile2(sem_plan* plan, int shape);
int sem_plan_get_tile2(const sem_plan* plan, int* shape, int* ex, int* ez);

/* Forward solve of a batch; replaces run() / _time_stepping_loop_torch
 * (core.py:334-435, 598-659).  Writes b->seis. */
int sem_forward(sem_plan* plan, const sem_batch* b, void* workspace, size_t ws_bytes, void* stream);

/* Forward + misfit + adjoint + gradient; replaces the node-level part of
 * run_forward_and_adjoint (core.py:1067-1278) including
 * accumulate_sensitivity_dJdc_torch (core.py:49-85).
 *   obs      device double [batch][nt][rec_pitch] observed data y_obs; the
 *            residual d_syn - y_obs is formed in double, as core.py:1224-1227
 *            forms it (the adjoint source then keeps the state precision)
 *   misfit   device double [batch]               0.5 * ||d_syn - y_obs||^2
 *   grad_c   device double [batch][N] or NULL     d loglik / d c at the nodes,
 *            loglik = -misfit * inv_sigma2 (core.py:1230; inv_sigma2 =
 *            1/(noise_std^2 + 1e-24) as core.py:1086)
 *   grad_c2  device double [batch][N] or NULL     d loglik / d (c^2)
 * The gradient-time integral is accumulated in double.  With both gradient
 * outputs NULL only the forward pass and the misfit run. */
int sem_forward_adjoint(sem_plan* plan, const sem_batch* b, const double* obs, double inv_sigma2,
                        double* misfit, double* grad_c, double* grad_c2,
                        void* workspace, size_t ws_bytes, void* stream);

/* Forward solve that also keeps wavefield snapshots of a node-index box:
 * snapshot k is u after step k*every (u^{k*every+1}), nodes ix0 <= ix < ix1,
 * iz0 <= jz < iz1, stored [batch][nsnap][iz1-iz0][ix1-ix0] (real), nsnap =
 * ceil(nt / every).  Replaces the save_wavefield path of run()
 * (Bspline_Coupling/SEM/sem_waveform/core.py:318-325, 400-401, 649-651): the
 * reference's ROI mask selects the nodes with x, z inside the ROI rectangle,
 * which on the structured grid is this box.  Seismograms as sem_forward. */
int sem_forward_snapshots(sem_plan* plan, const sem_batch* b, int every, int ix0, int ix1, int iz0, int iz1,
                          void* snapshots, void* workspace, size_t ws_bytes, void* stream);

/* Background velocity from a regular grid onto the nodes: vp [nzg][nxg] on
 * increasing axes xg [nxg], zg [nzg], node coordinates clipped into the grid,
 * bilinear (SEM_INTERP_LINEAR) or nearest grid value (SEM_INTERP_NEAREST,
 * ties to the lower index).  Replaces build_background_from_grid
 * (CO2_plume/one_stage/single_cubic/sem_waveform/velocity.py:212-274, scipy
 * RegularGridInterpolator with clip=True).  All arrays device double; out [N]. */
#define SEM_INTERP_LINEAR 0
#define SEM_INTERP_NEAREST 1
int sem_grid_to_nodes(sem_plan* plan, int nxg, int nzg, const double* xg, const double* zg, const double* vp,
                      int method, double* out, void* stream);

/* Tensor-product cubic B-spline velocity nodes (clamped uniform knots on the
 * domain; no reference counterpart -- nearest analogue is the gridded
 * background of CO2_plume/one_stage/single_cubic/sem_waveform/velocity.py:212-274):
 *   c[m] = base + Bz coef[m] Bx^T,   c2 = c^2 in the plan's precision.
 * coef: device double [nmodel][nbz][nbx]; base: device double [N] or NULL;
 * c: device double [nmodel][N] (may be NULL); c2: device real [nmodel][N]. */
int sem_bspline_map(sem_plan* plan, int nmodel, int nbz, int nbx, const double* coef,
                    const double* base, double* c, void* c2, void* stream);
/* Chain rule onto the nodes: g_coef[m] = Bz^T g_c[m] Bx. */
int sem_bspline_project(sem_plan* plan, int nmodel, int nbz, int nbx, const double* g_c,
                        double* g_coef, void* stream);
/* sem_bspline_project that allocates nothing: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  Its partial sums go
 * to an 8 MiB device array of the library (one per GPU, held until the process
 * ends), models processed in chunks that fit: SEM_E_NOMEM if one model needs
 * more (NZ * nbx > 2^20).  The basis tables of (nbz, nbx) must exist already --
 * made by an earlier sem_bspline_map or sem_bspline_project call with these
 * node counts -- else SEM_E_INVALID.  Calls that overlap on one GPU share the
 * array and race; calls on one stream do not.  Same results as
 * sem_bspline_project. */
int sem_bspline_project_noalloc(sem_plan* plan, int nmodel, int nbz, int nbx, const double* g_c,
                                double* g_coef, void* stream);

/* Closed-curve anomaly velocity, the reference's parameterisation: signed
 * distance to a closed cubic B-spline through 6 control points (nearest of
 * `samples` curve samples + `newton_steps` Newton steps, sign from a
 * point-in-polygon test), v = v_out + (v_in - v_out) sigmoid(-sd/tau).
 * Replaces build_velocity_on_sem_nodes / signed_distance_to_spline_newton
 * (Bspline_Coupling/SEM/sem_waveform/velocity.py:37-202).
 *   ctrl  device double [nmodel][6][2]  absolute control points (x, z)
 *   c     device double [nmodel][N] or NULL;  c2 device real [nmodel][N]
 *   aux   device double [nmodel][4][N]  (sd, u*, n_out.x, n_out.z) for the projection
 * samples <= 2048. */
int sem_curve_velocity(sem_plan* plan, int nmodel, const double* ctrl, double v_inside, double v_outside,
                       double tau, int samples, int newton_steps, double* c, void* c2, double* aux, void* stream);
/* Analytic projection of a nodal gradient dL/dv onto the 12 control
 * coordinates, interleaved [x0, z0, x1, z1, ...]; replaces
 * _analytical_gradient_projection_improved (core.py:902-970, 1357-1363).
 *   dL_dv device double [nmodel][N];  grad_ctrl device double [nmodel][12] */
int sem_curve_project(sem_plan* plan, int nmodel, const double* dL_dv, const double* aux, double v_inside,
                      double v_outside, double tau, double* grad_ctrl, void* stream);

/* Generalised closed-curve anomaly.  kind = SEM_CURVE_CLAMPED6 is the curve
 * above; SEM_CURVE_PERIODIC is the closed periodic cubic B-spline through K
 * control points of CO2_plume/one_stage/single_cubic/sem_waveform/velocity.py:9-152
 * (controls extended by their first 3, knots 0, 1, ..., K+6, parameter interval
 * [3, K+3), samples at 3 + kK/S, u* stored normalised to [0, 1)).  The anomaly
 * is blended over a per-node background (velocity.py:277-332):
 *   SEM_BLEND_SMOOTH : v = v_bg + (v_in - v_bg) sigmoid(-sd/tau)
 *   SEM_BLEND_REPLACE: v = v_in where sd < 0, else v_bg
 * with v_bg = background[node] (device double [N]) or v_outside if NULL.
 *   ctrl device double [nmodel][nctrl][2];  c, c2, aux as sem_curve_velocity. */
#define SEM_CURVE_CLAMPED6 0
#define SEM_CURVE_PERIODIC 1
#define SEM_BLEND_SMOOTH 0
#define SEM_BLEND_REPLACE 1
#define SEM_CURVE_MAX_CTRL 32
typedef struct {
    int kind;                  /* SEM_CURVE_*                                          */
    int nctrl;                 /* 6 for CLAMPED6; 4 .. SEM_CURVE_MAX_CTRL for PERIODIC  */
    int samples;               /* initial-guess curve samples, 2 .. 2048               */
    int newton_steps;
    double v_inside, v_outside, tau;
    const double* background;  /* device double [N] or NULL                            */
    int blend;                 /* SEM_BLEND_*                                          */
} sem_curve_desc;
int sem_curve_velocity_ex(sem_plan* plan, const sem_curve_desc* cd, int nmodel, const double* ctrl, double* c,
                          void* c2, double* aux, void* stream);
/* Projection of dL/dv onto the 2 nctrl control coordinates, interleaved
 * [x0, z0, x1, z1, ...]: the clamped rule above for CLAMPED6 (threshold
 * 1e-8 (v_bg - v_in)); for PERIODIC the rule of
 * CO2_plume/one_stage/single_cubic/sem_waveform/core.py:1075-1185 (periodic
 * basis with the 3 appended controls folded back, nodes kept where
 * |dv/dsd| > 1e-8 max(1, |v_in - v_bg|)).  grad_ctrl device double [nmodel][2 nctrl]. */
int sem_curve_project_ex(sem_plan* plan, const sem_curve_desc* cd, int nmodel, const double* dL_dv,
                         const double* aux, double* grad_ctrl, void* stream);

/* ---- continuous batching (slots) -------------------------------------------
 * A slot is one solve that advances through its own phases at its own time
 * step.  A slot table lives in device memory; sem_slots_step() advances every
 * slot whose phase matches by one step of its current segment, so requests of
 * one mesh size with different step counts, sources and receivers share each
 * launch and may join or leave between segments.  All pointers are device
 * pointers into the slot's own buffers (nodal arrays in the plan's padded
 * layout, sem_plan_layout()).  No reference counterpart: the reference runs one
 * solve per call (core.py:1067) and its drivers loop over samples and sources
 * (CO2_plume/one_stage/single_cubic/RQS_NF.py:1035-1072, 1488-1492). */
#define SEM_SLOT_IDLE 0
#define SEM_SLOT_FORWARD 1    /* u0/u1 ping-pong (by step parity, or by step pair:
                                 pair_parity), w increment                           */
#define SEM_SLOT_RECOMPUTE 2  /* u0 = checkpoint u of the segment, w = its w, seg    */
#define SEM_SLOT_ADJOINT 3    /* lambda in the ring, w = omega                        */
#define SEM_SLOT_FORWARD_HIST 4 /* forward into the full history: u^{n0+m} in seg slot m
                                   (seg = history + n0 * seg_stride), w increment; the
                                   adjoint then reads u from the history (no recompute) */

typedef struct {
    int phase;             /* SEM_SLOT_*                                              */
    int n0, len;           /* first time step of the current segment, its length     */
    int src_node;          /* injection node                                          */
    int rec_set;           /* receiver set id                                         */
    int record;            /* forward: sample seismograms during this segment         */
    void* u0;
    void* u1;
    void* w;
    const void* c2;        /* padded squared velocity of this solve                   */
    const void* wavelet;   /* [nt]                                                    */
    void* seis;            /* [nt][rec_pitch]                                         */
    void* ck_u;            /* forward: (u, w) at the end of the segment go here, or NULL */
    void* ck_w;
    void* seg;             /* recompute: slots 1..len of u;  forward_hist: slots 0..len;
                              adjoint: u^{it+1} at slot it-n0+1                        */
    long seg_stride;       /* elements between segment slots                          */
    const void* resid;     /* adjoint source [nt][rec_pitch]                          */
    double* G;             /* gradient accumulator, sem_plan_gradient_elems() doubles */
    void* ring;            /* adjoint: lambda_t in slot (t+1) % ring_slots, stride seg_stride */
    int ring_slots;        /* >= chunk_len + 1 of sem_slots_correlate                 */
    int pair_parity;       /* forward: 0 = u^n in u1 when n is odd, else in u0 (one step
                              per launch, sem_slots_step); 1 = u^n in u1 when (n + 1) / 2
                              is odd, else in u0 (two steps per launch, sem_slots_step2) */
    void* w2;              /* sem_slots_step2: second increment buffer (the increment
                              alternates w -> w2 -> w ... launch by launch)           */
} sem_slot;

/* Padded nodal layout: row pitch and per-array element count. */
int sem_plan_layout(const sem_plan* plan, int* pitch, long* array_elems);
/* Doubles in one solve's gradient accumulator (sem_slot.G, sem_gradient_finalize):
 * the padded nodal array followed by 9 per-element planes, so that every element
 * keeps its own running sums for its 25 nodes and assembly happens once per solve.
 * Replaces the per-step nodal assembly of accumulate_sensitivity_dJdc_torch
 * (core.py:49-85); zero-fill before the first adjoint chunk. */
int sem_plan_gradient_elems(const sem_plan* plan, long* elems);
/* One step of every slot whose phase is in phase_mask (bit 1 << SEM_SLOT_*):
 * forward/recompute slots run the forward step, adjoint slots the adjoint step
 * (a mask mixing both launches both kernels).  m = step index within the
 * current segments; slots with m >= len skip.  rec_pitch: seismogram row length.
 * FORWARD slots with pair_parity set are not advanced (sem_slots_step2 is
 * theirs). */
int sem_slots_step(sem_plan* plan, const sem_slot* slots_device, int nslots, int m, unsigned phase_mask,
                   int rec_pitch, void* stream);
/* Steps m and m + 1 (m even) of every FORWARD, FORWARD_HIST, RECOMPUTE and
 * ADJOINT slot in phase_mask, in one launch per kernel (two-step kernel, the
 * same bits as two sem_slots_step calls); a slot with m + 1 == len takes step
 * m alone.  The increment is read from w and written to w2 on launches with
 * (m / 2) even, the other way round on odd ones, so it ends a segment of len
 * steps in w2 when ceil(len / 2) is odd.  A FORWARD slot reads u^n from the
 * buffer pair_parity names and writes u^{n+2} (u^{n+1} when it takes a step
 * alone) to the other one; its checkpoint (ck_u, ck_w) receives the state
 * after the segment's last pair.  No reference counterpart (continuous
 * batching).
 * Conditions on the slot table (a slot that breaks the first three is not
 * advanced at all):
 *   - w2 is non-null and differs from w;
 *   - FORWARD slots have pair_parity set and an even n0, and an even len
 *     when ck_u is set;
 *   - ADJOINT slots have ring_slots >= 3 (a pair writes lambda_{t-1} and
 *     lambda_{t-2} while lambda_t is read);
 *   - the chunk_len of the sem_slots_correlate calls between these launches
 *     is even, so that no step pair straddles a correlation chunk (the pair
 *     would overwrite a ring slot the pending correlation still reads). */
int sem_slots_step2(sem_plan* plan, const sem_slot* slots_device, int nslots, int m, unsigned phase_mask,
                    int rec_pitch, void* stream);
/* Gradient correlation of chunk `chunk` (segment steps chunk*chunk_len ..
 * min((chunk+1)*chunk_len, len) - 1) of every ADJOINT slot, added to its G in
 * FP64; enqueue after those adjoint steps and before the ring slots are reused
 * (ring_slots >= chunk_len + 1).  Replaces the per-step correlation of
 * accumulate_sensitivity_dJdc_torch (core.py:49-85, 1265-1270). */
int sem_slots_correlate(sem_plan* plan, const sem_slot* slots_device, int nslots, int chunk, int chunk_len,
                        void* stream);
/* Receiver samples of u (final post-update sample) for slots in phase_mask:
 * forward slots sample u[(n0+len)&1] (forward_hist slots seg slot len) into
 * seis row n0+len-1. */
int sem_slots_record_final(sem_plan* plan, const sem_slot* slots_device, int nslots, unsigned phase_mask,
                           int rec_pitch, void* stream);

/* Receiver tables for captured launch sequences: reserve device capacity for
 * `max_entries` tile entries (each receiver appears in 1-4 tiles) and
 * `max_sets` sets so that later sem_receiver_set_add calls never move the
 * table; sem_receivers_sync uploads sets added since the last upload (only the
 * appended part is written, so launches in flight are unaffected). */
int sem_receivers_reserve(sem_plan* plan, long max_entries, int max_sets);
int sem_receivers_sync(sem_plan* plan);

/* Per-solve helpers for slot owners (device pointers):
 * copy [nmodel][N] squared velocities into the padded layout;  */
int sem_pad_nodal(sem_plan* plan, const void* src, void* dst_padded, int nmodel, void* stream);
/* res = seis - obs on the first nrec columns of nt rows (obs double, the difference
 * formed in double and stored in the state precision), misfit = 0.5 sum res^2 (double) */
int sem_residual(sem_plan* plan, const void* seis, const double* obs, void* res, double* misfit, int nrec, int nt,
                 int rec_pitch, void* stream);
/* grad_c2 = scale * G, grad_c = 2 sqrt(c2) grad_c2 for one solve, G assembled per
 * node from the accumulator (sem_plan_gradient_elems doubles), c2 padded; outputs [N] */
int sem_gradient_finalize(sem_plan* plan, const double* G, const void* c2_padded, double scale, double* grad_c,
                          double* grad_c2, void* stream);

/* ---- device-resident request loop -------------------------------------------
 * A lane of continuous batching whose requests, slot states and per-round slot
 * tables live in device memory.  The host pushes admitted requests into a
 * device queue (sem_loop_push, stream-ordered) and enqueues rounds
 * (sem_loop_round); each round, on the device: picks queued requests into free
 * slots in arrival order, writes the slot table, advances every slot by one
 * segment of `segment` steps (the slot kernels above, so a solve's arithmetic is
 * the batch path's), records the final samples and forms the residual and misfit
 * of the slots whose forward pass ended, assembles the nodal gradient of the
 * solves that completed, moves every slot to its next phase and frees the
 * completed ones, appending their ids to a completion list in host-mapped
 * memory (sem_loop_poll).  A round's launches and arguments never change, so a
 * captured graph of one round serves every round.  No reference counterpart:
 * the reference runs one solve per call (core.py:1067) and its drivers loop over
 * samples and sources (CO2_plume/one_stage/single_cubic/RQS_NF.py:1035-1072,
 * 1488-1492). */
typedef struct {
    long long id;          /* caller's request id, echoed in the completion list       */
    int nt;                /* time steps (<= the plan's nt)                            */
    int full_history;      /* 1: hist holds u^0..u^nt (no recompute); 0: checkpoints   */
    int src_node;          /* injection node (sem_source_node)                         */
    int rec_set;           /* receiver set id                                          */
    int nrec;              /* receivers of the set (columns of seis / resid / obs used) */
    int reserved;
    double inv_sigma2;     /* gradient scale 1 / (noise_std^2 + 1e-24)                 */
    const void* c2;        /* padded squared velocity                                  */
    const void* wavelet;   /* [nt]                                                     */
    void* seis;            /* [nt][rec_pitch] synthetic seismograms (state precision)  */
    void* resid;           /* [nt][rec_pitch] residual (state precision)               */
    const double* obs;     /* [nt][rec_pitch] observations                             */
    double* misfit;        /* [1] 0.5 ||resid||^2                                      */
    double* G;             /* gradient accumulator (sem_plan_gradient_elems), zeroed   */
    double* grad_c;        /* [N] d loglik / d c, written when the solve completes      */
    void* w;               /* forward increment, zeroed                                */
    void* w2;              /* second forward increment buffer (two-step loops), zeroed */
    void* om;              /* adjoint increment, zeroed                                */
    void* om2;             /* second adjoint increment buffer (two-step loops), zeroed */
    void* ring;            /* chunk + 1 adjoint states, zeroed                         */
    void* hist;            /* full history: nt + 1 states, u^0 zeroed                  */
    void* u0;              /* checkpoints: state pair, zeroed                          */
    void* u1;
    void* cku;             /* checkpoints: ceil(nt / segment) states of u and of w;    */
    void* ckw;             /*   entry 0 zeroed                                         */
    void* seg;             /* checkpoints: segment + 1 states                          */
} sem_request;

typedef struct sem_loop sem_loop;
/* nslots <= 256 slots advancing `segment` steps per round (two_step: every
 * slot two steps per launch; needs an even segment and correlation chunk);
 * queue_capacity: requests pushed and not yet
 * completed, at most.  Allocates the loop's device and host-mapped state. */
int sem_loop_create(sem_plan* plan, int nslots, int segment, int two_step, int rec_pitch, int queue_capacity,
                    sem_loop** out);
int sem_loop_destroy(sem_loop* loop);
/* Adjoint steps per gradient-correlation chunk of the loop (ring states per
 * request: chunk + 1). */
int sem_loop_chunk(const sem_loop* loop, int* chunk);
/* Enqueue one request on `stream` (after the work that fills its buffers). */
int sem_loop_push(sem_loop* loop, const sem_request* req, void* stream);
/* Enqueue one round on `stream`; allocates nothing, capturable. */
int sem_loop_round(sem_loop* loop, void* stream);
/* Completion records since index `since` (the list is cumulative): up to `max`
 * ids and the rounds that completed them; *total = completions so far.  Reads
 * host-mapped memory written by the device (no device call).  Records older than
 * the list's 4096 most recent are gone (SEM_E_INVALID). */
int sem_loop_poll(sem_loop* loop, long long since, long long* ids, long long* rounds, int max, long long* total);
/* Counters: requests pushed, sem_loop_round calls (a round captured into a graph
 * counts once however often it is replayed), rounds completed on the device,
 * slots occupied after the last completed round. */
int sem_loop_counts(const sem_loop* loop, long long* pushed, long long* rounds_enqueued, long long* rounds_done,
                    int* active);

/* A non-blocking HIP stream of default priority on the calling thread's
 * current device, and its release (waits for the work on it first).  No
 * reference counterpart: a serving lane takes one for its lifetime, so that
 * each lane's work sits on a hardware queue of its own; lanes on pooled
 * streams could share one and then run one after the other. */
int sem_stream_create(void** stream);
int sem_stream_destroy(void* stream);

/* Thread-local message of the last failing call on this thread. */
const char* sem_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MUN_SEM_CAPI_H */
