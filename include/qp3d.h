/*
 * qp3d — MI355X-native 3D pure-qP VTI forward modeling (C ABI).
 *
 * Grid layout everywhere: padded fields [nz][nx][ny], y unit-stride, linear index
 * z*nx*ny + x*ny + y, with nx = hx + 2*pml, ny = hy + 2*pml, nz = Nz0 + 2*pml
 * (array_new.c:2, GPU_kernel.cu:551, Main.c:125). Every pointer argument below is host
 * memory owned by the caller unless the comment says otherwise.
 */
#ifndef QP3D_H
#define QP3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the legacy void entry point records them in qp3d_last_error) ---- */
#define QP3D_OK 0
#define QP3D_ERR_ARG (-1)       /* invalid dimensions, pml, window or pointer */
#define QP3D_ERR_IO (-2)        /* missing or short input file, unwritable output */
#define QP3D_ERR_ALLOC (-3)     /* host or device allocation failed */
#define QP3D_ERR_HIP (-4)       /* a HIP runtime call failed */
#define QP3D_ERR_NONFINITE (-5) /* the wavefield became inf/NaN (CFL violation) */
#define QP3D_ERR_STATE (-6)     /* call order violated (e.g. run before upload) */
#define QP3D_ERR_TIMEOUT (-7)   /* a bounded device-side wait ran out (persistent batch launch) */

/* Message of the last failure on the calling thread ("" if none). */
const char *qp3d_last_error(void);
/* Internal helper shared by the host sources of the library (printf-style). */
void qp3d_set_error(const char *fmt, ...);

/* ---- drop-in solver entry point -------------------------------------------------------
 * Replaces void modeling3d(...) of 3dPureqPModeling_single_shot.h:4-6 (GPU implementation
 * 3dPureqPModeling_single_shot.cu:26-447, CPU twin modeling3D.c:17-153) with the same
 * argument list and meaning:
 *   vp, epsilon, delta, source  padded model and spatial source [nz][nx][ny];
 *   wavelet[nt]                  source time function;
 *   record                       out: [nt][nx-2pml][ny-2pml], record[k] = u(sz) after step k;
 *   dlr..ddfb[pml]               PML profiles (qp3d_pmlcoff); c, c2[36] FD tables (qp3d_fdcoff*);
 *   snapflag, ntsnap, snap_file  when snapflag == 1 and k % ntsnap == 0, writes
 *                                <snap_file>k.dat (interior cube) and <snap_file>_{xz,yz,xy}k.dat.
 * sx, sy are unused (the source position is carried by `source`), as in the reference.
 * Runs on HIP device 0, synchronously; failures are reported by qp3d_last_error(). A wavefield
 * that becomes non-finite (CFL violation) still leaves its record (and snapshot files) written,
 * holding the inf/NaN values, as the reference's would.
 * The library also exports this entry point with C++ linkage (include/qp3d_cxx.h), the symbol a
 * driver compiled with g++ against the reference's header binds. */
void modeling3d(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt, int pml,
                int snapflag, int sx, int sy, int sz, float *vp, float *epsilon, float *delta, float *source,
                float *wavelet, float *record, float *dlr, float *ddlr, float *dtb, float *ddtb, float *dfb,
                float *ddfb, float *c, float *c2, const char *snap_file);

/* Same as modeling3d() but returns a QP3D_* status. No reference counterpart: the
 * reference's entry point returns void and checks nothing. */
int qp3d_modeling3d(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt, int pml,
                    int snapflag, int sx, int sy, int sz, const float *vp, const float *epsilon,
                    const float *delta, const float *source, const float *wavelet, float *record,
                    const float *dlr, const float *ddlr, const float *dtb, const float *ddtb, const float *dfb,
                    const float *ddfb, const float *c, const float *c2, const char *snap_file);

/* Allocation-free variant of qp3d_modeling3d. No counterpart in the reference; it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments and meaning as qp3d_modeling3d, except:
 *   - every array (model, source, wavelet, tables and the record) is device memory of the
 *     device current on the calling thread;
 *   - snapflag must be 0 (snapshots would need memory this entry point does not allocate);
 *   - stream (a hipStream_t, NULL = the null stream) orders the run; the call returns after the
 *     nt steps have finished (QP3D_ERR_NONFINITE as qp3d_modeling3d).
 * It calls no hipMalloc/hipFree: the wavefields, PML state, converted model, PML table,
 * per-step maxima, flag, source box and work lists live in one fixed device array of
 * QP3D_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the process ends
 * (about 48 bytes per grid cell with rows padded to 32 cells, plus 40 bytes per y-face PML
 * cell: a 141^3 grid fits); larger grids are rejected with QP3D_ERR_ARG. Every call on a device
 * shares that device's array: calls that overlap on one GPU (other threads, other streams) race
 * and must be serialised by the caller. */
#define QP3D_NOALLOC_SCRATCH_BYTES (192ull << 20)
int qp3d_modeling3d_noalloc(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt,
                            int pml, int snapflag, int sx, int sy, int sz, const float *vp, const float *epsilon,
                            const float *delta, const float *source, const float *wavelet, float *record,
                            const float *dlr, const float *ddlr, const float *dtb, const float *ddtb,
                            const float *dfb, const float *ddfb, const float *c, const float *c2,
                            const char *snap_file, void *stream);

/* ---- host setup math (bit-identical to the reference's tables) ---- */
void qp3d_fdcoff1(float *c);  /* CPU_function.c:111-128 fdcoff1, c[36] */
void qp3d_fdcoff2(float *c);  /* CPU_function.c:131-148 fdcoff2, c[36] */
/* CPU_function.c:150-169 pmlcoff */
void qp3d_pmlcoff(int pml, float vpmax, float dx, float dy, float dz, float *dlr, float *ddlr, float *dtb,
                  float *ddtb, float *dfb, float *ddfb);
/* CPU_function.c:37-64 Source; either output may be NULL to skip it. */
void qp3d_source(int sx, int sy, int sz, float fm, float amp, float alpha, float dt, float dx, float dy, float dz,
                 float t0, int nt, int nx, int ny, int nz, float *source, float *wavelet);
/* CPU_function.c:67-108 extend_model */
void qp3d_extend_model(const float *in, float *out, int nx, int ny, int nz, int pml);
/* CPU_function.c:10-22 absMaxval over n cells */
float qp3d_absmax(const float *a, long n);
/* CPU_function.c:234-261 readmodel3d; returns a status (the reference checks nothing). */
int qp3d_readmodel3d(const char *filename, float *a, int ny, int nx, int nz, int fny, int fnx, int fnz, int vny,
                     int vnx, int vnz);
/* Synthetic layered VTI model in model-file layout [ny][nx][nz]; no reference counterpart
 * (the reference's model files are not distributed). */
int qp3d_genmodel(int nx, int ny, int nz, uint64_t seed, float sigma_cells, float amp_frac, float *vp, float *eps,
                  float *del);

/* ---- plan API: device-resident runs -------------------------------------------------
 * No reference counterpart; it splits modeling3d() into allocate / upload / run / fetch so
 * that a caller can keep the model in HBM across shots, time the loop, or read state. */
#define QP3D_STORAGE_FP32 0 /* wavefields and PML state in float */
#define QP3D_STORAGE_BF16 1 /* stored bfloat16, arithmetic float */
#define QP3D_STORAGE_FP16 2 /* stored half with a power-of-two scale, arithmetic float */

typedef struct qp3d_config {
  int nx, ny, nz, pml, nt;
  float dx, dy, dz, dt;
  int sz;        /* receiver plane, padded z index */
  int storage;   /* QP3D_STORAGE_* */
  int use_graph; /* 1: steps run as replays of one captured hipGraph of 10 steps (a group that
                    starts on an odd step or holds a snapshot step is launched directly, and so
                    is every step while instrumentation is on); with two streams the graph
                    forks each step's PML updates onto the second stream and joins them.
                    0: direct launches */
  int snapflag;  /* 1: extract the interior cube when k % ntsnap == 0 */
  int ntsnap;
  int zchunk;    /* z planes per workgroup in the stencil kernels, 0 = automatic */
  int device;    /* HIP device ordinal */
  int tile_update; /* tile shape of the interior update kernel, 1..QP3D_NTILES, 0 = default */
  int tile_grad;   /* tile shape of the gradient-max kernel, 1..QP3D_NTILES, 0 = default */
  int tile_pml;    /* tile shape of the PML update on the x and z faces, 1..QP3D_NTILES, 0 = default */
  int streams;     /* 2: run the interior and PML updates concurrently on two streams; 1: one stream;
                      0 = default (2; 1 with gmax_mode 2). The two updates write disjoint cells;
                      early in a shot both are bound by the march of their few active items, and
                      concurrent they overlap (576^3 shot: 1.046 against 1.113 ms per step).
                      The FWI gradient's steps and a batch's steps run on one stream. */
  int tile_pml_y;  /* tile shape of the PML update on the y faces (a launch of its own),
                      1..QP3D_NTILES, 0 = default */
  int snap_s;      /* 1 (with snapflag == 1): at every snapshot step k also keep the scalar-operator
                      field S of that step over the whole padded grid (read it with
                      qp3d_plan_snapshot_s); 0: not kept */
  int gmax_mode;   /* how a step obtains G, the gradient maximum that sets the scalar operator's
                      threshold 0.01*G: 1: a gradient pass over the grid before the update (the
                      reference's order); 2: the update kernels use the threshold predicted from
                      the previous steps' G, reduce this step's G themselves, and a fix-up
                      kernel recomputes the cells whose decision |grad u| > 0.01*G the prediction
                      got wrong (a marking pass over the grid finds them when the prediction
                      missed by more than 1/128); the same results up to the rounding of those
                      cells; implies streams = 1. 0 = default (1) */
} qp3d_config;

/* Tile shapes (wavefronts along x x rows per lane; a tile is 64 cells along y):
 * 1: 8x2  2: 16x1  3: 4x2  4: 8x1  5: 12x1  6: 16x2  7: 4x4 */
#define QP3D_NTILES 7

typedef struct qp3d_plan qp3d_plan;

/* Arrays handed to the plan, FWI and batch calls below may be host memory or device memory of
 * the plan's device (the kind is read from each pointer); outputs are written where they point.
 * The modeling3d() entry point above takes host memory, as the reference's does. */
int qp3d_plan_create(const qp3d_config *cfg, qp3d_plan **out);
int qp3d_plan_destroy(qp3d_plan *p);
/* Uploads model, source and tables (same meaning as the modeling3d() arguments). */
int qp3d_plan_upload(qp3d_plan *p, const float *vp, const float *eps, const float *del, const float *source,
                     const float *wavelet, const float *dlr, const float *ddlr, const float *dtb, const float *ddtb,
                     const float *dfb, const float *ddfb, const float *c, const float *c2);
/* Zeroes wavefields, PML state, per-step maxima and the device record. */
int qp3d_plan_reset(qp3d_plan *p);
/* Enqueues time steps [k0, k1) on the plan's stream (asynchronous). */
int qp3d_plan_run(qp3d_plan *p, int k0, int k1);
/* The same, ordered with the caller's HIP stream (a hipStream_t; NULL = the null stream): the
 * steps start after the work already queued on it, and work queued on it afterwards waits for
 * them. No host synchronisation; qp3d_plan_sync still reports a non-finite wavefield. */
int qp3d_plan_run_on(qp3d_plan *p, int k0, int k1, void *stream);
/* Waits for the plan's stream; returns QP3D_ERR_NONFINITE if the field blew up. */
int qp3d_plan_sync(qp3d_plan *p);
/* Copies the device record [nt][hx][hy] to `record` (host or device memory). */
int qp3d_plan_record(const qp3d_plan *p, float *record);
/* State after the last step run, padded [nz][nx][ny] floats. which: 0 = u after the last
 * step (u2), 1 = the step before (u1); 2/3 = w1 newest/older, 4/5 = w3 newest/older,
 * 6/7 = p newest/older, 8 = w21 (PML state in full-grid layout, zero outside the PML). */
int qp3d_plan_field(const qp3d_plan *p, int which, float *out);
/* Number of snapshots held and the copy of snapshot i ([hz][hx][hy]) with its step index. */
int qp3d_plan_snapshot_count(const qp3d_plan *p);
int qp3d_plan_snapshot(const qp3d_plan *p, int i, int *step, float *out);
/* The scalar-operator field S of snapshot i's step over the padded grid ([nz][nx][ny], 1 outside
 * [5,n-5)^3), as that step's update used it; needs cfg.snap_s. Replaces the S-field snapshot
 * the reference computes every step and has commented out of its writer
 * (3dPureqPModeling_single_shot.cu:381-409). */
int qp3d_plan_snapshot_s(const qp3d_plan *p, int i, int *step, float *out);
/* Per-step gradient maxima G[k] (float) of steps [0, nt). */
int qp3d_plan_step_maxima(const qp3d_plan *p, float *out);
/* Work items the plan would launch for a configuration, built on the host without a device
 * (for tests and tools): which = 0 gradient-max pass, 1 interior update, 2 PML update on the
 * x and z faces, 3 PML update on the y faces; capacity = resident workgroups the z-chunking
 * fills. Writes min(count, cap) items of 4 ints (x0 | x1 << 16, y0 | y1 << 16 with bit 31 set
 * for a folded y-face tile, z0, z1) in logical order and returns the count, or a QP3D_* error. */
int qp3d_work_items(const qp3d_config *cfg, int which, int capacity, int *out, int cap);
/* Interior extents hx, hy, hz of the plan's grid (any output may be NULL). */
int qp3d_plan_dims(const qp3d_plan *p, int *hx, int *hy, int *hz);
/* gmax_mode 2: the predicted G of every step, out[nt] (0 for step 0), from which each step's
 * threshold and marked band were taken (the prediction of step k is fitted to G_0..G_{k-1});
 * QP3D_ERR_STATE for a plan created with another mode. */
int qp3d_plan_step_predictions(const qp3d_plan *p, float *out);
/* How the plan's steps obtain G (qp3d_config.gmax_mode after defaults: 1 or 2; 1 while the
 * plan belongs to a batch), or a QP3D_* error. */
int qp3d_plan_gmax_mode(const qp3d_plan *p);
/* Amplitude bound of the gradient pass (gmax_mode 1; no reference counterpart): the update
 * kernels record max |u| per block of 8 x 8 x 32 cells (z, x, y), a small pass evaluates exactly
 * the tile that held the previous step's largest gradient component, and the pass skips every
 * tile whose bound 2 sum|c1|/d max|u| is below the G found so far, without reading it. From the
 * same maxima, after the gradient pass, the updates learn the blocks where S = 1 (they skip the
 * eps - del load there) and the blocks still at rest (u = 0 within 5 cells at every step since
 * the reset, and no source: they skip them). G, and so every result, is the same as with the
 * bound off. On by default; the setter turns it off (0) or on (1).
 * qp3d_plan_gradient_tiles: tiles evaluated (out[0]) and skipped (out[1]) since the last reset;
 * qp3d_plan_update_cells: out[9], per update kind (interior tiles, PML tiles over the full-grid
 * state, folded y-face tiles) the cells of its items, of them left out at rest, and updated
 * without eps - del, since the last reset. Both counted on instrumented steps
 * (qp3d_plan_set_instrument). */
int qp3d_plan_set_amp_bound(qp3d_plan *p, int on);
int qp3d_plan_gradient_tiles(const qp3d_plan *p, unsigned long long *out);
int qp3d_plan_update_cells(const qp3d_plan *p, unsigned long long *out);
/* 1: bracket every stencil launch with HIP events on the plan's stream; 2: only the interior
 * update (two events per step instead of four); 0: none. Resets the accumulated times. */
int qp3d_plan_set_instrument(qp3d_plan *p, int on);
/* Mean device time (ms) per time step over the instrumented steps: ms[0] gradient-max launch
 * (gmax_mode 2: the marking pass and the fix-up),
 * ms[1] interior update, ms[2] PML update, ms[3] the whole update stage (end of the gradient
 * maximum to the end of both updates; they overlap when streams == 2); ms[4] = steps counted.
 * With instrumentation mode 2 only ms[1] and ms[4] are measured (the others read 0). */
int qp3d_plan_kernel_times(qp3d_plan *p, double *ms);
/* Interior-update time (ms) of each step, ms[nt], as last measured on an instrumented step of
 * this plan (0 for steps never measured); harvested by qp3d_plan_sync. */
int qp3d_plan_step_times(const qp3d_plan *p, float *ms);
/* Algorithmic HBM bytes of one time step (interior cells 28 B, PML cells 68 B in FP32):
 * bytes[0] gradient-max launch (4 B per updated cell), bytes[1] interior update launch (24 B per
 * cell), bytes[2] PML update launch (64 B per PML cell, 24 B per interior cell it carries);
 * bytes[3], bytes[4] = cells of the interior and the PML update launches. */
int qp3d_plan_step_bytes(const qp3d_plan *p, double *bytes);

/* ---- adjoint-state FWI gradient ---------------------------------------------------------
 * No reference counterpart (the reference models forward only). Gradient of
 * J = 1/2 sum_k |record_k - observed_k|^2 with respect to the padded-grid vp, epsilon and delta
 * that were uploaded to the plan, for the discrete forward operator as implemented (PML
 * profiles held fixed, the 0.01*G threshold of the scalar operator treated as a constant).
 * The forward run stores the state every `interval` steps (checkpoint storage FP32 or BF16);
 * the gradient recomputes each segment from its checkpoint and back-propagates the residual
 * with the adjoint kernels, summing in FP32 within a segment and in FP64 over time.
 * interval >= nt stores every step of the run (full-wavefield storage). The plan must use
 * FP32 storage; the gradient's forward steps run on the plan's own stream. */
typedef struct qp3d_fwi qp3d_fwi;
int qp3d_fwi_create(qp3d_plan *p, int interval, int ckpt_storage, qp3d_fwi **out);
int qp3d_fwi_destroy(qp3d_fwi *f);
/* Forward run from rest with checkpoints; record, maxima and state as after qp3d_plan_run. */
int qp3d_fwi_forward(qp3d_fwi *f);
/* resid = record - observed, [nt][hx][hy]; outputs [nz][nx][ny] each (host or device memory). */
int qp3d_fwi_gradient(qp3d_fwi *f, const float *resid, double *g_vp, double *g_eps, double *g_del);
/* ms[0] checkpointed forward run, ms[1] recomputation, ms[2] adjoint sweep (device time). */
int qp3d_fwi_times(const qp3d_fwi *f, double *ms);
/* on != 0 (the default): the adjoint sweep keeps a per-block map of where its state has been
 * nonzero (from the receiver plane on) and leaves out the planes of each work item outside it,
 * where its launches would write zeros over zeros. The gradient is the same either way (up to
 * the sign of zeros). */
int qp3d_fwi_set_rest_map(qp3d_fwi *f, int on);
/* The forward run keeps the u_k of the last n segments resident in HBM (1 by default), and the
 * gradient back-propagates those without recomputing them (n = number of segments: no
 * recomputation at all, the checkpoints still taken); n < 0: as many as fit in the free device
 * memory less 2 GiB. Reallocates the segment buffer (interval * nx * ny * nz * 4 B per segment,
 * rows padded); run qp3d_fwi_forward again afterwards. *kept (may be NULL) = the number kept. On
 * an allocation failure one segment is kept where it fits (QP3D_ERR_ALLOC either way). */
int qp3d_fwi_set_resident_segments(qp3d_fwi *f, int n, int *kept);

/* ---- batched shots ------------------------------------------------------------------------
 * No reference counterpart (the reference runs its shots one after another, Main.c:179-247).
 * A batch advances several plans (one per shot: own grid window, model, source, PML profiles
 * and record; same nt, storage and device) in lockstep with one launch per stage and time step
 * for all shots. Build it after the plans' uploads; a plan uploaded again needs a new batch.
 * use_graph: steps run as replays of a captured 10-step graph. The plans' own run/sync are not
 * used while the batch drives them; records, fields and maxima are read through the plans. */
typedef struct qp3d_batch qp3d_batch;
int qp3d_batch_create(qp3d_plan *const *plans, int nshots, int use_graph, qp3d_batch **out);
int qp3d_batch_destroy(qp3d_batch *b);
int qp3d_batch_reset(qp3d_batch *b);
int qp3d_batch_run(qp3d_batch *b, int k0, int k1);
int qp3d_batch_sync(qp3d_batch *b);
/* Picks how a step goes out, timing `steps` steps of each candidate: one launch per stage for
 * all shots with each shot's work items cut for the whole machine (as its own plan cuts them)
 * or for its share of the machine, each with every stage's tile shape timed and the fastest
 * kept; each shot's own launches one after another; or one persistent launch per run, with
 * each tile set and either cut; then resets;
 * tiles_out[4] / ms_out[4] (may be NULL): the shared launches' tiles and stage times: gradient
 * max, interior, PML x/z faces, PML y faces. */
int qp3d_batch_autotune(qp3d_batch *b, int steps, int *tiles_out, double *ms_out);
/* 1 when the shared launches cut each shot's work items for its share of the machine, 0 for
 * the whole machine (the default; chosen by qp3d_batch_autotune). */
int qp3d_batch_share(const qp3d_batch *b);
/* 1 when the batch runs each shot's own launches (chosen by qp3d_batch_autotune), 0 when it
 * runs one launch per stage for all shots (the default); the setter forces either. */
int qp3d_batch_per_shot(const qp3d_batch *b);
int qp3d_batch_set_per_shot(qp3d_batch *b, int on);
/* Persistent mode (the shots of the reference's serial loop, Main.c:179-247, in one launch): each
 * qp3d_batch_run is a single launch over all its steps and shots. Resident workgroups take
 * work items (gradient-pass and update tiles of one shot and step) from one queue; two item
 * counters per shot order a shot's phases, so shots advance independently. Every
 * device-side wait is bounded (qp3d_batch_set_wait_limit, default 5 s): past the bound the
 * launch leaves and qp3d_batch_sync returns QP3D_ERR_TIMEOUT. qp3d_batch_persistent: the tile
 * set in use (0 .. 2), -1 when off; the setter turns it on with a tile set and cut (share 1:
 * each shot's items cut for its share of the machine) or off (set -1); qp3d_batch_autotune
 * chooses it when it is the fastest form. Bit-identical to the other forms. */
int qp3d_batch_persistent(const qp3d_batch *b);
int qp3d_batch_set_persistent(qp3d_batch *b, int set, int share);
int qp3d_batch_set_wait_limit(qp3d_batch *b, double us);
/* The persistent form (tile set, cut) the fastest in the last autotune, or the one set last. */
int qp3d_batch_persistent_choice(const qp3d_batch *b, int *set, int *share);
/* bytes[0] gradient-max bytes, bytes[1] update bytes per step (all shots), bytes[2] updated
 * cells, bytes[3] interior cells. */
int qp3d_batch_step_bytes(const qp3d_batch *b, double *bytes);

#ifdef __cplusplus
}
#endif
#endif /* QP3D_H */
