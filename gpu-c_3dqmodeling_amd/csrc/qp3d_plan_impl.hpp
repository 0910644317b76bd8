// Internal header of the qP3D HIP library: the plan object and the helpers shared by the
// forward time loop (qp3d_plan.hip) and the adjoint gradient (qp3d_fwi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "qp3d.h"
#include "qp3d_kernels.hpp"

using namespace qp3d;

#define HIP_TRY(expr)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) {                                                                            \
      qp3d_set_error("%s failed at %s:%d: %s", #expr, __FILE__, __LINE__, hipGetErrorString(e_));     \
      return QP3D_ERR_HIP;                                                                             \
    }                                                                                                  \
  } while (0)

namespace qp3d_impl {

constexpr int GRAPH_LEN = 10;   // time steps per captured graph
constexpr int AMP_ZCHUNK = 32;  // z planes per item of the gradient pass under the amplitude bound
constexpr int EV_PER_STEP = 5;  // instrumentation events per step
// words of the plan's small device block: largest-tile keys [2], two spare, bound-pass maxima
// (two 32-bit words), the rest-history word
constexpr int ARG_WORDS = 6;
constexpr size_t CNT_BYTES = sizeof(unsigned long long) * qp3d::CNT_SLOTS * qp3d::CNT_STRIDE;

struct WorkList {
  int4 *d = nullptr;
  int n = 0;       // launched workgroups (multiple of 8, padding items are empty)
  long cells = 0;  // cells updated by one launch
};

}  // namespace qp3d_impl

using qp3d_impl::WorkList;
using qp3d_impl::GRAPH_LEN;
using qp3d_impl::EV_PER_STEP;
using qp3d_impl::AMP_ZCHUNK;
using qp3d_impl::ARG_WORDS;
using qp3d_impl::CNT_BYTES;

struct qp3d_plan {
  qp3d_config cfg{};
  int hx = 0, hy = 0, hz = 0;
  long N = 0;   // cells of a device field: nz * nx * ys (rows padded to geo.ys)
  long Nd = 0;  // cells of a dense [nz][nx][ny] array (the caller's layout)
  Geo geo{};
  Coef coef{};
  int tile_u = 0, tile_g = 0, tile_p = 0;
  bool two_streams = false;
  bool cap_first = false;  // capturing the first step of a graph: u(k) needs no wait on ev_pml
  hipStream_t stream2 = nullptr;                          // PML updates when two_streams
  hipEvent_t ev_grad = nullptr, ev_pml = nullptr;         // cross-stream ordering, no timing
  hipEvent_t ev_in = nullptr, ev_out = nullptr;           // ordering with a caller's stream
  WorkList wl_grad, wl_int, wl_pml, wl_pmly;  // wl_pml: x and z faces; wl_pmly: y faces
  WorkList wl_grad_amp;                       // gradient pass under the amplitude bound
  int tile_py = 1;
  long n_int_cells = 0, n_pml_cells = 0;
  hipStream_t stream = nullptr;

  // State block: u[2], w1[2], w3[2], p[2], w21 (9 fields of N cells of esz bytes: float,
  // bfloat16 or half), then with folded y faces (yw > 0) the compact y-face PML state yw1[2],
  // yw3[2], ypp[2], yw21 (7 fields of ycells cells, YCompact layout); the full-grid fields'
  // y-face cells are then unused.
  char *state = nullptr;
  size_t esz = 4;
  size_t fstride = 0;      // bytes from one full-grid state field to the next (N*esz plus a pad)
  size_t ystride = 0;      // bytes from one compact field to the next (0 without compact arrays)
  size_t state_bytes = 0;  // 9 * fstride + 7 * ystride
  void *u[2]{}, *w1[2]{}, *w3[2]{}, *pp[2]{}, *w21 = nullptr;
  void *yw1[2]{}, *yw3[2]{}, *ypp[2]{}, *yw21 = nullptr;
  int yw = 0;           // cells per y face in the compact arrays (pml - 5), 0 = none
  unsigned ypln = 0;    // compact cells per z plane, (nx - 2 pml) * 2 yw
  long ycells = 0;      // nz * ypln
  int scale_log2 = 0;  // half storage: fields are stored times 2^scale_log2
  // model block: vdt2, eps, epsd over the grid, then their compact y-face copies
  float *model = nullptr;
  size_t model_bytes = 0;
  float *vdt2 = nullptr, *eps = nullptr, *epsd = nullptr;
  float *yvdt2 = nullptr, *yeps = nullptr, *yepsd = nullptr;
  float4 *pmltab = nullptr;
  unsigned *gmax = nullptr;  // [nt]
  int *flag = nullptr;
  // amplitude bound of the gradient pass (AmpGrid): block maxima of |u| in three buffers, the
  // key of each step's largest tile in two; amp_ok: every cell of the current u is covered by
  // its buffer (false after steps run without the bound, e.g. FWI recomputation)
  AmpGrid ampg{};
  unsigned long long *argkey = nullptr;
  unsigned *rflag = nullptr;  // nonzero while every step since the reset ran k_quiet (rest history)
  // counters of instrumented steps since the reset, CNT_SLOTS x CNT_STRIDE words (cnt_slot):
  // gradient tiles evaluated and skipped, update cells per kind
  unsigned long long *cnt = nullptr;
  bool amp_ok = true;
  bool amp_on = true;  // qp3d_plan_set_amp_bound
  bool spec = false;             // gmax_mode 2 (a batch turns it off for its plans)
  int in_batch = 0;              // 1 while a live batch drives the plan
  unsigned *fixmask = nullptr;   // gmax_mode 2: one bit per cell, all clear between steps
  float *gpred = nullptr;        // gmax_mode 2: [nt] predicted G per step
  long fixwords = 0;
  float *rec = nullptr;      // [nt][hx][hy]
  float *sbox = nullptr;
  int sb[6]{};               // z0 z1 x0 x1 y0 y1
  bool src_on = false, src_rim = false;
  float *wav = nullptr;      // [nt] wavelet on the device
  int *kdev = nullptr;       // step index read by graph-replayed launches

  // use_graph: GRAPH_LEN consecutive steps (starting at an even step) captured once and
  // replayed with the step index set on the device before each replay
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;

  std::vector<float *> snaps;
  std::vector<float *> snaps_s;  // S-field snapshots (cfg.snap_s), [nz][nx][ny] each
  std::vector<int> snap_steps;
  int nsnap_taken = 0;

  bool uploaded = false;
  unsigned upload_gen = 0;  // counts uploads (objects built on the plan check it)
  int steps_done = 0;  // one past the last step enqueued since the last reset

  int instrument = 0;  // 1: events around every stencil launch; 2: around the interior update only
  std::vector<hipEvent_t> ev;   // EV_PER_STEP per step: before gradmax, after gradmax, after interior,
                                // after PML, after the fix-up (gmax_mode 2)
  std::vector<int> ev_pending;  // steps with recorded events not yet harvested
  double t_grad = 0, t_int = 0, t_pml = 0, t_upd = 0;
  std::vector<float> t_int_step;  // [nt] interior-update time of each instrumented step (ms)
  long n_launch = 0;
};

// Byte stride between fields that one kernel streams side by side: the size rounded up to 256
// bytes plus 12.25 KiB, so that the same cell of different fields does not sit at the same
// offset modulo large powers of two (fields of exactly 2^k-multiple sizes would otherwise start
// on the same HBM channel and bank together).
inline size_t field_stride(size_t bytes) { return (bytes + 255) / 256 * 256 + 12544; }

// Interior-x extent (rows) of tile shape t.
inline int tile_tx(int t)
{
  static const int tx[] = {
#define QP3D_TX(wx, rx) wx * rx,
      QP3D_TILES(QP3D_TX)
#undef QP3D_TX
  };
  return tx[t];
}

// Calls f with a value of the storage policy of a QP3D_STORAGE_* mode.
template <class F>
inline void with_storage(int storage, F f)
{
  if (storage == QP3D_STORAGE_BF16)
    f(StoreBF16{});
  else if (storage == QP3D_STORAGE_FP16)
    f(StoreF16{});
  else
    f(StoreF32{});
}

// Kernel arguments of a step of parity `prev` (slot prev holds u(k-1) and receives u(k+1)).
template <class S>
inline void make_args(const qp3d_plan *p, int prev, Step st, GradArgs<S> &ga, UpdArgs<S> &ua)
{
  using T = typename S::T;
  const int cur = prev ^ 1;
  ga.g = p->geo;
  ga.c = p->coef;
  ga.ucur = (const T *)p->u[cur];
  ga.gmax = p->gmax;
  ga.flag = p->flag;
  ga.work = p->wl_grad.d;
  ga.step = st;
  ga.fixmask = p->fixmask;
  ga.gpred = p->gpred;
  ga.amp = p->ampg;  // the caller turns skipping on (ga.skip) and may turn the bound off
  ga.skip = 0;
  ga.argkey = p->argkey;
  ga.nwork = p->wl_grad.n;
  ga.gcount = p->cnt;
  ga.gseed = p->argkey ? (unsigned *)(p->argkey + 4) : nullptr;
  ga.gseed_out = nullptr;
  ua.g = p->geo;
  ua.c = p->coef;
  ua.ucur = (const T *)p->u[cur];
  ua.unext = (T *)p->u[prev];
  ua.vdt2 = p->vdt2; ua.eps = p->eps; ua.epsd = p->epsd;
  ua.w1p = (T *)p->w1[prev]; ua.w1c = (const T *)p->w1[cur];
  ua.w3p = (T *)p->w3[prev]; ua.w3c = (const T *)p->w3[cur];
  ua.pp = (T *)p->pp[prev]; ua.pc = (const T *)p->pp[cur];
  ua.w21 = (T *)p->w21;
  ua.yc.w1p = (T *)p->yw1[prev]; ua.yc.w1c = (const T *)p->yw1[cur];
  ua.yc.w3p = (T *)p->yw3[prev]; ua.yc.w3c = (const T *)p->yw3[cur];
  ua.yc.pp = (T *)p->ypp[prev]; ua.yc.pc = (const T *)p->ypp[cur];
  ua.yc.w21 = (T *)p->yw21;
  ua.yc.vdt2 = p->yvdt2; ua.yc.eps = p->yeps; ua.yc.epsd = p->yepsd;
  ua.yc.pln = p->ypln;
  ua.yc.yw = p->yw;
  ua.pmltab = p->pmltab;
  ua.gmax = p->gmax;
  ua.flag = p->flag;
  ua.fixmask = p->fixmask;
  ua.gpred = p->gpred;
  ua.wav = p->wav;
  ua.step = st;
  ua.src_on = p->src_on ? 1 : 0;
  ua.sb_z0 = p->sb[0]; ua.sb_z1 = p->sb[1]; ua.sb_x0 = p->sb[2]; ua.sb_x1 = p->sb[3];
  ua.sb_y0 = p->sb[4]; ua.sb_y1 = p->sb[5];
  ua.sbox = p->sbox;
  ua.rec = p->rec;
  ua.sz = p->cfg.sz;
  ua.work = nullptr;
  ua.amp = p->ampg;
  ua.amp.q = nullptr;  // the step turns the quiet map on after k_quiet has filled it
  ua.rest_off = nullptr;
  ua.ucount = nullptr;
}

// Sizes of the state and model blocks (esz and the geometry set): fstride, ystride,
// state_bytes, model_bytes (qp3d_plan_describe calls it). qp3d_bind_state points the fields into
// blocks of those sizes; qp3d_bind_ymodel only the compact model copies (3 fields of
// field_stride(4 * ycells) bytes).
void qp3d_state_layout(qp3d_plan *p);
void qp3d_bind_state(qp3d_plan *p, char *state, float *model);
void qp3d_bind_ymodel(qp3d_plan *p, float *ym);
// Fills the compact y-face model copies from the full-grid model (after every model change).
void qp3d_launch_ymodel(const qp3d_plan *p, hipStream_t s);
// Copies between a dense [nz][nx][ny] array of esz-byte elements (2, 4 or 8; host or device
// memory) and a device field in the plan's row-pitched layout, on stream s; to_field: dense ->
// field. Complete on return when the dense array is host memory; device to device copies are
// only queued. Needs no allocation when the dense array is device memory or the rows are not
// padded.
int qp3d_copy_field(const qp3d_plan *p, void *dense, void *field, size_t esz, bool to_field, hipStream_t s);
// dense array in device memory of the current device
bool qp3d_on_device(const void *ptr);

// Work lists (qp3d_plan.hip): a uniform list over [5,n-5)^3 for a tile shape.
int qp3d_build_uniform_list(const qp3d_plan *p, int tile, int occ, WorkList &w);
// The same items in host memory (logical order) for `capacity` resident workgroups.
void qp3d_items_uniform(const qp3d_plan *p, int tile, int capacity, std::vector<int4> &items, int zchunk = 0);
// Interior update items (tile_i, cap_i workgroups) into it, PML items of the x and z faces
// (tile_p, cap_p) into ip, of the y faces (tile_py, cap_py: a launch of their own, folded tiles
// when y_fold(pml)) into ipy.
void qp3d_items_update(const qp3d_plan *p, int tile_i, int tile_p, int cap_i, int cap_p, std::vector<int4> &it,
                       std::vector<int4> &ip, std::vector<int4> &ipy, int tile_py, int cap_py);
// Uploads items in XCD-aware order (batch: items hold z0 | z1 << 16 and a grid index).
int qp3d_upload_items(WorkList &w, const std::vector<int4> &logical, bool batch);
int qp3d_cu_count(const qp3d_plan *p);
// Enqueues forward time step k with direct launches (no capture, no snapshot, no events).
int qp3d_forward_step(qp3d_plan *p, int k);
// FWI steps: recompute = false, the forward run with the amplitude bound; true, a recomputed
// step that takes G_k from gmax[k] (left there by the forward run) instead of a gradient pass.
int qp3d_forward_step_fwi(qp3d_plan *p, int k, bool recompute, bool maps = false);
// The same on stream `on` (one stream), with the step index the kernels read given by `st`.
int qp3d_forward_step_on(qp3d_plan *p, int k, Step st, hipStream_t on);
// Validates a configuration (no device access); `who` prefixes the error message.
int qp3d_check_config(const qp3d_config &c, const char *who);
// Host-side description of a plan for a valid configuration (geometry, tiles, modes); no
// allocation.
void qp3d_plan_describe(qp3d_plan *p, const qp3d_config &c);
// Work items of the plan's four launches (gradient max, interior, PML x/z faces, PML y faces)
// in logical order, for the resident-workgroup capacity of the plan's device.
void qp3d_plan_items(const qp3d_plan *p, std::vector<int4> (&items)[4]);
// Items in XCD-aware launch order (padded to a multiple of 8); returns the cells they cover.
long qp3d_order_items(const std::vector<int4> &logical, bool batch, std::vector<int4> &ordered);
// PML coefficient table (3 x pml float4) and p->coef from profiles and FD tables in host or
// device memory.
int qp3d_make_tables(qp3d_plan *p, const float *dlr, const float *ddlr, const float *dtb, const float *ddtb,
                     const float *dfb, const float *ddfb, const float *c, const float *c2, std::vector<float4> &tab);
// In place over (vp, eps, del) copies: v2 = (vp dt)^2, ed = eps - del.
void qp3d_launch_model(float *v2, const float *eps, float *ed, float dt, long n, hipStream_t s);
// Bounding box {z0, z1, x0, x1, y0, y1} of the non-zero cells of a device source array into
// device memory `box`, which must hold {nz, -1, nx, -1, ny, -1} beforehand.
void qp3d_launch_source_box(const float *src, int nx, int ny, int nz, int *box, hipStream_t s);
