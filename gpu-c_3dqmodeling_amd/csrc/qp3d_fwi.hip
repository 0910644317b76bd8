// Adjoint-state FWI gradient on top of a forward plan: a checkpointed forward run, then
// segments in reverse order, each recomputed from its checkpoint (keeping u_k of every step
// of the segment) and back-propagated with the adjoint kernels of qp3d_adjoint.hpp. Gradient
// terms are summed in FP32 within a segment and accumulated over time in FP64.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "qp3d_adjoint.hpp"
#include "qp3d_plan_impl.hpp"

namespace {

constexpr int TILE_PROD = 1;   // Tile<16, 1> for k_adj_prod (interior and PML lists)
constexpr int TILE_APPLY = 3;  // Tile<8, 1> for k_adj_apply
using TP = Tile<16, 1>;
using TA = Tile<8, 1>;

int occupancy(const void *fn, int threads)
{
  int occ = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, threads, 0) != hipSuccess) occ = 1;
  return occ > 0 ? occ : 1;
}

}  // namespace

struct qp3d_fwi {
  qp3d_plan *p = nullptr;
  int interval = 0, nseg = 0;
  int ckpt_storage = QP3D_STORAGE_FP32;
  std::vector<void *> ckpt;      // state block at the start of each segment
  // FP32 checkpoints: the plan's block maxima of the segment's first u (one slot: max |u|, the
  // increment, |grad u|), its map and rest-history word, so that the recomputation runs with the
  // same maps as the forward run (none: recomputation without them)
  std::vector<void *> camp;
  std::vector<char> amp_saved;  // [nseg] the record of segment s holds the forward run's
  // [keep][interval][N] u_k of segments: slot i holds the u_k of segment slot_seg[i] (-1: none).
  // The forward run fills slot i with segment nseg - keep + i (the last `keep` segments stay
  // resident in HBM and are back-propagated without recomputation); a recomputed segment goes
  // to its own slot when it has one, else to slot 0.
  float *ubuf = nullptr;
  int keep = 1;
  std::vector<int> slot_seg;
  float *adj = nullptr;          // adjoint state: u[2], w1[2], p[2], w3[2], w21 (9 N)
  float *prod = nullptr;         // Pxx Pyy Pzz Qx Qy Qz (6 N)
  float *gacc = nullptr;         // per-segment FP32 sums w.r.t. v2, eps, eps-del (3 N)
  double *gsum = nullptr;        // FP64 sums (3 N)
  float *resid = nullptr;        // [nt][hx][hy]
  WorkList wl_i, wl_p, wl_py, wl_a;  // interior, PML x and z faces, PML y faces, all of [5,n-5)^3
  // touched map of the adjoint sweep (AdjMap) and its state before the sweep (the receiver
  // plane's blocks marked); rest_map: the adjoint launches leave out the planes it shows at rest
  unsigned char *tmap = nullptr;
  std::vector<unsigned char> tmap0;
  int nbz = 0, nbx = 0, nby = 0;
  bool rest_map = true;
  bool forward_done = false;
  unsigned gen = 0;              // the plan's upload count at the forward run
  hipEvent_t ev[4] = {};
  // the forward run's copies of u_k into the segment buffer, on a stream of their own so that
  // each overlaps the step that reads u_k: ce[0] u_k complete, ce[1] its copy done
  hipStream_t cs = nullptr;
  hipEvent_t ce[2] = {};
  double ms_forward = 0, ms_recompute = 0, ms_adjoint = 0;
};

// Bytes of one checkpoint's block-maximum record: three arrays of one slot, the map, the word.
static size_t amp_bytes(const qp3d_plan *p)
{
  return sizeof(unsigned) * (3 * (size_t)p->ampg.nblk + 1) + (size_t)p->ampg.nblk;
}

static void fwi_free(qp3d_fwi *f)
{
  if (!f) return;
  for (void *c : f->ckpt) (void)hipFree(c);
  for (void *c : f->camp) (void)hipFree(c);
  (void)hipFree(f->ubuf);
  (void)hipFree(f->adj);
  (void)hipFree(f->prod);
  (void)hipFree(f->gacc);
  (void)hipFree(f->gsum);
  (void)hipFree(f->resid);
  (void)hipFree(f->tmap);
  (void)hipFree(f->wl_i.d);
  (void)hipFree(f->wl_p.d);
  (void)hipFree(f->wl_py.d);
  (void)hipFree(f->wl_a.d);
  for (hipEvent_t e : f->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : f->ce)
    if (e) (void)hipEventDestroy(e);
  if (f->cs) (void)hipStreamDestroy(f->cs);
  delete f;
}

extern "C" int qp3d_fwi_create(qp3d_plan *p, int interval, int ckpt_storage, qp3d_fwi **out)
{
  if (!p || !out) {
    qp3d_set_error("fwi_create: null argument");
    return QP3D_ERR_ARG;
  }
  *out = nullptr;
  const qp3d_config &c = p->cfg;
  if (c.storage != QP3D_STORAGE_FP32) {
    qp3d_set_error("fwi_create: the gradient needs a plan with FP32 storage");
    return QP3D_ERR_ARG;
  }
  if (c.nt < 1 || interval < 1 || (ckpt_storage != QP3D_STORAGE_FP32 && ckpt_storage != QP3D_STORAGE_BF16)) {
    qp3d_set_error("fwi_create: need nt >= 1, interval >= 1, checkpoint storage FP32 or BF16");
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(c.device));
  qp3d_fwi *f = new qp3d_fwi();
  f->p = p;
  f->interval = std::min(interval, c.nt);
  f->nseg = (c.nt + f->interval - 1) / f->interval;
  f->ckpt_storage = ckpt_storage;
  auto fail = [&](int rc) {
    fwi_free(f);
    return rc;
  };
  auto dmalloc = [&](void **ptr, size_t bytes, const char *what) -> int {
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
      qp3d_set_error("fwi_create: hipMalloc of %zu bytes for %s failed: %s", bytes, what, hipGetErrorString(e));
      *ptr = nullptr;
      return QP3D_ERR_ALLOC;
    }
    return 0;
  };
  const size_t N = (size_t)p->N;
  const size_t cb = p->state_bytes / (ckpt_storage == QP3D_STORAGE_FP32 ? 1 : 2);
  int rc;
  for (int s = 1; s < f->nseg; s++) {  // segment 0 starts from rest
    void *d = nullptr;
    if ((rc = dmalloc(&d, cb, "checkpoint"))) return fail(rc);
    f->ckpt.push_back(d);
    if (ckpt_storage == QP3D_STORAGE_FP32 && p->ampg.b && p->rflag) {
      void *m = nullptr;
      if ((rc = dmalloc(&m, amp_bytes(p), "checkpoint block maxima"))) return fail(rc);
      f->camp.push_back(m);
    }
  }
  if ((rc = dmalloc((void **)&f->ubuf, sizeof(float) * N * f->interval, "segment wavefields"))) return fail(rc);
  f->slot_seg.assign(1, -1);
  if ((rc = dmalloc((void **)&f->adj, sizeof(float) * 9 * N, "adjoint state"))) return fail(rc);
  if ((rc = dmalloc((void **)&f->prod, sizeof(float) * 6 * N, "product fields"))) return fail(rc);
  if ((rc = dmalloc((void **)&f->gacc, sizeof(float) * 3 * N, "segment gradient"))) return fail(rc);
  if ((rc = dmalloc((void **)&f->gsum, sizeof(double) * 3 * N, "gradient"))) return fail(rc);
  if ((rc = dmalloc((void **)&f->resid, sizeof(float) * (size_t)c.nt * p->hx * p->hy, "residuals"))) return fail(rc);
  // product fields stay zero outside [5,n-5)^3, where k_adj_prod never writes
  HIP_TRY(hipMemset(f->prod, 0, sizeof(float) * 6 * N));
  f->nbz = (c.nz + (1 << AB_Z) - 1) >> AB_Z;
  f->nbx = (c.nx + (1 << AB_X) - 1) >> AB_X;
  f->nby = (c.ny + (1 << AB_Y) - 1) >> AB_Y;
  const size_t nblk = (size_t)f->nbz * f->nbx * f->nby;
  if ((rc = dmalloc((void **)&f->tmap, nblk, "adjoint touched map"))) return fail(rc);
  f->tmap0.assign(nblk, 0);
  for (int xb = c.pml >> AB_X; xb <= (c.nx - c.pml - 1) >> AB_X; xb++)
    for (int yb = c.pml >> AB_Y; yb <= (c.ny - c.pml - 1) >> AB_Y; yb++)
      f->tmap0[((size_t)xb * f->nby + yb) * f->nbz + (c.sz >> AB_Z)] = 1;
  const int oi = occupancy(reinterpret_cast<const void *>(&k_adj_prod<TP, false>), TP::NT);
  const int op = occupancy(reinterpret_cast<const void *>(&k_adj_prod<TP, true>), TP::NT);
  const int oa = occupancy(reinterpret_cast<const void *>(&k_adj_apply<TA>), TA::NT);
  {
    // as in the forward plan, the y-face PML tiles go to a launch of their own
    std::vector<int4> it, ip, ipy;
    const int ncu = qp3d_cu_count(p);
    qp3d_items_update(p, TILE_PROD, TILE_PROD, oi * ncu, op * ncu, it, ip, ipy, TILE_PROD, op * ncu);
    if ((rc = qp3d_upload_items(f->wl_i, it, false)) || (rc = qp3d_upload_items(f->wl_p, ip, false)) ||
        (rc = qp3d_upload_items(f->wl_py, ipy, false)))
      return fail(rc);
  }
  if ((rc = qp3d_build_uniform_list(p, TILE_APPLY, oa, f->wl_a))) return fail(rc);
  for (hipEvent_t &e : f->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(QP3D_ERR_HIP);
  for (hipEvent_t &e : f->ce)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail(QP3D_ERR_HIP);
  if (hipStreamCreateWithFlags(&f->cs, hipStreamNonBlocking) != hipSuccess) return fail(QP3D_ERR_HIP);
  *out = f;
  return QP3D_OK;
}

extern "C" int qp3d_fwi_destroy(qp3d_fwi *f)
{
  if (!f) return QP3D_OK;
  (void)hipSetDevice(f->p->cfg.device);
  (void)hipStreamSynchronize(f->p->stream);
  fwi_free(f);
  return QP3D_OK;
}

static int grid_of(long n) { return (int)std::min<long>((n + 255) / 256, 8192); }

// The block-maximum record at the start of segment s (step k0 = s * interval): slot k0 % 3 of
// the three arrays (the maxima of u_k0 that step k0's map reads), the map, the history word.
static int save_amp(qp3d_fwi *f, int s)
{
  qp3d_plan *p = f->p;
  if (f->camp.empty() || !p->amp_ok) return QP3D_OK;
  const size_t nb = p->ampg.nblk, sl = (size_t)((s * f->interval) % 3);
  unsigned *d = (unsigned *)f->camp[s - 1];
  for (int a = 0; a < 3; a++)
    HIP_TRY(hipMemcpyAsync(d + a * nb, p->ampg.b + (3 * a + sl) * nb, sizeof(unsigned) * nb, hipMemcpyDeviceToDevice,
                           p->stream));
  HIP_TRY(hipMemcpyAsync(d + 3 * nb, p->rflag, sizeof(unsigned), hipMemcpyDeviceToDevice, p->stream));
  HIP_TRY(hipMemcpyAsync(d + 3 * nb + 1, p->ampg.q, nb, hipMemcpyDeviceToDevice, p->stream));
  return QP3D_OK;
}

// Restores the record of segment s (s = 0: the reset's), with the other two slots cleared;
// false when the segment has none (the recomputation then runs without the maps).
static bool load_amp(qp3d_fwi *f, int s, int &rc)
{
  qp3d_plan *p = f->p;
  rc = QP3D_OK;
  if (f->camp.empty() || !f->amp_saved[s]) return false;
  const size_t nb = p->ampg.nblk, sl = (size_t)((s * f->interval) % 3);
  auto fail = [&](hipError_t e) {
    qp3d_set_error("fwi_gradient: restoring the block maxima failed: %s", hipGetErrorString(e));
    rc = QP3D_ERR_HIP;
    return false;
  };
  hipError_t e = hipMemsetAsync(p->ampg.b, 0, sizeof(unsigned) * 9 * nb, p->stream);
  if (e != hipSuccess) return fail(e);
  if (s == 0) {
    if ((e = hipMemsetAsync(p->ampg.q, 0, nb, p->stream)) != hipSuccess ||
        (e = hipMemsetAsync(p->rflag, 1, sizeof(unsigned), p->stream)) != hipSuccess)
      return fail(e);
  } else {
    const unsigned *d = (const unsigned *)f->camp[s - 1];
    for (int a = 0; a < 3; a++)
      if ((e = hipMemcpyAsync(p->ampg.b + (3 * a + sl) * nb, d + a * nb, sizeof(unsigned) * nb,
                              hipMemcpyDeviceToDevice, p->stream)) != hipSuccess)
        return fail(e);
    if ((e = hipMemcpyAsync(p->rflag, d + 3 * nb, sizeof(unsigned), hipMemcpyDeviceToDevice, p->stream)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(p->ampg.q, d + 3 * nb + 1, nb, hipMemcpyDeviceToDevice, p->stream)) != hipSuccess)
      return fail(e);
  }
  p->amp_ok = true;
  return true;
}

static int save_ckpt(qp3d_fwi *f, int s)
{
  qp3d_plan *p = f->p;
  const long n = (long)(p->state_bytes / sizeof(float));  // the whole state block, pads included
  if (f->ckpt_storage == QP3D_STORAGE_FP32) {
    HIP_TRY(hipMemcpyAsync(f->ckpt[s - 1], p->state, p->state_bytes, hipMemcpyDeviceToDevice, p->stream));
    f->amp_saved[s] = !f->camp.empty() && p->amp_ok;
    int rc = save_amp(f, s);
    if (rc) return rc;
  } else {
    hipLaunchKernelGGL(k_pack_bf16, dim3(grid_of(n)), dim3(256), 0, p->stream, (unsigned short *)f->ckpt[s - 1],
                       (const float *)p->state, n);
    HIP_TRY(hipGetLastError());
  }
  return QP3D_OK;
}

static int load_ckpt(qp3d_fwi *f, int s)
{
  qp3d_plan *p = f->p;
  const long n = (long)(p->state_bytes / sizeof(float));
  if (s == 0) {
    HIP_TRY(hipMemsetAsync(p->state, 0, p->state_bytes, p->stream));
  } else if (f->ckpt_storage == QP3D_STORAGE_FP32) {
    HIP_TRY(hipMemcpyAsync(p->state, f->ckpt[s - 1], p->state_bytes, hipMemcpyDeviceToDevice, p->stream));
  } else {
    hipLaunchKernelGGL(k_unpack_bf16, dim3(grid_of(n)), dim3(256), 0, p->stream, (float *)p->state,
                       (const unsigned short *)f->ckpt[s - 1], n);
    HIP_TRY(hipGetLastError());
  }
  return QP3D_OK;
}

// Copies u_k (the plan's current level before step k) to dst on the copy stream, once the work
// queued on the plan's stream so far has run, so that it overlaps step k; step k + 1, which
// writes u_{k+2} over u_k, and every reader of dst wait for f->ce[1] on the plan's stream.
static int copy_uk(qp3d_fwi *f, float *dst, int k)
{
  qp3d_plan *p = f->p;
  HIP_TRY(hipEventRecord(f->ce[0], p->stream));
  HIP_TRY(hipStreamWaitEvent(f->cs, f->ce[0], 0));
  HIP_TRY(hipMemcpyAsync(dst, p->u[(k & 1) ^ 1], sizeof(float) * (size_t)p->N, hipMemcpyDeviceToDevice, f->cs));
  HIP_TRY(hipEventRecord(f->ce[1], f->cs));
  return QP3D_OK;
}

// Forward run from rest with a checkpoint at the start of every segment; the plan then holds
// the record, the per-step maxima and the final state as after qp3d_plan_run.
extern "C" int qp3d_fwi_forward(qp3d_fwi *f)
{
  if (!f) return QP3D_ERR_ARG;
  qp3d_plan *p = f->p;
  if (!p->uploaded) {
    qp3d_set_error("fwi_forward: upload the model first");
    return QP3D_ERR_STATE;
  }
  if (!f->ubuf || f->keep < 1) {
    qp3d_set_error("fwi_forward: no segment buffer (qp3d_fwi_set_resident_segments failed)");
    return QP3D_ERR_STATE;
  }
  HIP_TRY(hipSetDevice(p->cfg.device));
  int rc = qp3d_plan_reset(p);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(f->ev[0], p->stream));
  // segment 0 starts from the reset (its record: zeros, an empty map, a whole history)
  f->amp_saved.assign((size_t)f->nseg, 0);
  f->amp_saved[0] = !f->camp.empty() && p->amp_ok;
  const int klast = (f->nseg - f->keep) * f->interval;  // steps of the resident segments keep u_k
  const size_t N = (size_t)p->N;
  bool copying = false;  // a copy of u_{k-1} is under way on the copy stream
  for (int k = 0; k < p->cfg.nt; k++) {
    if (k % f->interval == 0 && k > 0 && (rc = save_ckpt(f, k / f->interval))) return rc;
    // step k writes u_{k+1} over u_{k-1}, whose copy must have finished
    if (copying) HIP_TRY(hipStreamWaitEvent(p->stream, f->ce[1], 0));
    if (k >= klast) {
      if ((rc = copy_uk(f, f->ubuf + (size_t)(k - klast) * N, k))) return rc;
      copying = true;
    }
    if ((rc = qp3d_forward_step_fwi(p, k, false))) return rc;
  }
  if (copying) HIP_TRY(hipStreamWaitEvent(p->stream, f->ce[1], 0));
  for (int i = 0; i < f->keep; i++) f->slot_seg[i] = f->nseg - f->keep + i;
  f->gen = p->upload_gen;
  p->steps_done = p->cfg.nt;
  HIP_TRY(hipEventRecord(f->ev[1], p->stream));
  rc = qp3d_plan_sync(p);
  if (rc) return rc;
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, f->ev[0], f->ev[1]));
  f->ms_forward = ms;
  f->forward_done = true;
  return QP3D_OK;
}

static int adjoint_step(qp3d_fwi *f, int k, const float *uk)
{
  qp3d_plan *p = f->p;
  const size_t N = (size_t)p->N;
  const int nw = k & 1, cw = nw ^ 1;  // slot of the newest level / of level k
  float *A = f->adj;
  AdjArgs a;
  a.g = p->geo;
  a.c = p->coef;
  a.m = AdjMap{f->rest_map ? f->tmap : nullptr, f->nbz, f->nbx, f->nby};
  a.u = uk;
  a.vdt2 = p->vdt2; a.eps = p->eps; a.epsd = p->epsd;
  a.pmltab = p->pmltab;
  a.gmax = p->gmax + k;
  a.an = A + nw * N; a.ac = A + cw * N;
  a.o1n = A + (2 + nw) * N; a.o1c = A + (2 + cw) * N;
  a.pin = A + (4 + nw) * N; a.pic = A + (4 + cw) * N;
  a.o3n = A + (6 + nw) * N; a.o3c = A + (6 + cw) * N;
  a.o21 = A + 8 * N;
  a.pxx = f->prod; a.pyy = f->prod + N; a.pzz = f->prod + 2 * N;
  a.qx = f->prod + 3 * N; a.qy = f->prod + 4 * N; a.qz = f->prod + 5 * N;
  a.gv2 = f->gacc; a.ge = f->gacc + N; a.ged = f->gacc + 2 * N;
  a.resid = f->resid + (size_t)k * p->hx * p->hy;
  a.sz = p->cfg.sz;
  if (f->wl_i.n) {
    a.work = f->wl_i.d;
    hipLaunchKernelGGL((k_adj_prod<TP, false>), dim3(f->wl_i.n), dim3(TY, TP::WX), 0, p->stream, a);
    HIP_TRY(hipGetLastError());
  }
  for (const WorkList *w : {&f->wl_p, &f->wl_py}) {
    if (!w->n) continue;
    a.work = w->d;
    hipLaunchKernelGGL((k_adj_prod<TP, true>), dim3(w->n), dim3(TY, TP::WX), 0, p->stream, a);
    HIP_TRY(hipGetLastError());
  }
  ApplyArgs b;
  b.g = p->geo;
  b.c = p->coef;
  b.pxx = a.pxx; b.pyy = a.pyy; b.pzz = a.pzz; b.qx = a.qx; b.qy = a.qy; b.qz = a.qz;
  b.ac = a.ac;
  b.m = a.m;
  b.work = f->wl_a.d;
  hipLaunchKernelGGL((k_adj_apply<TA>), dim3(f->wl_a.n), dim3(TY, TA::WX), 0, p->stream, b);
  HIP_TRY(hipGetLastError());
  return QP3D_OK;
}

// Gradient of J = 1/2 sum |record - observed|^2 given resid = record - observed ([nt][hx][hy],
// host), with respect to the padded-grid vp, epsilon and delta (host, [nz][nx][ny] each).
extern "C" int qp3d_fwi_gradient(qp3d_fwi *f, const float *resid, double *g_vp, double *g_eps, double *g_del)
{
  if (!f || !resid || !g_vp || !g_eps || !g_del) {
    qp3d_set_error("fwi_gradient: null argument");
    return QP3D_ERR_ARG;
  }
  if (!f->forward_done || f->gen != f->p->upload_gen) {
    qp3d_set_error("fwi_gradient: run qp3d_fwi_forward first (and again after every upload)");
    return QP3D_ERR_STATE;
  }
  qp3d_plan *p = f->p;
  const qp3d_config &c = p->cfg;
  const size_t N = (size_t)p->N;
  HIP_TRY(hipSetDevice(c.device));
  // resid and the gradient arrays may be host or device memory
  HIP_TRY(hipMemcpyAsync(f->resid, resid, sizeof(float) * (size_t)c.nt * p->hx * p->hy, hipMemcpyDefault,
                         p->stream));
  HIP_TRY(hipMemsetAsync(f->adj, 0, sizeof(float) * 9 * N, p->stream));
  HIP_TRY(hipMemsetAsync(f->gsum, 0, sizeof(double) * 3 * N, p->stream));
  if (f->rest_map) {
    // the map's invariant: every nonzero adjoint value lies in a marked block (the product
    // fields hold the last sweep's values until cleared)
    HIP_TRY(hipMemsetAsync(f->prod, 0, sizeof(float) * 6 * N, p->stream));
    HIP_TRY(hipMemcpyAsync(f->tmap, f->tmap0.data(), f->tmap0.size(), hipMemcpyHostToDevice, p->stream));
  }
  double ms_re = 0, ms_adj = 0;
  int rc, last_re = -1;  // the last segment recomputed
  for (int s = f->nseg - 1; s >= 0; s--) {
    const int k0 = s * f->interval, k1 = std::min(k0 + f->interval, c.nt);
    HIP_TRY(hipEventRecord(f->ev[0], p->stream));
    const int slot = s >= f->nseg - f->keep ? s - (f->nseg - f->keep) : 0;
    float *const ub = f->ubuf + (size_t)slot * f->interval * N;
    if (f->slot_seg[slot] != s) {
      if ((rc = load_ckpt(f, s))) return rc;
      const bool maps = load_amp(f, s, rc);
      if (rc) return rc;
      for (int k = k0; k < k1; k++) {
        // step k writes u_{k+1} over u_{k-1}, whose copy must have finished
        if (k > k0) HIP_TRY(hipStreamWaitEvent(p->stream, f->ce[1], 0));
        if ((rc = copy_uk(f, ub + (size_t)(k - k0) * N, k))) return rc;
        // G_k from the forward run; the forward run's maps when its record was restored
        if ((rc = qp3d_forward_step_fwi(p, k, true, maps))) return rc;
      }
      HIP_TRY(hipStreamWaitEvent(p->stream, f->ce[1], 0));  // the adjoint sweep reads ub
      f->slot_seg[slot] = s;
      last_re = s;
    }
    HIP_TRY(hipEventRecord(f->ev[1], p->stream));
    HIP_TRY(hipMemsetAsync(f->gacc, 0, sizeof(float) * 3 * N, p->stream));
    for (int k = k1 - 1; k >= k0; k--)
      if ((rc = adjoint_step(f, k, ub + (size_t)(k - k0) * N))) return rc;
    hipLaunchKernelGGL(k_fold, dim3(grid_of(3 * (long)N)), dim3(256), 0, p->stream, f->gsum, (const float *)f->gacc,
                       3 * (long)N);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(f->ev[2], p->stream));
    HIP_TRY(hipEventSynchronize(f->ev[2]));
    float a = 0, b = 0;
    HIP_TRY(hipEventElapsedTime(&a, f->ev[0], f->ev[1]));
    HIP_TRY(hipEventElapsedTime(&b, f->ev[1], f->ev[2]));
    ms_re += a;
    ms_adj += b;
  }
  hipLaunchKernelGGL(k_grad_finish, dim3(grid_of((long)N)), dim3(256), 0, p->stream, f->gsum,
                     (const float *)p->vdt2, c.dt, (long)N);
  HIP_TRY(hipGetLastError());
  // [nz][nx][ny] outputs from the row-pitched sums
  double *outs[3] = {g_vp, g_eps, g_del};
  for (int i = 0; i < 3; i++)
    if ((rc = qp3d_copy_field(p, outs[i], f->gsum + i * N, sizeof(double), false, p->stream))) return rc;
  HIP_TRY(hipStreamSynchronize(p->stream));
  f->ms_recompute = ms_re;
  f->ms_adjoint = ms_adj;
  // the plan's state holds the end of the last segment recomputed (or of the forward run)
  if (last_re >= 0) p->steps_done = std::min((last_re + 1) * f->interval, c.nt);
  return QP3D_OK;
}

// Segments whose u_k the forward run keeps resident in HBM (the last n; 1 by default), so that
// the gradient back-propagates them without recomputation; n < 0: as many as fit in the free
// device memory less 2 GiB. Reallocates the segment buffer; the next gradient needs a forward
// run after it. *kept (may be null) receives the number chosen.
extern "C" int qp3d_fwi_set_resident_segments(qp3d_fwi *f, int n, int *kept)
{
  if (!f) return QP3D_ERR_ARG;
  qp3d_plan *p = f->p;
  HIP_TRY(hipSetDevice(p->cfg.device));
  HIP_TRY(hipStreamSynchronize(p->stream));
  const size_t seg = sizeof(float) * (size_t)p->N * f->interval;
  if (n < 0) {
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    const size_t margin = (size_t)2 << 30, have = fr + seg * f->keep;
    n = have > margin ? (int)std::min<size_t>((have - margin) / seg, (size_t)f->nseg) : 1;
  }
  if (n < 1 || n > f->nseg) {
    qp3d_set_error("fwi_set_resident_segments: need 1 <= n <= %d segments (got %d)", f->nseg, n);
    return QP3D_ERR_ARG;
  }
  if (n != f->keep) {
    // the old buffer goes first: both may not fit at once
    (void)hipFree(f->ubuf);
    f->ubuf = nullptr;
    hipError_t e = hipMalloc((void **)&f->ubuf, seg * n);
    if (e != hipSuccess) {
      f->ubuf = nullptr;
      (void)hipGetLastError();
      const int want = n;
      n = 1;
      e = hipMalloc((void **)&f->ubuf, seg);
      f->keep = e == hipSuccess ? 1 : 0;
      f->slot_seg.assign(1, -1);
      f->forward_done = false;
      qp3d_set_error("fwi_set_resident_segments: hipMalloc of %zu bytes for %d segments failed%s", seg * want, want,
                     e == hipSuccess ? " (one segment kept)" : "; no segment buffer left");
      return QP3D_ERR_ALLOC;
    }
    f->keep = n;
  }
  f->slot_seg.assign((size_t)f->keep, -1);
  f->forward_done = false;
  if (kept) *kept = f->keep;
  return QP3D_OK;
}

// Whether the adjoint sweep leaves out the planes that its touched map shows at rest (on by
// default; the gradient is the same either way, up to the sign of zeros).
extern "C" int qp3d_fwi_set_rest_map(qp3d_fwi *f, int on)
{
  if (!f) return QP3D_ERR_ARG;
  f->rest_map = on != 0;
  return QP3D_OK;
}

// Device time (ms) of the last checkpointed forward run, of the recomputation and of the
// adjoint sweep of the last gradient.
extern "C" int qp3d_fwi_times(const qp3d_fwi *f, double *ms)
{
  if (!f || !ms) return QP3D_ERR_ARG;
  ms[0] = f->ms_forward;
  ms[1] = f->ms_recompute;
  ms[2] = f->ms_adjoint;
  return QP3D_OK;
}
