// Host side of the qP3D HIP library: plan lifecycle, uploads, the time loop, and the
// C-ABI entry points declared in include/qp3d.h (including the drop-in modeling3d()).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "qp3d.h"
#include "qp3d_kernels.hpp"

using namespace qp3d;

// ---------------------------------------------------------------------------------------
// errors
static thread_local char g_err[1024];

extern "C" const char *qp3d_last_error(void) { return g_err; }

extern "C" void qp3d_set_error(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      qp3d_set_error("%s failed at %s:%d: %s", #expr, __FILE__, __LINE__, hipGetErrorString(e_)); \
      return QP3D_ERR_HIP;                                                                   \
    }                                                                                        \
  } while (0)

// ---------------------------------------------------------------------------------------
// plan
struct qp3d_plan {
  qp3d_config cfg{};
  int hx = 0, hy = 0, hz = 0;
  long N = 0;
  Geo geo{};
  Coef coef{};
  int nblk = 0;  // launched workgroups (multiple of 8)
  hipStream_t stream = nullptr;

  float *state = nullptr;  // u[2], w1[2], w3[2], p[2], w21: 9 fields of N floats
  float *u[2]{}, *w1[2]{}, *w3[2]{}, *pp[2]{}, *w21 = nullptr;
  float *model = nullptr;  // vdt2, eps, epsd
  float *vdt2 = nullptr, *eps = nullptr, *epsd = nullptr;
  float4 *pmltab = nullptr;
  unsigned *gmax = nullptr;  // [nt]
  int *flag = nullptr;
  float *rec = nullptr;      // [nt][hx][hy]
  float *sbox = nullptr;
  int sb[6]{};               // z0 z1 x0 x1 y0 y1
  bool src_on = false, src_rim = false;
  std::vector<float> wav;

  std::vector<float *> snaps;
  std::vector<int> snap_steps;
  int nsnap_taken = 0;

  bool uploaded = false;
  int steps_done = 0;  // one past the last step enqueued since the last reset

  bool instrument = false;
  std::vector<hipEvent_t> ev;  // 3 per step: before gradmax, between, after update
  std::vector<int> ev_pending; // steps with recorded events not yet harvested
  double t_grad = 0, t_upd = 0;
  long n_launch = 0;
};

static int pick_zchunk(const qp3d_config &c, int ntile, int nz5)
{
  if (c.zchunk > 0) return c.zchunk;
  // Aim for >= 4 workgroups per CU on a 256-CU part, but keep chunks long enough that the
  // 10-plane queue priming stays a small fraction of the work.
  int want = (1024 + ntile - 1) / ntile;
  int zc = (nz5 + want - 1) / want;
  if (zc < 32) zc = 32;
  if (zc > nz5) zc = nz5;
  return zc;
}

extern "C" int qp3d_plan_create(const qp3d_config *cfg, qp3d_plan **out)
{
  if (!cfg || !out) {
    qp3d_set_error("plan_create: null argument");
    return QP3D_ERR_ARG;
  }
  *out = nullptr;
  const qp3d_config &c = *cfg;
  if (c.pml < 6 || c.nx < 2 * c.pml + 1 || c.ny < 2 * c.pml + 1 || c.nz < 2 * c.pml + 1 || c.nt < 0) {
    qp3d_set_error("plan_create: need pml >= 6 and nx, ny, nz > 2*pml (got nx=%d ny=%d nz=%d pml=%d nt=%d)", c.nx,
                   c.ny, c.nz, c.pml, c.nt);
    return QP3D_ERR_ARG;
  }
  if (!(c.dx > 0 && c.dy > 0 && c.dz > 0 && c.dt > 0)) {
    qp3d_set_error("plan_create: spacings and dt must be positive");
    return QP3D_ERR_ARG;
  }
  if (c.sz < 0 || c.sz >= c.nz) {
    qp3d_set_error("plan_create: receiver plane sz=%d outside [0,%d)", c.sz, c.nz);
    return QP3D_ERR_ARG;
  }
  if (c.snapflag == 1 && c.ntsnap <= 0) {
    qp3d_set_error("plan_create: ntsnap must be positive when snapflag == 1");
    return QP3D_ERR_ARG;
  }
  if (c.storage != QP3D_STORAGE_FP32) {
    qp3d_set_error("plan_create: storage mode %d is not available in this build", c.storage);
    return QP3D_ERR_ARG;
  }
  const double cells = (double)c.nx * c.ny * c.nz;
  if (cells > 4.0e9) {
    qp3d_set_error("plan_create: grid of %.3g cells exceeds the 4e9-cell limit", cells);
    return QP3D_ERR_ARG;
  }
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (c.device < 0 || c.device >= ndev) {
    qp3d_set_error("plan_create: device %d not present (%d devices)", c.device, ndev);
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(c.device));

  qp3d_plan *p = new qp3d_plan();
  p->cfg = c;
  p->hx = c.nx - 2 * c.pml;
  p->hy = c.ny - 2 * c.pml;
  p->hz = c.nz - 2 * c.pml;
  p->N = (long)c.nx * c.ny * c.nz;
  Geo &g = p->geo;
  g.nx = c.nx; g.ny = c.ny; g.nz = c.nz; g.pml = c.pml;
  g.pln = (long)c.nx * c.ny;
  const int nx5 = c.nx - 2 * H, ny5 = c.ny - 2 * H, nz5 = c.nz - 2 * H;
  g.tiles_y = (ny5 + TY - 1) / TY;
  g.tiles_x = (nx5 + TX - 1) / TX;
  g.zchunk = pick_zchunk(c, g.tiles_x * g.tiles_y, nz5);
  const int nzc = (nz5 + g.zchunk - 1) / g.zchunk;
  g.nlog = g.tiles_x * g.tiles_y * nzc;
  p->nblk = (g.nlog + 7) / 8 * 8;

  auto fail = [&](int rc) {
    qp3d_plan_destroy(p);
    return rc;
  };
  auto dmalloc = [&](void **ptr, size_t bytes, const char *what) -> int {
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
      qp3d_set_error("plan_create: hipMalloc of %zu bytes for %s failed: %s", bytes, what, hipGetErrorString(e));
      *ptr = nullptr;
      return QP3D_ERR_ALLOC;
    }
    return 0;
  };
  const size_t fb = sizeof(float) * (size_t)p->N;
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    qp3d_set_error("plan_create: hipStreamCreate failed");
    return fail(QP3D_ERR_HIP);
  }
  int rc;
  if ((rc = dmalloc((void **)&p->state, 9 * fb, "wavefields and PML state"))) return fail(rc);
  float *s = p->state;
  p->u[0] = s; p->u[1] = s + p->N;
  p->w1[0] = s + 2 * p->N; p->w1[1] = s + 3 * p->N;
  p->w3[0] = s + 4 * p->N; p->w3[1] = s + 5 * p->N;
  p->pp[0] = s + 6 * p->N; p->pp[1] = s + 7 * p->N;
  p->w21 = s + 8 * p->N;
  if ((rc = dmalloc((void **)&p->model, 3 * fb, "model"))) return fail(rc);
  p->vdt2 = p->model; p->eps = p->model + p->N; p->epsd = p->model + 2 * p->N;
  if ((rc = dmalloc((void **)&p->pmltab, sizeof(float4) * 3 * c.pml, "PML table"))) return fail(rc);
  if ((rc = dmalloc((void **)&p->gmax, sizeof(unsigned) * (size_t)(c.nt > 0 ? c.nt : 1), "step maxima")))
    return fail(rc);
  if ((rc = dmalloc((void **)&p->flag, sizeof(int), "flag"))) return fail(rc);
  const size_t recb = sizeof(float) * (size_t)c.nt * p->hx * p->hy;
  if (recb && (rc = dmalloc((void **)&p->rec, recb, "record"))) return fail(rc);
  if (c.snapflag == 1) {
    const size_t sb = sizeof(float) * (size_t)p->hx * p->hy * p->hz;
    for (int k = 0; k < c.nt; k += c.ntsnap) {
      float *d = nullptr;
      if ((rc = dmalloc((void **)&d, sb, "snapshot"))) return fail(rc);
      p->snaps.push_back(d);
      p->snap_steps.push_back(k);
    }
  }
  *out = p;
  return QP3D_OK;
}

extern "C" int qp3d_plan_destroy(qp3d_plan *p)
{
  if (!p) return QP3D_OK;
  (void)hipSetDevice(p->cfg.device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  (void)hipFree(p->state);
  (void)hipFree(p->model);
  (void)hipFree(p->pmltab);
  (void)hipFree(p->gmax);
  (void)hipFree(p->flag);
  (void)hipFree(p->rec);
  (void)hipFree(p->sbox);
  for (float *d : p->snaps) (void)hipFree(d);
  for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return QP3D_OK;
}

extern "C" int qp3d_plan_upload(qp3d_plan *p, const float *vp, const float *eps, const float *del,
                                const float *source, const float *wavelet, const float *dlr, const float *ddlr,
                                const float *dtb, const float *ddtb, const float *dfb, const float *ddfb,
                                const float *c, const float *c2)
{
  if (!p || !vp || !eps || !del || !source || !wavelet || !dlr || !ddlr || !dtb || !ddtb || !dfb || !ddfb || !c ||
      !c2) {
    qp3d_set_error("plan_upload: null argument");
    return QP3D_ERR_ARG;
  }
  const qp3d_config &cf = p->cfg;
  HIP_TRY(hipSetDevice(cf.device));
  const long N = p->N;
  // model: (vp*dt)^2, epsilon, epsilon-delta
  std::vector<float> h(3 * (size_t)N);
  for (long i = 0; i < N; i++) {
    const double v = (double)vp[i] * cf.dt;
    h[i] = (float)(v * v);
    h[N + i] = eps[i];
    h[2 * N + i] = eps[i] - del[i];
  }
  HIP_TRY(hipMemcpy(p->model, h.data(), sizeof(float) * 3 * (size_t)N, hipMemcpyHostToDevice));
  // PML coefficient table per face family (x: dlr, y: dfb, z: dtb) and layer j:
  //   K1 = (1+d dt)^2 - 3, K2 = 2 d dt - 1, K3 = 1 - dt d, E = d'   (GPU_kernel.cu:568-576)
  std::vector<float4> tab(3 * (size_t)cf.pml);
  const float *dd[3] = {dlr, dfb, dtb}, *de[3] = {ddlr, ddfb, ddtb};
  for (int f = 0; f < 3; f++)
    for (int j = 0; j < cf.pml; j++) {
      const float d = dd[f][j], dt = cf.dt;
      float4 t;
      t.x = (float)(std::pow((double)(d * dt + 1), 2) - 3);
      t.y = 2 * d * dt - 1;
      t.z = 1 - dt * d;
      t.w = de[f][j];
      tab[(size_t)f * cf.pml + j] = t;
    }
  HIP_TRY(hipMemcpy(p->pmltab, tab.data(), sizeof(float4) * tab.size(), hipMemcpyHostToDevice));
  // FD coefficients (10th-order row of the 6x6 tables, CPU_function.c:123-127,143-147)
  Coef &k = p->coef;
  for (int m = 1; m <= H; m++) {
    k.cxx[m - 1] = c2[30 + m] / (cf.dx * cf.dx);
    k.cyy[m - 1] = c2[30 + m] / (cf.dy * cf.dy);
    k.czz[m - 1] = c2[30 + m] / (cf.dz * cf.dz);
    k.cx[m - 1] = c[30 + m] / cf.dx;
    k.cy[m - 1] = c[30 + m] / cf.dy;
    k.cz[m - 1] = c[30 + m] / cf.dz;
  }
  k.dt = cf.dt;
  // source: bounding box of the non-zero cells of the full-grid spatial source
  int z0 = cf.nz, z1 = -1, x0 = cf.nx, x1 = -1, y0 = cf.ny, y1 = -1;
  for (int z = 0; z < cf.nz; z++)
    for (int x = 0; x < cf.nx; x++) {
      const float *row = source + (long)z * cf.nx * cf.ny + (long)x * cf.ny;
      for (int y = 0; y < cf.ny; y++)
        if (row[y] != 0.0f) {
          z0 = std::min(z0, z); z1 = std::max(z1, z);
          x0 = std::min(x0, x); x1 = std::max(x1, x);
          y0 = std::min(y0, y); y1 = std::max(y1, y);
        }
    }
  (void)hipFree(p->sbox);
  p->sbox = nullptr;
  p->src_on = z1 >= 0;
  p->src_rim = false;
  if (p->src_on) {
    const int bz = z1 - z0 + 1, bx = x1 - x0 + 1, by = y1 - y0 + 1;
    std::vector<float> b((size_t)bz * bx * by);
    for (int z = 0; z < bz; z++)
      for (int x = 0; x < bx; x++)
        memcpy(&b[((size_t)z * bx + x) * by], source + (long)(z + z0) * cf.nx * cf.ny + (long)(x + x0) * cf.ny + y0,
               sizeof(float) * by);
    HIP_TRY(hipMalloc((void **)&p->sbox, sizeof(float) * b.size()));
    HIP_TRY(hipMemcpy(p->sbox, b.data(), sizeof(float) * b.size(), hipMemcpyHostToDevice));
    p->sb[0] = z0; p->sb[1] = z1; p->sb[2] = x0; p->sb[3] = x1; p->sb[4] = y0; p->sb[5] = y1;
    p->src_rim = z0 < H || x0 < H || y0 < H || z1 >= cf.nz - H || x1 >= cf.nx - H || y1 >= cf.ny - H;
  }
  p->wav.assign(wavelet, wavelet + cf.nt);
  p->uploaded = true;
  return QP3D_OK;
}

extern "C" int qp3d_plan_reset(qp3d_plan *p)
{
  if (!p) {
    qp3d_set_error("plan_reset: null plan");
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(p->cfg.device));
  HIP_TRY(hipMemsetAsync(p->state, 0, sizeof(float) * 9 * (size_t)p->N, p->stream));
  HIP_TRY(hipMemsetAsync(p->gmax, 0, sizeof(unsigned) * (size_t)(p->cfg.nt > 0 ? p->cfg.nt : 1), p->stream));
  HIP_TRY(hipMemsetAsync(p->flag, 0, sizeof(int), p->stream));
  if (p->rec) HIP_TRY(hipMemsetAsync(p->rec, 0, sizeof(float) * (size_t)p->cfg.nt * p->hx * p->hy, p->stream));
  p->nsnap_taken = 0;
  p->steps_done = 0;
  return QP3D_OK;
}

static int harvest_events(qp3d_plan *p)
{
  for (int k : p->ev_pending) {
    hipEvent_t *e = &p->ev[3 * (size_t)k];
    float a = 0, b = 0;
    HIP_TRY(hipEventElapsedTime(&a, e[0], e[1]));
    HIP_TRY(hipEventElapsedTime(&b, e[1], e[2]));
    p->t_grad += a;
    p->t_upd += b;
    p->n_launch += 1;
  }
  p->ev_pending.clear();
  return QP3D_OK;
}

static int launch_step(qp3d_plan *p, int k)
{
  const qp3d_config &cf = p->cfg;
  const int prev = k & 1, cur = prev ^ 1;  // u[prev] holds u(k-1) and receives u(k+1)
  hipEvent_t *e = nullptr;
  if (p->instrument) {
    if (p->ev.size() < 3 * (size_t)cf.nt) {
      size_t old = p->ev.size();
      p->ev.resize(3 * (size_t)cf.nt);
      for (size_t i = old; i < p->ev.size(); i++) HIP_TRY(hipEventCreate(&p->ev[i]));
    }
    e = &p->ev[3 * (size_t)k];
    HIP_TRY(hipEventRecord(e[0], p->stream));
  }
  GradArgs<StoreF32> ga;
  ga.g = p->geo;
  ga.c = p->coef;
  ga.ucur = p->u[cur];
  ga.gmax = p->gmax + k;
  ga.flag = p->flag;
  hipLaunchKernelGGL(k_gradmax<StoreF32>, dim3(p->nblk), dim3(TY, WX), 0, p->stream, ga);
  HIP_TRY(hipGetLastError());
  if (e) HIP_TRY(hipEventRecord(e[1], p->stream));

  UpdArgs<StoreF32> ua;
  ua.g = p->geo;
  ua.c = p->coef;
  ua.ucur = p->u[cur];
  ua.unext = p->u[prev];
  ua.vdt2 = p->vdt2; ua.eps = p->eps; ua.epsd = p->epsd;
  ua.w1p = p->w1[prev]; ua.w1c = p->w1[cur];
  ua.w3p = p->w3[prev]; ua.w3c = p->w3[cur];
  ua.pp = p->pp[prev]; ua.pc = p->pp[cur];
  ua.w21 = p->w21;
  ua.pmltab = p->pmltab;
  ua.gmax = p->gmax + k;
  ua.wk = p->wav[k];
  ua.src_on = p->src_on ? 1 : 0;
  ua.sb_z0 = p->sb[0]; ua.sb_z1 = p->sb[1]; ua.sb_x0 = p->sb[2]; ua.sb_x1 = p->sb[3];
  ua.sb_y0 = p->sb[4]; ua.sb_y1 = p->sb[5];
  ua.sbox = p->sbox;
  ua.rec = p->rec + (size_t)k * p->hx * p->hy;
  ua.sz = cf.sz;
  hipLaunchKernelGGL(k_update<StoreF32>, dim3(p->nblk), dim3(TY, WX), 0, p->stream, ua);
  HIP_TRY(hipGetLastError());
  if (e) {
    HIP_TRY(hipEventRecord(e[2], p->stream));
    p->ev_pending.push_back(k);
  }
  if (p->src_rim) {
    const long nb = (long)(p->sb[1] - p->sb[0] + 1) * (p->sb[3] - p->sb[2] + 1) * (p->sb[5] - p->sb[4] + 1);
    const int blocks = (int)std::min<long>((nb + 255) / 256, 4096);
    hipLaunchKernelGGL(k_rim_source<StoreF32>, dim3(blocks), dim3(256), 0, p->stream, (const float *)p->u[cur],
                       p->u[prev], p->geo, p->sb[0], p->sb[1], p->sb[2], p->sb[3], p->sb[4], p->sb[5],
                       (const float *)p->sbox, p->wav[k]);
    HIP_TRY(hipGetLastError());
  }
  if (cf.snapflag == 1 && k % cf.ntsnap == 0) {
    const int si = k / cf.ntsnap;
    if (si < (int)p->snaps.size()) {
      hipLaunchKernelGGL(k_snapshot<StoreF32>, dim3((p->hy + 255) / 256, p->hx, p->hz), dim3(256), 0, p->stream,
                         (const float *)p->u[prev], p->snaps[si], p->geo);
      HIP_TRY(hipGetLastError());
      p->nsnap_taken = std::max(p->nsnap_taken, si + 1);
    }
  }
  return QP3D_OK;
}

extern "C" int qp3d_plan_run(qp3d_plan *p, int k0, int k1)
{
  if (!p) {
    qp3d_set_error("plan_run: null plan");
    return QP3D_ERR_ARG;
  }
  if (!p->uploaded) {
    qp3d_set_error("plan_run: upload the model first");
    return QP3D_ERR_STATE;
  }
  if (k0 < 0 || k1 > p->cfg.nt || k0 > k1) {
    qp3d_set_error("plan_run: steps [%d,%d) outside [0,%d)", k0, k1, p->cfg.nt);
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(p->cfg.device));
  for (int k = k0; k < k1; k++) {
    int rc = launch_step(p, k);
    if (rc) return rc;
    p->steps_done = k + 1;
  }
  return QP3D_OK;
}

extern "C" int qp3d_plan_sync(qp3d_plan *p)
{
  if (!p) {
    qp3d_set_error("plan_sync: null plan");
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(p->cfg.device));
  HIP_TRY(hipStreamSynchronize(p->stream));
  int rc = harvest_events(p);
  if (rc) return rc;
  int flag = 0;
  HIP_TRY(hipMemcpy(&flag, p->flag, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) {
    qp3d_set_error("wavefield became non-finite (inf/NaN): dt violates the stability limit for this model");
    return QP3D_ERR_NONFINITE;
  }
  return QP3D_OK;
}

extern "C" int qp3d_plan_record(const qp3d_plan *p, float *record)
{
  if (!p || !record) {
    qp3d_set_error("plan_record: null argument");
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(p->cfg.device));
  HIP_TRY(hipStreamSynchronize(p->stream));
  if (p->rec)
    HIP_TRY(hipMemcpy(record, p->rec, sizeof(float) * (size_t)p->cfg.nt * p->hx * p->hy, hipMemcpyDeviceToHost));
  return QP3D_OK;
}

// After the step k the newest wavefield sits in slot k&1 (launch_step writes u(k+1) over u(k-1)).
extern "C" int qp3d_plan_field(const qp3d_plan *p, int which, float *out)
{
  if (!p || !out || which < 0 || which > 8) {
    qp3d_set_error("plan_field: bad argument");
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(p->cfg.device));
  HIP_TRY(hipStreamSynchronize(p->stream));
  const int newest = (p->steps_done - 1) & 1, older = newest ^ 1;
  const float *src = nullptr;
  switch (which) {
  case 0: src = p->u[newest]; break;
  case 1: src = p->u[older]; break;
  case 2: src = p->w1[newest]; break;
  case 3: src = p->w1[older]; break;
  case 4: src = p->w3[newest]; break;
  case 5: src = p->w3[older]; break;
  case 6: src = p->pp[newest]; break;
  case 7: src = p->pp[older]; break;
  case 8: src = p->w21; break;
  }
  HIP_TRY(hipMemcpy(out, src, sizeof(float) * (size_t)p->N, hipMemcpyDeviceToHost));
  return QP3D_OK;
}

extern "C" int qp3d_plan_snapshot_count(const qp3d_plan *p) { return p ? p->nsnap_taken : 0; }

extern "C" int qp3d_plan_snapshot(const qp3d_plan *p, int i, int *step, float *out)
{
  if (!p || i < 0 || i >= p->nsnap_taken || !out) {
    qp3d_set_error("plan_snapshot: index %d outside [0,%d)", i, p ? p->nsnap_taken : 0);
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(p->cfg.device));
  HIP_TRY(hipStreamSynchronize(p->stream));
  if (step) *step = p->snap_steps[i];
  HIP_TRY(hipMemcpy(out, p->snaps[i], sizeof(float) * (size_t)p->hx * p->hy * p->hz, hipMemcpyDeviceToHost));
  return QP3D_OK;
}

extern "C" int qp3d_plan_step_maxima(const qp3d_plan *p, float *out)
{
  if (!p || !out) {
    qp3d_set_error("plan_step_maxima: null argument");
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(p->cfg.device));
  HIP_TRY(hipStreamSynchronize(p->stream));
  HIP_TRY(hipMemcpy(out, p->gmax, sizeof(float) * (size_t)p->cfg.nt, hipMemcpyDeviceToHost));
  return QP3D_OK;
}

extern "C" int qp3d_plan_set_instrument(qp3d_plan *p, int on)
{
  if (!p) return QP3D_ERR_ARG;
  p->instrument = on != 0;
  p->t_grad = p->t_upd = 0;
  p->n_launch = 0;
  p->ev_pending.clear();
  return QP3D_OK;
}

extern "C" int qp3d_plan_kernel_times(qp3d_plan *p, double *ms_gradmax, double *ms_update, long *launches)
{
  if (!p) return QP3D_ERR_ARG;
  int rc = harvest_events(p);
  if (rc) return rc;
  const long n = p->n_launch;
  if (ms_gradmax) *ms_gradmax = n ? p->t_grad / n : 0;
  if (ms_update) *ms_update = n ? p->t_upd / n : 0;
  if (launches) *launches = n;
  return QP3D_OK;
}

extern "C" int qp3d_plan_step_bytes(const qp3d_plan *p, double *bytes_gradmax, double *bytes_update)
{
  if (!p) return QP3D_ERR_ARG;
  const qp3d_config &c = p->cfg;
  const double nupd = (double)(c.nx - 10) * (c.ny - 10) * (c.nz - 10);
  const double nint = (double)p->hx * p->hy * p->hz;
  const double npml = nupd - nint;
  // gradient-max: one read of u per updated cell. update: interior u_cur, u_prev, 3 model
  // fields, u_next (24 B); PML u_cur, 3 model fields, u_next, 7 state reads and 4 state writes (64 B).
  if (bytes_gradmax) *bytes_gradmax = 4.0 * nupd;
  if (bytes_update) *bytes_update = 24.0 * nint + 64.0 * npml;
  return QP3D_OK;
}

// ---------------------------------------------------------------------------------------
// drop-in entry points
static int write_floats(const char *name, const float *a, size_t n)
{
  FILE *fp = fopen(name, "wb");
  if (!fp) {
    qp3d_set_error("cannot open %s for writing", name);
    return QP3D_ERR_IO;
  }
  size_t w = fwrite(a, sizeof(float), n, fp);
  fclose(fp);
  if (w != n) {
    qp3d_set_error("short write to %s", name);
    return QP3D_ERR_IO;
  }
  return QP3D_OK;
}

// 3dPureqPModeling_single_shot.cu:340-377 / modeling3D.c:89-124: the interior cube and its
// three centre slices, xz at y'=(hy-1)/2, yz at x'=(hx-1)/2, xy at z'=(hz-1)/2.
static int write_snapshot_files(const char *prefix, int k, const float *snap, int hx, int hy, int hz)
{
  char name[4096];
  std::vector<float> xz((size_t)hz * hx), yz((size_t)hz * hy), xy((size_t)hx * hy);
  for (int z = 0; z < hz; z++)
    for (int x = 0; x < hx; x++) xz[(size_t)z * hx + x] = snap[((size_t)z * hx + x) * hy + (hy - 1) / 2];
  for (int z = 0; z < hz; z++)
    for (int y = 0; y < hy; y++) yz[(size_t)z * hy + y] = snap[((size_t)z * hx + (hx - 1) / 2) * hy + y];
  for (int x = 0; x < hx; x++)
    for (int y = 0; y < hy; y++) xy[(size_t)x * hy + y] = snap[((size_t)((hz - 1) / 2) * hx + x) * hy + y];
  int rc;
  snprintf(name, sizeof name, "%s%d.dat", prefix, k);
  if ((rc = write_floats(name, snap, (size_t)hx * hy * hz))) return rc;
  snprintf(name, sizeof name, "%s_xz%d.dat", prefix, k);
  if ((rc = write_floats(name, xz.data(), xz.size()))) return rc;
  snprintf(name, sizeof name, "%s_yz%d.dat", prefix, k);
  if ((rc = write_floats(name, yz.data(), yz.size()))) return rc;
  snprintf(name, sizeof name, "%s_xy%d.dat", prefix, k);
  if ((rc = write_floats(name, xy.data(), xy.size()))) return rc;
  return QP3D_OK;
}

extern "C" int qp3d_modeling3d(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt,
                               int pml, int snapflag, int sx, int sy, int sz, const float *vp, const float *epsilon,
                               const float *delta, const float *source, const float *wavelet, float *record,
                               const float *dlr, const float *ddlr, const float *dtb, const float *ddtb,
                               const float *dfb, const float *ddfb, const float *c, const float *c2,
                               const char *snap_file)
{
  (void)sx; (void)sy;
  if (!record) {
    qp3d_set_error("modeling3d: record is null");
    return QP3D_ERR_ARG;
  }
  if (snapflag == 1 && !snap_file) {
    qp3d_set_error("modeling3d: snapflag == 1 needs snap_file");
    return QP3D_ERR_ARG;
  }
  qp3d_config cfg{};
  cfg.nx = nx; cfg.ny = ny; cfg.nz = nz; cfg.pml = pml; cfg.nt = nt;
  cfg.dx = dx; cfg.dy = dy; cfg.dz = dz; cfg.dt = dt;
  cfg.sz = sz;
  cfg.storage = QP3D_STORAGE_FP32;
  cfg.snapflag = snapflag == 1 ? 1 : 0;
  cfg.ntsnap = ntsnap;
  cfg.device = 0;
  qp3d_plan *p = nullptr;
  int rc = qp3d_plan_create(&cfg, &p);
  if (rc) return rc;
  if ((rc = qp3d_plan_upload(p, vp, epsilon, delta, source, wavelet, dlr, ddlr, dtb, ddtb, dfb, ddfb, c, c2)) ||
      (rc = qp3d_plan_reset(p)) || (rc = qp3d_plan_run(p, 0, nt)) || (rc = qp3d_plan_sync(p)) ||
      (rc = qp3d_plan_record(p, record))) {
    qp3d_plan_destroy(p);
    return rc;
  }
  if (cfg.snapflag) {
    std::vector<float> snap((size_t)p->hx * p->hy * p->hz);
    for (int i = 0; i < qp3d_plan_snapshot_count(p); i++) {
      int k = 0;
      if ((rc = qp3d_plan_snapshot(p, i, &k, snap.data())) ||
          (rc = write_snapshot_files(snap_file, k, snap.data(), p->hx, p->hy, p->hz))) {
        qp3d_plan_destroy(p);
        return rc;
      }
    }
  }
  qp3d_plan_destroy(p);
  return QP3D_OK;
}

extern "C" void modeling3d(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt,
                           int pml, int snapflag, int sx, int sy, int sz, float *vp, float *epsilon, float *delta,
                           float *source, float *wavelet, float *record, float *dlr, float *ddlr, float *dtb,
                           float *ddtb, float *dfb, float *ddfb, float *c, float *c2, const char *snap_file)
{
  int rc = qp3d_modeling3d(nx, ny, nz, nt, ntsnap, dx, dy, dz, dt, pml, snapflag, sx, sy, sz, vp, epsilon, delta,
                           source, wavelet, record, dlr, ddlr, dtb, ddtb, dfb, ddfb, c, c2, snap_file);
  if (rc) fprintf(stderr, "modeling3d: %s\n", qp3d_last_error());
}
