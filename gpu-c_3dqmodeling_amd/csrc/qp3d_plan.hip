// Host side of the qP3D HIP library: plan lifecycle, uploads, work decomposition, the time
// loop, and the qp3d_plan_* entry points declared in include/qp3d.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qp3d_launch.hpp"
#include "qp3d_plan_impl.hpp"

// ---------------------------------------------------------------------------------------
// work decomposition

namespace {
struct Seg {
  int a, b;
};

// [a,b) cut into pieces of at most w.
void cut(std::vector<Seg> &out, int a, int b, int w)
{
  for (int s = a; s < b; s += w) out.push_back({s, std::min(s + w, b)});
}

// [a,b) cut at the multiples of w (pieces of at most w that start on a w-cell boundary, except
// the first), so that the lanes of a 64-wide tile start on a cache line (see Lane).
void cut_aligned(std::vector<Seg> &out, int a, int b, int w)
{
  for (int s = a; s < b; s = (s / w + 1) * w) out.push_back({s, std::min((s / w + 1) * w, b)});
}

int4 pack(int x0, int x1, int y0, int y1, int z0, int z1)
{
  return make_int4(x0 | (x1 << 16), y0 | (y1 << 16), z0, z1);
}

// z planes per item for `ntile` xy tiles over a z extent `len`: the count of items should fill
// whole rounds of `capacity` resident workgroups while each item stays long enough for the
// 10-plane queue priming to be small.
int choose_zchunk(int ntile, int len, int capacity, int forced)
{
  if (forced > 0) return std::min(forced, len);
  double best = -1;
  int bestc = len;
  for (int nc = 1; nc <= std::max(1, len / 16); nc++) {
    const int zc = (len + nc - 1) / nc;
    const long items = (long)ntile * ((len + zc - 1) / zc);
    const long rounds = (items + capacity - 1) / capacity;
    const double fill = (double)items / ((double)rounds * capacity);
    const double prime = (double)zc / (zc + 2 * H);
    const double score = fill * prime;
    if (score > best + 1e-9) {
      best = score;
      bestc = zc;
    }
  }
  return bestc;
}

// Order items so that the workgroups dispatched to one XCD (ids b with equal b % 8) take a
// contiguous range of the logical order, and pad to a multiple of 8 with empty items.
std::vector<int4> xcd_order(const std::vector<int4> &logical)
{
  const int n = (int)logical.size();
  const int np = (n + 7) / 8 * 8, per = np / 8;
  std::vector<int4> out(np, make_int4(0, 0, 0, 0));
  for (int b = 0; b < np; b++) {
    const int l = (b & 7) * per + (b >> 3);
    if (l < n) out[b] = logical[l];
  }
  return out;
}
}  // namespace

long qp3d_order_items(const std::vector<int4> &logical, bool batch, std::vector<int4> &ordered)
{
  ordered = xcd_order(logical);
  long cells = 0;
  for (const int4 &t : logical) {
    const long ny_ = (long)((((unsigned)t.y >> 16) & 0x7fff) - (t.y & 0xffff)) * (t.y < 0 ? 2 : 1);
    const long nz_ = batch ? (long)(((unsigned)t.z >> 16) - (t.z & 0xffff)) : (long)(t.w - t.z);
    cells += (long)(((unsigned)t.x >> 16) - (t.x & 0xffff)) * ny_ * nz_;
  }
  return cells;
}

int qp3d_upload_items(WorkList &w, const std::vector<int4> &logical, bool batch)
{
  std::vector<int4> v;
  w.cells = qp3d_order_items(logical, batch, v);
  w.n = (int)v.size();
  (void)hipFree(w.d);
  w.d = nullptr;
  if (w.n == 0) return QP3D_OK;
  HIP_TRY(hipMalloc((void **)&w.d, sizeof(int4) * v.size()));
  HIP_TRY(hipMemcpy(w.d, v.data(), sizeof(int4) * v.size(), hipMemcpyHostToDevice));
  return QP3D_OK;
}

static int upload_list(WorkList &w, const std::vector<int4> &logical) { return qp3d_upload_items(w, logical, false); }

int qp3d_cu_count(const qp3d_plan *p)
{
  int ncu = 256;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, p->cfg.device);
  return ncu;
}

// Uniform tiles over [5,n-5)^3 (the gradient-max pass); `capacity` is the number of
// workgroups resident at once that the items should fill.
void qp3d_items_uniform(const qp3d_plan *p, int tile, int capacity, std::vector<int4> &items, int zchunk)
{
  const qp3d_config &c = p->cfg;
  const int TX = tile_tx(tile);
  std::vector<Seg> xs, ys;
  cut(xs, H, c.nx - H, TX);
  cut_aligned(ys, H, c.ny - H, TY);
  const int len = c.nz - 2 * H;
  const int zc = choose_zchunk((int)(xs.size() * ys.size()), len, capacity, zchunk > 0 ? zchunk : c.zchunk);
  for (int z = H; z < c.nz - H; z += zc)
    for (const Seg &x : xs)
      for (const Seg &y : ys) items.push_back(pack(x.a, x.b, y.a, y.b, z, std::min(z + zc, c.nz - H)));
}

int qp3d_build_uniform_list(const qp3d_plan *p, int tile, int occ, WorkList &w)
{
  std::vector<int4> items;
  qp3d_items_uniform(p, tile, occ * qp3d_cu_count(p), items);
  return upload_list(w, items);
}

// Update tiles aligned to the PML boundary in x and y, z split at the PML planes: interior
// tiles (tile shape tile_i) hold only interior cells, PML tiles (tile_p) one face family each.
void qp3d_items_update(const qp3d_plan *p, int tile_i, int tile_p, int cap_i, int cap_p, std::vector<int4> &it,
                       std::vector<int4> &ip, std::vector<int4> &ipy, int tile_py, int cap_py)
{
  const qp3d_config &c = p->cfg;
  const int pml = c.pml;
  auto segs = [](std::vector<Seg> &out, std::vector<bool> &inner, int n, int pml, int w) {
    const int bounds[4] = {H, pml, n - pml, n - H};
    for (int s = 0; s < 3; s++) {
      const size_t before = out.size();
      cut(out, bounds[s], bounds[s + 1], w);
      inner.insert(inner.end(), out.size() - before, s == 1);
    }
  };
  // interior kernel: its own x tiling of the interior; PML kernel: x tiling with its tile width
  // (the y faces with their own when they go to a launch of their own)
  std::vector<Seg> xu, xs, xsy, ys, ysa;
  std::vector<bool> xui, xi, xiy, yi;
  segs(xu, xui, c.nx, pml, tile_tx(tile_i));
  segs(xs, xi, c.nx, pml, tile_tx(tile_p));
  segs(xsy, xiy, c.nx, pml, tile_tx(tile_py));
  segs(ys, yi, c.ny, pml, TY);
  // x-face tiles own every y (edges and corners included), so their y cuts need not follow the
  // y faces: 64-cell pieces on 64-cell boundaries
  cut_aligned(ysa, H, c.ny - H, TY);
  int nint_tiles = 0;
  for (size_t a = 0; a < xu.size(); a++)
    for (size_t b = 0; b < ys.size(); b++) nint_tiles += xui[a] && yi[b];
  // The two y faces of an x-interior tile column share one folded tile when each face fits
  // half a wavefront with its 5-cell halo.
  const bool fold = y_fold(pml);
  std::vector<std::pair<int4, bool>> ptiles, ytiles;  // PML xy tiles over the whole z range
  for (size_t a = 0; a < xs.size(); a++)
    if (!xi[a])
      for (size_t b = 0; b < ysa.size(); b++)
        ptiles.push_back({pack(xs[a].a, xs[a].b, ysa[b].a, ysa[b].b, 0, 0), false});
  for (size_t a = 0; a < xsy.size(); a++) {
    if (!xiy[a]) continue;
    if (fold) {
      ytiles.push_back({pack(xsy[a].a, xsy[a].b, H, pml, 0, 0), true});
    } else {
      for (size_t b = 0; b < ys.size(); b++)
        if (!yi[b]) ytiles.push_back({pack(xsy[a].a, xsy[a].b, ys[b].a, ys[b].b, 0, 0), false});
    }
  }
  const int zmid = c.nz - 2 * pml, zpml = pml - H;
  const int zc_i = choose_zchunk(nint_tiles, zmid, cap_i, c.zchunk);
  const int zc_p = choose_zchunk((int)ptiles.size(), c.nz - 2 * H, cap_p, c.zchunk);
  if (!ytiles.empty()) {
    const int zc_y = choose_zchunk((int)ytiles.size(), c.nz - 2 * H, cap_py, c.zchunk);
    for (int z = H; z < c.nz - H; z += zc_y)
      for (const auto &t : ytiles) {
        int4 w = t.first;
        w.z = z;
        w.w = std::min(z + zc_y, c.nz - H);
        if (t.second) w.y = (int)((unsigned)w.y | 0x80000000u);
        ipy.push_back(w);
      }
  }
  for (int z = pml; z < c.nz - pml; z += zc_i)
    for (size_t a = 0; a < xu.size(); a++)
      for (size_t b = 0; b < ys.size(); b++)
        if (xui[a] && yi[b]) it.push_back(pack(xu[a].a, xu[a].b, ys[b].a, ys[b].b, z, std::min(z + zc_i, c.nz - pml)));
  // PML tiles over the whole z range, then the interior tiles' top and bottom z faces
  for (int z = H; z < c.nz - H; z += zc_p)
    for (const auto &t : ptiles) {
      int4 w = t.first;
      w.z = z;
      w.w = std::min(z + zc_p, c.nz - H);
      if (t.second) w.y = (int)((unsigned)w.y | 0x80000000u);
      ip.push_back(w);
    }
  for (int face = 0; face < 2; face++) {
    const int za = face == 0 ? H : c.nz - pml, zb = face == 0 ? pml : c.nz - H;
    const int zc = std::max(1, std::min(zpml, zc_p));
    for (int z = za; z < zb; z += zc)
      for (size_t a = 0; a < xs.size(); a++)
        for (size_t b = 0; b < ys.size(); b++)
          if (xi[a] && yi[b]) ip.push_back(pack(xs[a].a, xs[a].b, ys[b].a, ys[b].b, z, std::min(z + zc, zb)));
  }
}

// Default gradient-pass tile (QP3D_TILES index): the 4x4 run tile, or 16x1 with BF16 storage
// (measured at 576^3: 0.39 ms FP16 and 0.25 ms FP32 with 4x4; BF16 0.45 ms with 16x1 against
// 0.54 ms with 4x4).
static int default_tile_grad(int storage) { return storage == QP3D_STORAGE_BF16 ? 1 : 6; }

void qp3d_plan_items(const qp3d_plan *p, std::vector<int4> (&items)[4])
{
  int occ_g = 1, occ_i = 1, occ_p = 1, occ_py = 1;
  with_storage(p->cfg.storage, [&](auto sp) {
    using S = decltype(sp);
    occ_g = occupancy_gradmax<S>(p->tile_g);
    occ_i = occupancy_update<S, 0>(p->tile_u, p->spec);
    occ_p = occupancy_update<S, 1>(p->tile_p, p->spec);
    occ_py = p->yw ? occupancy_update<S, 2>(p->tile_py, p->spec) : occupancy_update<S, 1>(p->tile_py, p->spec);
  });
  const int ncu = qp3d_cu_count(p);
  for (std::vector<int4> &v : items) v.clear();
  qp3d_items_uniform(p, p->tile_g, occ_g * ncu, items[0]);
  // The PML update is two launches: x and z faces, then the y faces (folded tiles fill half a
  // wavefront per face and run best with a narrower tile; measured at 576^3).
  qp3d_items_update(p, p->tile_u, p->tile_p, occ_i * ncu, occ_p * ncu, items[1], items[2], items[3], p->tile_py,
                    occ_py * ncu);
}

static int build_work(qp3d_plan *p)
{
  std::vector<int4> items[4];
  qp3d_plan_items(p, items);
  int rc;
  if ((rc = upload_list(p->wl_grad, items[0]))) return rc;
  if (p->ampg.b) {
    // the gradient pass under the amplitude bound: short items, so that a quiet stretch of a
    // tile column is skipped on its own
    std::vector<int4> ia;
    qp3d_items_uniform(p, p->tile_g, 1, ia, AMP_ZCHUNK);
    if ((rc = upload_list(p->wl_grad_amp, ia))) return rc;
  }
  if ((rc = upload_list(p->wl_int, items[1]))) return rc;
  if ((rc = upload_list(p->wl_pml, items[2]))) return rc;
  if ((rc = upload_list(p->wl_pmly, items[3]))) return rc;
  p->n_int_cells = p->wl_int.cells;
  p->n_pml_cells = p->wl_pml.cells + p->wl_pmly.cells;
  return QP3D_OK;
}

// ---------------------------------------------------------------------------------------
// lifecycle

// Validates a configuration (no device access); `who` prefixes the error message.
int qp3d_check_config(const qp3d_config &c, const char *who)
{
  if (c.pml < 6 || c.nx < 2 * c.pml + 1 || c.ny < 2 * c.pml + 1 || c.nz < 2 * c.pml + 1 || c.nt < 0) {
    qp3d_set_error("%s: need pml >= 6 and nx, ny, nz > 2*pml (got nx=%d ny=%d nz=%d pml=%d nt=%d)", who, c.nx,
                   c.ny, c.nz, c.pml, c.nt);
    return QP3D_ERR_ARG;
  }
  if (!(c.dx > 0 && c.dy > 0 && c.dz > 0 && c.dt > 0)) {
    qp3d_set_error("%s: spacings and dt must be positive", who);
    return QP3D_ERR_ARG;
  }
  if (c.sz < 0 || c.sz >= c.nz) {
    qp3d_set_error("%s: receiver plane sz=%d outside [0,%d)", who, c.sz, c.nz);
    return QP3D_ERR_ARG;
  }
  if (c.snapflag == 1 && c.ntsnap <= 0) {
    qp3d_set_error("%s: ntsnap must be positive when snapflag == 1", who);
    return QP3D_ERR_ARG;
  }
  if (c.snap_s != 0 && (c.snap_s != 1 || c.snapflag != 1)) {
    qp3d_set_error("%s: snap_s = 1 needs snapflag == 1 (got snap_s=%d snapflag=%d)", who, c.snap_s, c.snapflag);
    return QP3D_ERR_ARG;
  }
  if (c.tile_update < 0 || c.tile_update > NTILES || c.tile_grad < 0 || c.tile_grad > NTILES || c.tile_pml < 0 ||
      c.tile_pml > NTILES || c.tile_pml_y < 0 || c.tile_pml_y > NTILES || c.streams < 0 || c.streams > 2) {
    qp3d_set_error("%s: tile shapes must be in [0,%d] and streams in [0,2]", who, NTILES);
    return QP3D_ERR_ARG;
  }
  if (c.gmax_mode < 0 || c.gmax_mode > 2) {
    qp3d_set_error("%s: gmax_mode %d outside [0,2]", who, c.gmax_mode);
    return QP3D_ERR_ARG;
  }
  if (c.storage != QP3D_STORAGE_FP32 && c.storage != QP3D_STORAGE_BF16 && c.storage != QP3D_STORAGE_FP16) {
    qp3d_set_error("%s: unknown storage mode %d", who, c.storage);
    return QP3D_ERR_ARG;
  }
  const double cells = (double)c.nx * row_pitch(c.ny) * c.nz;  // rows padded to 32 cells
  if (cells >= 1073741824.0 || c.nx > 32767 || c.ny > 32767) {
    qp3d_set_error("%s: grid of %.4g cells (rows padded to 32 cells) exceeds the 2^30-cell limit of one plan", who,
                   cells);
    return QP3D_ERR_ARG;
  }
  return QP3D_OK;
}

// The host-side description of a plan for a valid configuration: geometry, tile shapes and
// modes (allocates nothing).
void qp3d_plan_describe(qp3d_plan *p, const qp3d_config &c)
{
  p->cfg = c;
  p->hx = c.nx - 2 * c.pml;
  p->hy = c.ny - 2 * c.pml;
  p->hz = c.nz - 2 * c.pml;
  p->geo.nx = c.nx; p->geo.ny = c.ny; p->geo.nz = c.nz; p->geo.pml = c.pml;
  p->geo.ys = row_pitch(c.ny);
  p->geo.pln = (unsigned)c.nx * (unsigned)p->geo.ys;
  p->N = (long)c.nz * c.nx * p->geo.ys;
  p->Nd = (long)c.nz * c.nx * c.ny;
  // defaults measured fastest on MI355X at 576^3: 8x2 interior, 4x4 gradient max (16x1 with
  // BF16 storage), 16x2 PML x and z faces, 16x1 PML y faces
  p->tile_u = c.tile_update > 0 ? c.tile_update - 1 : 0;
  p->tile_g = c.tile_grad > 0 ? c.tile_grad - 1 : default_tile_grad(c.storage);
  p->tile_p = c.tile_pml > 0 ? c.tile_pml - 1 : 5;
  p->tile_py = c.tile_pml_y > 0 ? c.tile_pml_y - 1 : 1;  // 16x1
  p->spec = c.gmax_mode == 2;
  // two streams by default; a step whose fix-up follows all updates runs on one stream; a
  // captured loop forks the PML updates onto the second stream and joins them back in the graph
  p->two_streams = (c.streams == 2 || c.streams == 0) && !p->spec;
  p->esz = c.storage == QP3D_STORAGE_FP32 ? 4 : 2;
  // folded y faces keep their model and PML state in compact arrays (YCompact)
  p->yw = y_fold(c.pml) ? c.pml - H : 0;
  p->ypln = (unsigned)(c.nx - 2 * c.pml) * (unsigned)(2 * p->yw);
  p->ycells = (long)c.nz * p->ypln;
  qp3d_state_layout(p);
}

static size_t ymodel_stride(const qp3d_plan *p) { return p->yw ? field_stride(sizeof(float) * (size_t)p->ycells) : 0; }

void qp3d_state_layout(qp3d_plan *p)
{
  p->fstride = field_stride(p->esz * (size_t)p->N);
  p->ystride = p->yw ? field_stride(p->esz * (size_t)p->ycells) : 0;
  p->state_bytes = 9 * p->fstride + 7 * p->ystride;
  p->model_bytes = 3 * field_stride(sizeof(float) * (size_t)p->N) + 3 * ymodel_stride(p);
}

// model == nullptr: the state only (the caller places the model arrays).
void qp3d_bind_state(qp3d_plan *p, char *s, float *model)
{
  const size_t fs = p->fstride, ys = p->ystride;
  p->state = s;
  p->u[0] = s; p->u[1] = s + fs;
  p->w1[0] = s + 2 * fs; p->w1[1] = s + 3 * fs;
  p->w3[0] = s + 4 * fs; p->w3[1] = s + 5 * fs;
  p->pp[0] = s + 6 * fs; p->pp[1] = s + 7 * fs;
  p->w21 = s + 8 * fs;
  if (p->yw) {
    char *y = s + 9 * fs;
    p->yw1[0] = y; p->yw1[1] = y + ys;
    p->yw3[0] = y + 2 * ys; p->yw3[1] = y + 3 * ys;
    p->ypp[0] = y + 4 * ys; p->ypp[1] = y + 5 * ys;
    p->yw21 = y + 6 * ys;
  }
  if (model) {
    const size_t ms = field_stride(sizeof(float) * (size_t)p->N) / sizeof(float);
    p->model = model;
    p->vdt2 = model; p->eps = model + ms; p->epsd = model + 2 * ms;
    if (p->yw) qp3d_bind_ymodel(p, model + 3 * ms);
  }
}

// The compact y-face model copies at ym (3 * ymodel_stride bytes).
void qp3d_bind_ymodel(qp3d_plan *p, float *ym)
{
  const size_t yms = ymodel_stride(p) / sizeof(float);
  p->yvdt2 = ym; p->yeps = ym + yms; p->yepsd = ym + 2 * yms;
}


extern "C" int qp3d_plan_create(const qp3d_config *cfg, qp3d_plan **out)
{
  if (!cfg || !out) {
    qp3d_set_error("plan_create: null argument");
    return QP3D_ERR_ARG;
  }
  *out = nullptr;
  const qp3d_config &c = *cfg;
  int rc0 = qp3d_check_config(c, "plan_create");
  if (rc0) return rc0;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (c.device < 0 || c.device >= ndev) {
    qp3d_set_error("plan_create: device %d not present (%d devices)", c.device, ndev);
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(c.device));

  qp3d_plan *p = new qp3d_plan();
  qp3d_plan_describe(p, c);

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
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    qp3d_set_error("plan_create: hipStreamCreate failed");
    return fail(QP3D_ERR_HIP);
  }
  if (p->two_streams) {
    // second stream (default priority) for the PML update; ev_grad orders it after the
    // gradient maximum, ev_pml orders the next step (and the end of every run) after it
    if (hipStreamCreateWithFlags(&p->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&p->ev_grad, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&p->ev_pml, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(p->ev_pml, p->stream2) != hipSuccess) {
      qp3d_set_error("plan_create: creating the second stream or its events failed");
      return fail(QP3D_ERR_HIP);
    }
  }
  int rc;
  char *state = nullptr;
  float *model = nullptr;
  if ((rc = dmalloc((void **)&state, p->state_bytes, "wavefields and PML state"))) return fail(rc);
  p->state = state;
  if ((rc = dmalloc((void **)&model, p->model_bytes, "model"))) return fail(rc);
  qp3d_bind_state(p, state, model);
  if (hipMemset(model, 0, p->model_bytes) != hipSuccess) {  // row padding of the model stays zero
    qp3d_set_error("plan_create: clearing the model failed");
    return fail(QP3D_ERR_HIP);
  }
  if ((rc = dmalloc((void **)&p->pmltab, sizeof(float4) * 3 * c.pml, "PML table"))) return fail(rc);
  if ((rc = dmalloc((void **)&p->gmax, sizeof(unsigned) * (size_t)(c.nt > 0 ? c.nt : 1), "step maxima")))
    return fail(rc);
  if ((rc = dmalloc((void **)&p->flag, sizeof(int), "flag"))) return fail(rc);
  if ((rc = dmalloc((void **)&p->kdev, sizeof(int), "step index"))) return fail(rc);
  if (!p->spec) {
    // amplitude bound of the gradient pass: three buffers of block maxima, two step keys
    AmpGrid &ag = p->ampg;
    ag.nbz = (c.nz + (1 << AB_Z) - 1) >> AB_Z;
    ag.nbx = (c.nx + (1 << AB_X) - 1) >> AB_X;
    ag.nby = (c.ny + (1 << AB_Y) - 1) >> AB_Y;
    ag.nblk = (unsigned)ag.nbz * ag.nbx * ag.nby;
    // three rotating slots each of max |u|, max |u_{k+1} - u_k|, max gradient component of u_k
    if ((rc = dmalloc((void **)&ag.b, sizeof(unsigned) * 9 * (size_t)ag.nblk, "block maxima"))) return fail(rc);
    if ((rc = dmalloc((void **)&ag.q, (size_t)ag.nblk, "quiet map"))) return fail(rc);
    // keys of the largest tile [2], two spare, bound-pass maxima (two words), the rest history
    // word; the counters of instrumented steps
    if ((rc = dmalloc((void **)&p->argkey, ARG_WORDS * sizeof(unsigned long long), "largest-tile keys")))
      return fail(rc);
    if ((rc = dmalloc((void **)&p->cnt, CNT_BYTES, "counters"))) return fail(rc);
    p->rflag = (unsigned *)(p->argkey + 5);
    if (hipMemset(ag.b, 0, sizeof(unsigned) * 9 * (size_t)ag.nblk) != hipSuccess ||
        hipMemset(ag.q, 0, (size_t)ag.nblk) != hipSuccess ||
        hipMemset(p->argkey, 0, ARG_WORDS * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(p->cnt, 0, CNT_BYTES) != hipSuccess || hipMemset(p->rflag, 1, sizeof(unsigned)) != hipSuccess) {
      qp3d_set_error("plan_create: clearing the block maxima failed");
      return fail(QP3D_ERR_HIP);
    }
  }
  if (p->spec) {
    p->fixwords = (p->N + 31) / 32;
    if ((rc = dmalloc((void **)&p->fixmask, sizeof(unsigned) * (size_t)p->fixwords, "cell mask"))) return fail(rc);
    if ((rc = dmalloc((void **)&p->gpred, sizeof(float) * (size_t)(c.nt > 0 ? c.nt : 1), "predicted maxima")))
      return fail(rc);
    if (hipMemset(p->fixmask, 0, sizeof(unsigned) * (size_t)p->fixwords) != hipSuccess ||
        hipMemset(p->gpred, 0, sizeof(float) * (size_t)(c.nt > 0 ? c.nt : 1)) != hipSuccess) {
      qp3d_set_error("plan_create: clearing the cell mask failed");
      return fail(QP3D_ERR_HIP);
    }
  }
  if ((rc = dmalloc((void **)&p->wav, sizeof(float) * (size_t)(c.nt > 0 ? c.nt : 1), "wavelet"))) return fail(rc);
  const size_t recb = sizeof(float) * (size_t)c.nt * p->hx * p->hy;
  if (recb && (rc = dmalloc((void **)&p->rec, recb, "record"))) return fail(rc);
  if (c.snapflag == 1) {
    const size_t sb = sizeof(float) * (size_t)p->hx * p->hy * p->hz;
    for (int k = 0; k < c.nt; k += c.ntsnap) {
      float *d = nullptr;
      if ((rc = dmalloc((void **)&d, sb, "snapshot"))) return fail(rc);
      p->snaps.push_back(d);
      p->snap_steps.push_back(k);
      if (c.snap_s) {
        float *ds = nullptr;
        if ((rc = dmalloc((void **)&ds, sizeof(float) * (size_t)p->Nd, "S-field snapshot"))) return fail(rc);
        p->snaps_s.push_back(ds);
      }
    }
  }
  if ((rc = build_work(p))) return fail(rc);
  *out = p;
  return QP3D_OK;
}

extern "C" int qp3d_plan_destroy(qp3d_plan *p)
{
  if (!p) return QP3D_OK;
  (void)hipSetDevice(p->cfg.device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->stream2) (void)hipStreamSynchronize(p->stream2);
  (void)hipFree(p->state);
  (void)hipFree(p->model);
  (void)hipFree(p->pmltab);
  (void)hipFree(p->gmax);
  (void)hipFree(p->flag);
  (void)hipFree(p->fixmask);
  (void)hipFree(p->gpred);
  (void)hipFree(p->kdev);
  (void)hipFree(p->wav);
  (void)hipFree(p->ampg.b);
  (void)hipFree(p->ampg.q);
  (void)hipFree(p->argkey);
  (void)hipFree(p->cnt);
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  (void)hipFree(p->rec);
  (void)hipFree(p->sbox);
  (void)hipFree(p->wl_grad.d);
  (void)hipFree(p->wl_grad_amp.d);
  (void)hipFree(p->wl_int.d);
  (void)hipFree(p->wl_pml.d);
  (void)hipFree(p->wl_pmly.d);
  for (float *d : p->snaps) (void)hipFree(d);
  for (float *d : p->snaps_s) (void)hipFree(d);
  for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
  if (p->ev_grad) (void)hipEventDestroy(p->ev_grad);
  if (p->ev_in) (void)hipEventDestroy(p->ev_in);
  if (p->ev_out) (void)hipEventDestroy(p->ev_out);
  if (p->ev_pml) (void)hipEventDestroy(p->ev_pml);
  if (p->stream2) (void)hipStreamDestroy(p->stream2);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return QP3D_OK;
}

// Whether ptr is device memory (arrays handed to the plan may live on the host or the device).
bool qp3d_on_device(const void *ptr)
{
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, ptr) != hipSuccess) {
    (void)hipGetLastError();  // memory the runtime does not know: plain host memory
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}
static bool on_device(const void *ptr) { return qp3d_on_device(ptr); }

// rows of ny elements from a src with row stride spitch to a dst with row stride dpitch
template <class E>
static __global__ void k_rows(E *dst, const E *src, long rows, int ny, int dpitch, int spitch)
{
  const long n = rows * ny;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long r = i / ny;
    const int y = (int)(i - r * ny);
    dst[r * dpitch + y] = src[r * spitch + y];
  }
}

template <class E>
static void launch_rows(void *dst, const void *src, long rows, int ny, int dpitch, int spitch, hipStream_t s)
{
  const long n = rows * ny;
  hipLaunchKernelGGL(k_rows<E>, dim3((unsigned)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, (E *)dst,
                     (const E *)src, rows, ny, dpitch, spitch);
}

int qp3d_copy_field(const qp3d_plan *p, void *dense, void *field, size_t esz, bool to_field, hipStream_t s)
{
  const int ny = p->cfg.ny, ys = p->geo.ys;
  const long rows = (long)p->cfg.nx * p->cfg.nz;
  const size_t nbytes = esz * (size_t)p->Nd;
  const bool dev = on_device(dense);
  if (ys == ny) {  // no row padding: one plain copy
    if (to_field)
      HIP_TRY(hipMemcpyAsync(field, dense, nbytes, hipMemcpyDefault, s));
    else
      HIP_TRY(hipMemcpyAsync(dense, field, nbytes, hipMemcpyDefault, s));
    if (!dev) HIP_TRY(hipStreamSynchronize(s));
    return QP3D_OK;
  }
  auto rows_copy = [&](void *dst, const void *src, int dp, int sp) {
    if (esz == 2) launch_rows<unsigned short>(dst, src, rows, ny, dp, sp, s);
    else if (esz == 8) launch_rows<double>(dst, src, rows, ny, dp, sp, s);
    else launch_rows<float>(dst, src, rows, ny, dp, sp, s);
  };
  if (dev) {
    if (to_field) rows_copy(field, dense, ys, ny);
    else rows_copy(dense, field, ny, ys);
    HIP_TRY(hipGetLastError());
    return QP3D_OK;
  }
  // host memory: through a dense device copy
  void *tmp = nullptr;
  hipError_t e = hipMalloc(&tmp, nbytes);
  if (e != hipSuccess) {
    qp3d_set_error("copying a field: hipMalloc of %zu bytes failed: %s", nbytes, hipGetErrorString(e));
    return QP3D_ERR_ALLOC;
  }
  if (to_field) {
    e = hipMemcpyAsync(tmp, dense, nbytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
      rows_copy(field, tmp, ys, ny);
      e = hipGetLastError();
    }
  } else {
    rows_copy(tmp, field, ny, ys);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(dense, tmp, nbytes, hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(tmp);
  if (e != hipSuccess) {
    qp3d_set_error("copying a field failed: %s", hipGetErrorString(e));
    return QP3D_ERR_HIP;
  }
  return QP3D_OK;
}

// In place over the three model arrays, holding vp, eps and del on entry:
// (vp dt)^2 in double rounded to float, eps, eps - del (the arithmetic of the host setup).
static __global__ void k_model(float *v2, const float *eps, float *ed, float dt, long n)
{
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double v = (double)v2[i] * dt;
    v2[i] = (float)(v * v);
    ed[i] = eps[i] - ed[i];
  }
}

// Bounding box {z0, z1, x0, x1, y0, y1} of the non-zero cells of a source array in device
// memory; box starts at {nz, -1, nx, -1, ny, -1}.
static __global__ void k_source_box(const float *src, int nx, int ny, int nz, int *box)
{
  const long pln = (long)nx * ny, n = pln * nz;
  int b[6] = {nz, -1, nx, -1, ny, -1};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    if (src[i] != 0.0f) {
      const int z = (int)(i / pln), x = (int)((i - z * pln) / ny), y = (int)(i - z * pln - (long)x * ny);
      b[0] = min(b[0], z); b[1] = max(b[1], z);
      b[2] = min(b[2], x); b[3] = max(b[3], x);
      b[4] = min(b[4], y); b[5] = max(b[5], y);
    }
  if (b[1] >= 0) {
    atomicMin(&box[0], b[0]); atomicMax(&box[1], b[1]);
    atomicMin(&box[2], b[2]); atomicMax(&box[3], b[3]);
    atomicMin(&box[4], b[4]); atomicMax(&box[5], b[5]);
  }
}

// Full-grid cell index of compact y-face cell i (YCompact layout; planes of ypln cells).
static __device__ __forceinline__ long ycompact_to_full(long i, Geo g, int yw, unsigned ypln)
{
  const int z = (int)(i / ypln);
  const int rem = (int)(i - (long)z * ypln);
  const int xr = rem / (2 * yw), col = rem - xr * (2 * yw);
  const int y = col < yw ? H + col : g.ny - g.pml + (col - yw);
  return (long)z * g.pln + (long)(xr + g.pml) * g.ys + y;
}

// Compact y-face copies of the three model arrays.
static __global__ void k_ymodel(const float *v2, const float *eps, const float *ed, float *yv2, float *yeps, float *yed,
                                Geo g, int yw, unsigned ypln, long n)
{
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long f = ycompact_to_full(i, g, yw, ypln);
    yv2[i] = v2[f];
    yeps[i] = eps[f];
    yed[i] = ed[f];
  }
}

// A compact y-face state field written into the y-face cells of its full-grid field.
template <class E>
static __global__ void k_yexpand(const E *cmp, E *full, Geo g, int yw, unsigned ypln, long n)
{
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    full[ycompact_to_full(i, g, yw, ypln)] = cmp[i];
}

static unsigned grid_blocks(long n) { return (unsigned)std::min<long>((n + 255) / 256, 8192); }

void qp3d_launch_ymodel(const qp3d_plan *p, hipStream_t s)
{
  if (!p->yw) return;
  hipLaunchKernelGGL(k_ymodel, dim3(grid_blocks(p->ycells)), dim3(256), 0, s, (const float *)p->vdt2,
                     (const float *)p->eps, (const float *)p->epsd, p->yvdt2, p->yeps, p->yepsd, p->geo, p->yw, p->ypln,
                     p->ycells);
}

static __global__ void k_scale(float *a, long n, float f)
{
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] *= f;
}

// n floats from host or device memory into a host vector.
static int to_host(const float *src, size_t n, std::vector<float> &v)
{
  v.resize(n);
  if (n) HIP_TRY(hipMemcpy(v.data(), src, sizeof(float) * n, hipMemcpyDefault));  // lands before return
  return QP3D_OK;
}

// PML coefficient table per face family (x: dlr, y: dfb, z: dtb) and layer j, and the FD
// coefficients of the plan (p->coef), from profiles and tables in host or device memory:
//   K1 = (1+d dt)^2 - 3, K2 = 2 d dt - 1, K3 = 1 - dt d, E = d'   (GPU_kernel.cu:568-576)
int qp3d_make_tables(qp3d_plan *p, const float *dlr, const float *ddlr, const float *dtb, const float *ddtb,
                     const float *dfb, const float *ddfb, const float *c, const float *c2, std::vector<float4> &tab)
{
  const qp3d_config &cf = p->cfg;
  std::vector<float> hdlr, hddlr, hdtb, hddtb, hdfb, hddfb, hc, hc2;
  int rc;
  if ((rc = to_host(dlr, (size_t)cf.pml, hdlr)) || (rc = to_host(ddlr, (size_t)cf.pml, hddlr)) ||
      (rc = to_host(dtb, (size_t)cf.pml, hdtb)) || (rc = to_host(ddtb, (size_t)cf.pml, hddtb)) ||
      (rc = to_host(dfb, (size_t)cf.pml, hdfb)) || (rc = to_host(ddfb, (size_t)cf.pml, hddfb)) ||
      (rc = to_host(c, 36, hc)) || (rc = to_host(c2, 36, hc2)))
    return rc;
  tab.resize(3 * (size_t)cf.pml);
  const float *dd[3] = {hdlr.data(), hdfb.data(), hdtb.data()}, *de[3] = {hddlr.data(), hddfb.data(), hddtb.data()};
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
  // FD coefficients: the 10th-order rows of the 6x6 tables (CPU_function.c:123-127,143-147)
  Coef &k = p->coef;
  for (int m = 1; m <= H; m++) {
    k.c1[m - 1] = hc[30 + m];
    k.c2[m - 1] = hc2[30 + m];
  }
  k.ix = 1.0f / cf.dx; k.iy = 1.0f / cf.dy; k.iz = 1.0f / cf.dz;
  k.ixx = (float)(1.0 / ((double)cf.dx * cf.dx));
  k.iyy = (float)(1.0 / ((double)cf.dy * cf.dy));
  k.izz = (float)(1.0 / ((double)cf.dz * cf.dz));
  for (int m = 0; m < H; m++) {
    k.g1x[m] = k.c1[m] * k.ix;
    k.g1y[m] = k.c1[m] * k.iy;
    k.g1z[m] = k.c1[m] * k.iz;
  }
  k.dt = cf.dt;
  return QP3D_OK;
}

void qp3d_launch_model(float *v2, const float *eps, float *ed, float dt, long n, hipStream_t s)
{
  hipLaunchKernelGGL(k_model, dim3((unsigned)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, v2, eps, ed, dt,
                     n);
}

void qp3d_launch_source_box(const float *src, int nx, int ny, int nz, int *box, hipStream_t s)
{
  const long n = (long)nx * ny * nz;
  hipLaunchKernelGGL(k_source_box, dim3((unsigned)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, src, nx, ny,
                     nz, box);
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
  HIP_TRY(hipStreamSynchronize(p->stream));
  const long N = p->N;
  // Every array may be host or device memory. Model: vp, eps, del copied into the plan's model
  // arrays (rows padded to the pitch; the padding stays zero), then (vp*dt)^2, epsilon,
  // epsilon-delta formed in place.
  // (copies on the plan's stream: a device-to-device hipMemcpy may return before it lands)
  int rc;
  if ((rc = qp3d_copy_field(p, const_cast<float *>(vp), p->vdt2, sizeof(float), true, p->stream)) ||
      (rc = qp3d_copy_field(p, const_cast<float *>(eps), p->eps, sizeof(float), true, p->stream)) ||
      (rc = qp3d_copy_field(p, const_cast<float *>(del), p->epsd, sizeof(float), true, p->stream)))
    return rc;
  qp3d_launch_model(p->vdt2, p->eps, p->epsd, cf.dt, N, p->stream);
  HIP_TRY(hipGetLastError());
  qp3d_launch_ymodel(p, p->stream);
  HIP_TRY(hipGetLastError());
  std::vector<float4> tab;
  if ((rc = qp3d_make_tables(p, dlr, ddlr, dtb, ddtb, dfb, ddfb, c, c2, tab))) return rc;
  HIP_TRY(hipMemcpyAsync(p->pmltab, tab.data(), sizeof(float4) * tab.size(), hipMemcpyHostToDevice, p->stream));
  std::vector<float> hw;
  if ((rc = to_host(wavelet, (size_t)cf.nt, hw))) return rc;
  wavelet = hw.data();
  // source: bounding box of the non-zero cells of the full-grid spatial source
  int z0 = cf.nz, z1 = -1, x0 = cf.nx, x1 = -1, y0 = cf.ny, y1 = -1;
  const bool src_dev = on_device(source);
  if (src_dev) {
    int box[6] = {cf.nz, -1, cf.nx, -1, cf.ny, -1}, *dbox = nullptr;
    HIP_TRY(hipMalloc((void **)&dbox, sizeof box));
    hipError_t e = hipMemcpyAsync(dbox, box, sizeof box, hipMemcpyHostToDevice, p->stream);
    if (e == hipSuccess) {
      qp3d_launch_source_box(source, cf.nx, cf.ny, cf.nz, dbox, p->stream);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e == hipSuccess) e = hipMemcpy(box, dbox, sizeof box, hipMemcpyDeviceToHost);
    (void)hipFree(dbox);
    if (e != hipSuccess) {
      qp3d_set_error("plan_upload: source box on the device failed: %s", hipGetErrorString(e));
      return QP3D_ERR_HIP;
    }
    z0 = box[0]; z1 = box[1]; x0 = box[2]; x1 = box[3]; y0 = box[4]; y1 = box[5];
  } else {
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
  }
  (void)hipFree(p->sbox);
  p->sbox = nullptr;
  p->src_on = z1 >= 0;
  p->src_rim = false;
  p->scale_log2 = 0;
  if (p->src_on) {
    const int bz = z1 - z0 + 1, bx = x1 - x0 + 1, by = y1 - y0 + 1;
    std::vector<float> b((size_t)bz * bx * by);
    for (int z = 0; z < bz; z++) {
      const float *s0 = source + (long)(z + z0) * cf.nx * cf.ny + (long)x0 * cf.ny + y0;
      if (src_dev)
        HIP_TRY(hipMemcpy2D(&b[(size_t)z * bx * by], sizeof(float) * by, s0, sizeof(float) * cf.ny, sizeof(float) * by, bx,
                            hipMemcpyDeviceToHost));
      else
        for (int x = 0; x < bx; x++) memcpy(&b[((size_t)z * bx + x) * by], s0 + (long)x * cf.ny, sizeof(float) * by);
    }
    if (cf.storage == QP3D_STORAGE_FP16) {
      // Half storage: scale the fields (through the source) so that the peak injected value
      // sits at 2^7. The wavefield peaks near 8x the peak of source*wavelet and the PML state
      // below that, which leaves 2^5 of headroom under the half range and keeps far-field
      // amplitudes out of the subnormals.
      double ms = 0, mw = 0;
      for (float v : b) ms = std::max(ms, (double)std::fabs(v));
      for (int k = 0; k < cf.nt; k++) mw = std::max(mw, (double)std::fabs(wavelet[k]));
      if (ms * mw > 0) p->scale_log2 = std::max(-100, std::min(100, (int)std::floor(std::log2(128.0 / (ms * mw)))));
      for (float &v : b) v = std::ldexp(v, p->scale_log2);
    }
    HIP_TRY(hipMalloc((void **)&p->sbox, sizeof(float) * b.size()));
    HIP_TRY(hipMemcpyAsync(p->sbox, b.data(), sizeof(float) * b.size(), hipMemcpyHostToDevice, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));  // b is about to go out of scope
    p->sb[0] = z0; p->sb[1] = z1; p->sb[2] = x0; p->sb[3] = x1; p->sb[4] = y0; p->sb[5] = y1;
    p->src_rim = z0 < H || x0 < H || y0 < H || z1 >= cf.nz - H || x1 >= cf.nx - H || y1 >= cf.ny - H;
  }
  if (cf.nt > 0) HIP_TRY(hipMemcpyAsync(p->wav, wavelet, sizeof(float) * cf.nt, hipMemcpyHostToDevice, p->stream));
  // the copies and the model kernel are on the plan's stream; done before the host vectors go
  HIP_TRY(hipStreamSynchronize(p->stream));
  // a captured loop holds the source box pointer and bounds: capture again on the next run
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->gexec = nullptr;
  p->graph = nullptr;
  p->uploaded = true;
  p->upload_gen++;
  return QP3D_OK;
}

extern "C" int qp3d_plan_reset(qp3d_plan *p)
{
  if (!p) {
    qp3d_set_error("plan_reset: null plan");
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(p->cfg.device));
  HIP_TRY(hipMemsetAsync(p->state, 0, p->state_bytes, p->stream));
  HIP_TRY(hipMemsetAsync(p->gmax, 0, sizeof(unsigned) * (size_t)(p->cfg.nt > 0 ? p->cfg.nt : 1), p->stream));
  HIP_TRY(hipMemsetAsync(p->flag, 0, sizeof(int), p->stream));
  if (p->fixmask) HIP_TRY(hipMemsetAsync(p->fixmask, 0, sizeof(unsigned) * (size_t)p->fixwords, p->stream));
  if (p->gpred)
    HIP_TRY(hipMemsetAsync(p->gpred, 0, sizeof(float) * (size_t)(p->cfg.nt > 0 ? p->cfg.nt : 1), p->stream));
  if (p->rec) HIP_TRY(hipMemsetAsync(p->rec, 0, sizeof(float) * (size_t)p->cfg.nt * p->hx * p->hy, p->stream));
  if (p->ampg.b) {
    HIP_TRY(hipMemsetAsync(p->ampg.b, 0, sizeof(unsigned) * 9 * (size_t)p->ampg.nblk, p->stream));
    HIP_TRY(hipMemsetAsync(p->argkey, 0, ARG_WORDS * sizeof(unsigned long long), p->stream));
    HIP_TRY(hipMemsetAsync(p->cnt, 0, CNT_BYTES, p->stream));
    // every block at rest, and a history from here
    HIP_TRY(hipMemsetAsync(p->ampg.q, 0, (size_t)p->ampg.nblk, p->stream));
    HIP_TRY(hipMemsetAsync(p->rflag, 1, sizeof(unsigned), p->stream));
  }
  p->amp_ok = true;  // u = 0 everywhere, and so are the block maxima
  p->nsnap_taken = 0;
  p->steps_done = 0;
  return QP3D_OK;
}

// ---------------------------------------------------------------------------------------
// time loop

static int harvest_events(qp3d_plan *p)
{
  for (int k : p->ev_pending) {
    hipEvent_t *e = &p->ev[EV_PER_STEP * (size_t)k];
    // e0..e1 gradient max; e1..e2 interior update; PML update ends at e3 and starts at e1
    // (concurrent, two streams) or e2 (one stream); with gmax_mode 2 there is no gradient pass
    // (e0 = e1) and e3..e4 is the marking pass and the fix-up. Mode 2 records only e1 and e2.
    float a = 0, b = 0, c = 0;
    HIP_TRY(hipEventElapsedTime(&b, e[1], e[2]));
    if (p->instrument == 1) {
      HIP_TRY(hipEventElapsedTime(&a, p->spec ? e[3] : e[0], p->spec ? e[4] : e[1]));
      HIP_TRY(hipEventElapsedTime(&c, p->two_streams ? e[1] : e[2], e[3]));
    }
    p->t_grad += a;
    p->t_int += b;
    if (p->t_int_step.size() < (size_t)p->cfg.nt) p->t_int_step.assign((size_t)p->cfg.nt, 0.f);
    p->t_int_step[(size_t)k] = b;
    p->t_pml += c;
    if (p->instrument == 1) p->t_upd += p->two_streams ? std::max(b, c) : b + c;
    p->n_launch += 1;
  }
  p->ev_pending.clear();
  return QP3D_OK;
}

// Enqueues time step k. `st` is the step index the kernels use (k itself, or a device index
// plus an offset when the launches are being captured; k then only sets the slot parity).
// A bare step (graph capture, recomputation) leaves out the snapshot and the instrumentation
// events. amp: the gradient pass skips tiles under the amplitude bound (when the block maxima
// cover the current u) and the updates refresh the block maxima; off for the steps that other
// objects drive (FWI recomputation, batches, the allocation-free entry point).
constexpr int GBOUND_BLOCKS = 64;  // workgroups of the bound pass (k_gbound)

template <class S>
static int launch_step_s(qp3d_plan *p, int k, Step st, bool bare, hipStream_t on, bool amp, bool gknown = false)
{
  // on: launch everything on this stream (and without the plan's second stream) instead of on
  // the plan's own streams
  hipStream_t const s0 = on ? on : p->stream;
  const bool two = p->two_streams && !on;  // never with gmax_mode 2
  using T = typename S::T;
  const qp3d_config &cf = p->cfg;
  const int prev = k & 1, cur = prev ^ 1;  // u[prev] holds u(k-1) and receives u(k+1)
  hipEvent_t *e = nullptr;
  if (p->instrument && !bare) {
    if (p->ev.size() < EV_PER_STEP * (size_t)cf.nt) {
      size_t old = p->ev.size();
      p->ev.resize(EV_PER_STEP * (size_t)cf.nt);
      for (size_t i = old; i < p->ev.size(); i++) HIP_TRY(hipEventCreate(&p->ev[i]));
    }
    e = &p->ev[EV_PER_STEP * (size_t)k];
  }
  // u(k) is complete only when the previous step's PML update has finished too (a graph's
  // first step: the launch of the graph waits for it instead)
  if (two && !p->cap_first) HIP_TRY(hipStreamWaitEvent(s0, p->ev_pml, 0));
  p->cap_first = false;
  const bool all = p->instrument == 1;  // 2: events around the interior update only
  if (e && all) HIP_TRY(hipEventRecord(e[0], s0));
  GradArgs<S> ga;
  UpdArgs<S> ua;
  make_args<S>(p, k & 1, st, ga, ua);
  const bool use_amp = amp && p->amp_on && p->ampg.b && !p->spec;
  if (!use_amp) {
    ga.amp = AmpGrid{};
    ga.argkey = nullptr;
    ga.gcount = nullptr;
    ua.amp = AmpGrid{};
  }
  if (!p->spec && !gknown) {
    int nwork = p->wl_grad.n;
    if (use_amp) {
      ga.work = p->wl_grad_amp.d;
      ga.nwork = nwork = p->wl_grad_amp.n;
      ga.skip = p->amp_ok ? 1 : 0;
      // tile counts only on fully instrumented steps
      if (p->instrument != 1 || bare) ga.gcount = nullptr;
      launch_gbound<S>(p->tile_g, GBOUND_BLOCKS, s0, ga);
      HIP_TRY(hipGetLastError());
    }
    launch_gradmax<S>(p->tile_g, nwork, s0, ga);
    HIP_TRY(hipGetLastError());
    if (use_amp && ga.skip) {
      // G_k is complete: mark the blocks where S = 1, and those at rest, for the updates
      const SrcBox sb{p->src_on ? 1 : 0, p->sb[0], p->sb[1], p->sb[2], p->sb[3], p->sb[4], p->sb[5]};
      hipLaunchKernelGGL(k_quiet, dim3((p->ampg.nblk + 255) / 256), dim3(256), 0, s0, p->ampg, p->geo, p->coef,
                         (const unsigned *)p->gmax, st, (const unsigned *)p->rflag, sb);
      HIP_TRY(hipGetLastError());
      ua.amp.q = p->ampg.q;
    }
  } else if (gknown && use_amp && p->amp_ok) {
    // G_k known from an earlier run of the same step (FWI recomputation from a checkpoint whose
    // block maxima and maps were restored): no gradient pass, the same maps
    const SrcBox sb{p->src_on ? 1 : 0, p->sb[0], p->sb[1], p->sb[2], p->sb[3], p->sb[4], p->sb[5]};
    const dim3 grid((p->ampg.nblk + 255) / 256);
    hipLaunchKernelGGL(k_amp_clear, grid, dim3(256), 0, s0, p->ampg, st);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_quiet, grid, dim3(256), 0, s0, p->ampg, p->geo, p->coef, (const unsigned *)p->gmax, st,
                       (const unsigned *)p->rflag, sb);
    HIP_TRY(hipGetLastError());
    ua.amp.q = p->ampg.q;
  }
  // a step without the map breaks the rest history until the next reset
  if (!ua.amp.q) ua.rest_off = p->rflag;
  if (ua.amp.q && p->instrument == 1 && !bare) ua.ucount = p->cnt + 2;  // all events: counters too
  p->amp_ok = use_amp;  // u(k+1) is covered by the block maxima the updates write
  if (e) HIP_TRY(hipEventRecord(e[1], s0));
  if (two) {
    HIP_TRY(hipEventRecord(p->ev_grad, s0));
    HIP_TRY(hipStreamWaitEvent(p->stream2, p->ev_grad, 0));
  }

  if (p->wl_int.n) {
    ua.work = p->wl_int.d;
    launch_update<S, 0>(p->tile_u, p->wl_int.n, s0, ua, p->spec);
    HIP_TRY(hipGetLastError());
  }
  if (e) HIP_TRY(hipEventRecord(e[2], s0));
  hipStream_t ps = two ? p->stream2 : s0;
  if (p->wl_pml.n) {
    ua.work = p->wl_pml.d;
    launch_update<S, 1>(p->tile_p, p->wl_pml.n, ps, ua, p->spec);
    HIP_TRY(hipGetLastError());
  }
  if (p->wl_pmly.n) {
    ua.work = p->wl_pmly.d;
    if (p->yw)
      launch_update<S, 2>(p->tile_py, p->wl_pmly.n, ps, ua, p->spec);
    else
      launch_update<S, 1>(p->tile_py, p->wl_pmly.n, ps, ua, p->spec);
    HIP_TRY(hipGetLastError());
  }
  if (e && all) HIP_TRY(hipEventRecord(e[3], ps));
  if (p->spec) {
    // after every update of the step (G_k complete): the marking pass (its workgroups exit at
    // once when the prediction held), then the fix-up of the marked cells
    launch_mark<S>(p->tile_g, p->wl_grad.n, s0, ga);
    HIP_TRY(hipGetLastError());
    const int blocks = (int)std::min<long>((p->fixwords / 4 + 256) / 256, 8192);
    hipLaunchKernelGGL(k_spec_fix<S>, dim3(blocks), dim3(256), 0, s0, ua, p->fixwords, cf.nt);
    HIP_TRY(hipGetLastError());
    if (e && all) HIP_TRY(hipEventRecord(e[4], s0));
  }
  if (e) p->ev_pending.push_back(k);
  // The rim source and the snapshot touch cells no PML tile writes (the outer 5 cells; the
  // interior cube), so they follow the interior update on the main stream.
  if (two) HIP_TRY(hipEventRecord(p->ev_pml, p->stream2));
  if (p->src_rim) {
    const long nb = (long)(p->sb[1] - p->sb[0] + 1) * (p->sb[3] - p->sb[2] + 1) * (p->sb[5] - p->sb[4] + 1);
    const int blocks = (int)std::min<long>((nb + 255) / 256, 4096);
    hipLaunchKernelGGL(k_rim_source<S>, dim3(blocks), dim3(256), 0, s0, (const T *)p->u[cur],
                       (T *)p->u[prev], p->geo, p->sb[0], p->sb[1], p->sb[2], p->sb[3], p->sb[4], p->sb[5],
                       (const float *)p->sbox, (const float *)p->wav, st, ua.amp);
    HIP_TRY(hipGetLastError());
  }
  if (!bare && cf.snapflag == 1 && k % cf.ntsnap == 0) {
    const int si = k / cf.ntsnap;
    if (si < (int)p->snaps.size()) {
      hipLaunchKernelGGL(k_snapshot<S>, dim3((p->hy + 255) / 256, p->hx, p->hz), dim3(256), 0, s0,
                         (const T *)p->u[prev], p->snaps[si], p->geo);
      HIP_TRY(hipGetLastError());
      if (si < (int)p->snaps_s.size()) {
        // S of this step: from u(k) (slot cur, which this step does not write) and G(k)
        const int blocks = (int)std::min<long>((p->Nd + 255) / 256, 8192);
        hipLaunchKernelGGL(k_sfield<S>, dim3(blocks), dim3(256), 0, s0, (const T *)p->u[cur],
                           (const float *)p->eps, (const float *)p->epsd, (const unsigned *)p->gmax, st, p->geo,
                           p->coef, p->snaps_s[si]);
        HIP_TRY(hipGetLastError());
      }
      p->nsnap_taken = std::max(p->nsnap_taken, si + 1);
    }
  }
  return QP3D_OK;
}

static int launch_step(qp3d_plan *p, int k, Step st, bool bare, hipStream_t on, bool amp, bool gknown = false)
{
  int rc = QP3D_OK;
  with_storage(p->cfg.storage,
               [&](auto sp) { rc = launch_step_s<decltype(sp)>(p, k, st, bare, on, amp, gknown); });
  return rc;
}

// FWI: the first forward run with the amplitude bound, or a recomputed step whose G_k the first
// run already left in gmax[k] (no gradient pass; the update reads it)
// (on the plan's own stream only: the checkpoints and the adjoint follow on it). maps: a
// recomputed step whose checkpoint restored the block maxima, maps and rest history of the
// first run uses them as that run did.
int qp3d_forward_step_fwi(qp3d_plan *p, int k, bool recompute, bool maps)
{
  if (recompute && p->spec) return launch_step(p, k, Step{nullptr, k}, true, p->stream, false);
  return recompute ? launch_step(p, k, Step{nullptr, k}, true, p->stream, maps, true)
                   : launch_step(p, k, Step{nullptr, k}, true, p->stream, true);
}

int qp3d_forward_step(qp3d_plan *p, int k) { return launch_step(p, k, Step{nullptr, k}, true, nullptr, false); }

int qp3d_forward_step_on(qp3d_plan *p, int k, Step st, hipStream_t on)
{
  return launch_step(p, k, st, true, on, false);
}

static bool snapshot_in(const qp3d_plan *p, int k0, int k1)
{
  if (p->cfg.snapflag != 1) return false;
  const int first = (k0 + p->cfg.ntsnap - 1) / p->cfg.ntsnap * p->cfg.ntsnap;  // first multiple >= k0
  return first < k1 && first / p->cfg.ntsnap < (int)p->snaps.size();
}

// Captures GRAPH_LEN steps (even start) into p->gexec; kernels read the step from p->kdev.
static int capture_steps(qp3d_plan *p)
{
  HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  int rc = QP3D_OK;
  // every replay starts with the block maxima covering u (qp3d_plan_run checks amp_ok)
  p->amp_ok = true;
  // two streams: each step forks the PML updates onto the second stream after the gradient
  // pass (ev_grad) and the next step joins them (ev_pml), both recorded inside the capture;
  // the last step's PML updates join the capturing stream before the capture ends
  p->cap_first = true;
  for (int i = 0; i < GRAPH_LEN && rc == QP3D_OK; i++) rc = launch_step(p, i, Step{p->kdev, i}, true, nullptr, true);
  p->cap_first = false;
  if (rc == QP3D_OK && p->two_streams) {
    const hipError_t ej = hipStreamWaitEvent(p->stream, p->ev_pml, 0);
    if (ej != hipSuccess) {
      qp3d_set_error("plan_run: joining the second stream in the capture failed: %s", hipGetErrorString(ej));
      rc = QP3D_ERR_HIP;
    }
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(p->stream, &g);
  if (rc == QP3D_OK && e != hipSuccess) {
    qp3d_set_error("plan_run: graph capture failed: %s", hipGetErrorString(e));
    rc = QP3D_ERR_HIP;
  }
  if (rc == QP3D_OK) {
    e = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      qp3d_set_error("plan_run: graph instantiation failed: %s", hipGetErrorString(e));
      p->gexec = nullptr;
      rc = QP3D_ERR_HIP;
    }
  }
  if (rc == QP3D_OK)
    p->graph = g;
  else if (g)
    (void)hipGraphDestroy(g);
  return rc;
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
  const bool graph = p->cfg.use_graph && !p->instrument;
  for (int k = k0; k < k1;) {
    // a replay assumes the block maxima cover u (a direct step renews them first when not)
    if (graph && (k & 1) == 0 && k + GRAPH_LEN <= k1 && !snapshot_in(p, k, k + GRAPH_LEN) &&
        (p->amp_ok || !p->ampg.b)) {
      if (!p->gexec) {
        int rc = capture_steps(p);
        if (rc) return rc;
      }
      hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, p->stream, p->kdev, k);
      HIP_TRY(hipGetLastError());
      // the replay's first step reads u(k), which the last direct step's PML update finishes
      if (p->two_streams) HIP_TRY(hipStreamWaitEvent(p->stream, p->ev_pml, 0));
      HIP_TRY(hipGraphLaunch(p->gexec, p->stream));
      k += GRAPH_LEN;
    } else {
      int rc = launch_step(p, k, Step{nullptr, k}, false, nullptr, true);
      if (rc) return rc;
      k++;
    }
    p->steps_done = k;
  }
  // everything enqueued so far is ordered before later work on the plan's stream
  if (p->two_streams) HIP_TRY(hipStreamWaitEvent(p->stream, p->ev_pml, 0));
  return QP3D_OK;
}

// qp3d_plan_run ordered with a caller's stream: the steps start after the work queued on
// `stream` so far, and work queued on `stream` afterwards waits for them (events, no host wait).
extern "C" int qp3d_plan_run_on(qp3d_plan *p, int k0, int k1, void *stream)
{
  if (!p) {
    qp3d_set_error("plan_run_on: null plan");
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(p->cfg.device));
  if (!p->ev_in) {
    HIP_TRY(hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming));
  }
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipEventRecord(p->ev_in, s));
  HIP_TRY(hipStreamWaitEvent(p->stream, p->ev_in, 0));
  const int rc = qp3d_plan_run(p, k0, k1);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(p->ev_out, p->stream));
  HIP_TRY(hipStreamWaitEvent(s, p->ev_out, 0));
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

// ---------------------------------------------------------------------------------------
// results

// Takes the half-storage scale out of values read back from the device.
static void unscale(const qp3d_plan *p, float *a, size_t n)
{
  if (!p->scale_log2) return;
  const float f = std::ldexp(1.0f, -p->scale_log2);
  for (size_t i = 0; i < n; i++) a[i] *= f;
}

// n floats of plan data (device) into out, host or device memory, with the FP16 storage scale
// taken out.
static int copy_out(const qp3d_plan *p, float *out, const float *src, size_t n)
{
  if (!on_device(out)) {
    HIP_TRY(hipMemcpy(out, src, sizeof(float) * n, hipMemcpyDeviceToHost));
    unscale(p, out, n);
    return QP3D_OK;
  }
  // on the plan's stream, so that the scaling follows the copy, and complete on return
  HIP_TRY(hipMemcpyAsync(out, src, sizeof(float) * n, hipMemcpyDeviceToDevice, p->stream));
  if (p->scale_log2) {
    hipLaunchKernelGGL(k_scale, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)), dim3(256), 0, p->stream, out,
                       (long)n, std::ldexp(1.0f, -p->scale_log2));
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(p->stream));
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
  const size_t n = (size_t)p->cfg.nt * p->hx * p->hy;
  return p->rec ? copy_out(p, record, p->rec, n) : QP3D_OK;
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
  void *src = nullptr, *ysrc = nullptr;
  switch (which) {
  case 0: src = p->u[newest]; break;
  case 1: src = p->u[older]; break;
  case 2: src = p->w1[newest]; ysrc = p->yw1[newest]; break;
  case 3: src = p->w1[older]; ysrc = p->yw1[older]; break;
  case 4: src = p->w3[newest]; ysrc = p->yw3[newest]; break;
  case 5: src = p->w3[older]; ysrc = p->yw3[older]; break;
  case 6: src = p->pp[newest]; ysrc = p->ypp[newest]; break;
  case 7: src = p->pp[older]; ysrc = p->ypp[older]; break;
  case 8: src = p->w21; ysrc = p->yw21; break;
  }
  if (ysrc && p->yw) {
    // the y-face cells of a PML state field live in its compact array; the same cells of the
    // full-grid field are otherwise unused, so the compact values are written there first
    if (p->esz == 4)
      hipLaunchKernelGGL(k_yexpand<float>, dim3(grid_blocks(p->ycells)), dim3(256), 0, p->stream,
                         (const float *)ysrc, (float *)src, p->geo, p->yw, p->ypln, p->ycells);
    else
      hipLaunchKernelGGL(k_yexpand<unsigned short>, dim3(grid_blocks(p->ycells)), dim3(256), 0, p->stream,
                         (const unsigned short *)ysrc, (unsigned short *)src, p->geo, p->yw, p->ypln, p->ycells);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(p->stream));
  }
  const size_t nd = (size_t)p->Nd;
  int rc;
  if (p->esz == 4) {
    // dense copy on the plan's stream, then the FP16 scale taken out where the copy landed
    if ((rc = qp3d_copy_field(p, out, src, sizeof(float), false, p->stream))) return rc;
    if (!on_device(out)) {
      unscale(p, out, nd);
      return QP3D_OK;
    }
    if (p->scale_log2) {
      hipLaunchKernelGGL(k_scale, dim3((unsigned)std::min<size_t>((nd + 255) / 256, 8192)), dim3(256), 0, p->stream,
                         out, (long)nd, std::ldexp(1.0f, -p->scale_log2));
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(p->stream));
    return QP3D_OK;
  }
  const bool dev = on_device(out);
  std::vector<float> hout(dev ? nd : 0);
  float *const dst = out;
  if (dev) out = hout.data();
  {
    std::vector<unsigned short> h(nd);
    if ((rc = qp3d_copy_field(p, h.data(), src, sizeof(unsigned short), false, p->stream))) return rc;
    if (p->cfg.storage == QP3D_STORAGE_BF16)
      for (size_t i = 0; i < nd; i++) {
        const unsigned bits = (unsigned)h[i] << 16;
        memcpy(&out[i], &bits, 4);
      }
    else
      for (size_t i = 0; i < nd; i++) {
        _Float16 v;
        memcpy(&v, &h[i], 2);
        out[i] = (float)v;
      }
  }
  unscale(p, out, nd);
  if (dev) HIP_TRY(hipMemcpy(dst, out, sizeof(float) * nd, hipMemcpyHostToDevice));
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
  return copy_out(p, out, p->snaps[i], (size_t)p->hx * p->hy * p->hz);
}

extern "C" int qp3d_plan_snapshot_s(const qp3d_plan *p, int i, int *step, float *out)
{
  if (!p || i < 0 || i >= p->nsnap_taken || i >= (int)p->snaps_s.size() || !out) {
    qp3d_set_error("plan_snapshot_s: no S-field snapshot %d (snapshots taken %d, S fields kept: %s)", i,
                   p ? p->nsnap_taken : 0, p && !p->snaps_s.empty() ? "yes" : "no");
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(p->cfg.device));
  HIP_TRY(hipStreamSynchronize(p->stream));
  if (step) *step = p->snap_steps[i];
  HIP_TRY(hipMemcpyAsync(out, p->snaps_s[i], sizeof(float) * (size_t)p->Nd, hipMemcpyDefault, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  return QP3D_OK;
}

extern "C" int qp3d_work_items(const qp3d_config *cfg, int which, int capacity, int *out, int cap)
{
  if (!cfg || which < 0 || which > 3 || capacity < 1 || cap < 0 || (cap > 0 && !out)) {
    qp3d_set_error("work_items: bad argument");
    return QP3D_ERR_ARG;
  }
  const qp3d_config &c = *cfg;
  if (c.pml < 6 || c.nx < 2 * c.pml + 1 || c.ny < 2 * c.pml + 1 || c.nz < 2 * c.pml + 1 || c.tile_update < 0 ||
      c.tile_update > NTILES || c.tile_grad < 0 || c.tile_grad > NTILES || c.tile_pml < 0 || c.tile_pml > NTILES ||
      c.tile_pml_y < 0 || c.tile_pml_y > NTILES) {
    qp3d_set_error("work_items: bad configuration");
    return QP3D_ERR_ARG;
  }
  qp3d_plan p;  // host-side description only: nothing is allocated
  qp3d_plan_describe(&p, c);
  std::vector<int4> it, ip, ipy;
  if (which == 0) {
    qp3d_items_uniform(&p, p.tile_g, capacity, it);
  } else {
    qp3d_items_update(&p, p.tile_u, p.tile_p, capacity, capacity, it, ip, ipy, p.tile_py, capacity);
  }
  const std::vector<int4> &v = which == 3 ? ipy : which == 2 ? ip : it;
  for (size_t i = 0; i < v.size() && (int)i < cap; i++) {
    out[4 * i] = v[i].x;
    out[4 * i + 1] = v[i].y;
    out[4 * i + 2] = v[i].z;
    out[4 * i + 3] = v[i].w;
  }
  return (int)v.size();
}

extern "C" int qp3d_plan_dims(const qp3d_plan *p, int *hx, int *hy, int *hz)
{
  if (!p) return QP3D_ERR_ARG;
  if (hx) *hx = p->hx;
  if (hy) *hy = p->hy;
  if (hz) *hz = p->hz;
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
  return copy_out(p, out, (const float *)p->gmax, (size_t)p->cfg.nt);
}

extern "C" int qp3d_plan_step_predictions(const qp3d_plan *p, float *out)
{
  if (!p || !out) {
    qp3d_set_error("plan_step_predictions: null argument");
    return QP3D_ERR_ARG;
  }
  if (!p->gpred) {
    qp3d_set_error("plan_step_predictions: the plan was created with gmax_mode != 2");
    return QP3D_ERR_STATE;
  }
  HIP_TRY(hipSetDevice(p->cfg.device));
  return copy_out(p, out, (const float *)p->gpred, (size_t)p->cfg.nt);
}

extern "C" int qp3d_plan_gmax_mode(const qp3d_plan *p)
{
  if (!p) {
    qp3d_set_error("plan_gmax_mode: null plan");
    return QP3D_ERR_ARG;
  }
  return p->spec ? 2 : 1;
}

// Amplitude bound of the gradient pass on (the default) or off; results are the same either way.
extern "C" int qp3d_plan_set_amp_bound(qp3d_plan *p, int on)
{
  if (!p || on < 0 || on > 1) {
    qp3d_set_error("plan_set_amp_bound: bad arguments");
    return QP3D_ERR_ARG;
  }
  if (p->amp_on != (on == 1)) {
    HIP_TRY(hipSetDevice(p->cfg.device));
    HIP_TRY(hipStreamSynchronize(p->stream));
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->gexec = nullptr;
    p->graph = nullptr;
    p->amp_on = on == 1;
    p->amp_ok = false;  // the block maxima were not kept up while off
  }
  return QP3D_OK;
}

// Gradient-pass tiles evaluated and skipped under the amplitude bound since the last reset
// (out[0], out[1]), counted on instrumented steps (qp3d_plan_set_instrument); both 0 when the
// bound is unavailable (gmax_mode 2).
// The counter slots after the plan's work so far (zeros when the plan keeps none).
static int read_counters(const qp3d_plan *p, unsigned long long *c)
{
  std::memset(c, 0, CNT_BYTES);
  if (!p->cnt) return QP3D_OK;
  HIP_TRY(hipSetDevice(p->cfg.device));
  HIP_TRY(hipStreamSynchronize(p->stream));
  if (p->two_streams) HIP_TRY(hipStreamSynchronize(p->stream2));
  HIP_TRY(hipMemcpy(c, p->cnt, CNT_BYTES, hipMemcpyDeviceToHost));
  return QP3D_OK;
}

extern "C" int qp3d_plan_gradient_tiles(const qp3d_plan *p, unsigned long long *out)
{
  if (!p || !out) return QP3D_ERR_ARG;
  out[0] = out[1] = 0;
  unsigned long long c[CNT_SLOTS * CNT_STRIDE];
  int rc = read_counters(p, c);
  if (rc) return rc;
  for (int s = 0; s < CNT_SLOTS; s++)
    for (int j = 0; j < 2; j++) out[j] += c[s * CNT_STRIDE + j];
  return QP3D_OK;
}

extern "C" int qp3d_plan_update_cells(const qp3d_plan *p, unsigned long long *out)
{
  if (!p || !out) return QP3D_ERR_ARG;
  for (int i = 0; i < 9; i++) out[i] = 0;
  unsigned long long c[CNT_SLOTS * CNT_STRIDE];
  int rc = read_counters(p, c);
  if (rc) return rc;
  for (int s = 0; s < CNT_SLOTS; s++)
    for (int j = 0; j < 9; j++) out[j] += c[s * CNT_STRIDE + 2 + j];
  return QP3D_OK;
}

extern "C" int qp3d_plan_set_instrument(qp3d_plan *p, int on)
{
  if (!p) return QP3D_ERR_ARG;
  if (on < 0 || on > 2) {
    qp3d_set_error("plan_set_instrument: mode %d outside [0,2]", on);
    return QP3D_ERR_ARG;
  }
  p->instrument = on;
  p->t_grad = p->t_int = p->t_pml = p->t_upd = 0;
  p->n_launch = 0;
  p->ev_pending.clear();
  return QP3D_OK;
}

extern "C" int qp3d_plan_step_times(const qp3d_plan *p, float *ms)
{
  if (!p || !ms) return QP3D_ERR_ARG;
  for (int k = 0; k < p->cfg.nt; k++) ms[k] = (size_t)k < p->t_int_step.size() ? p->t_int_step[(size_t)k] : 0.f;
  return QP3D_OK;
}

extern "C" int qp3d_plan_kernel_times(qp3d_plan *p, double *ms)
{
  if (!p || !ms) return QP3D_ERR_ARG;
  int rc = harvest_events(p);
  if (rc) return rc;
  const long n = p->n_launch;
  ms[0] = n ? p->t_grad / n : 0;
  ms[1] = n ? p->t_int / n : 0;
  ms[2] = n ? p->t_pml / n : 0;
  ms[3] = n ? p->t_upd / n : 0;
  ms[4] = (double)n;
  return QP3D_OK;
}

extern "C" int qp3d_plan_step_bytes(const qp3d_plan *p, double *bytes)
{
  if (!p || !bytes) return QP3D_ERR_ARG;
  const qp3d_config &c = p->cfg;
  const double nupd = (double)(c.nx - 10) * (c.ny - 10) * (c.nz - 10);
  const double nint = (double)p->hx * p->hy * p->hz;
  const double npml = nupd - nint;
  // gradient pass: one read of u per updated cell. Interior update: u_cur, u_prev, 3 model
  // fields, u_next (24 B). PML update: u_cur, 3 model fields, u_next, 7 state reads and 4 state
  // writes (64 B); the PML launch also carries the interior cells of mixed tiles (24 B).
  // Stored fields take esz bytes (4 float, 2 bfloat16/half); the model stays float.
  const double pml_launch_int = (double)p->n_pml_cells - npml;
  const double e = (double)p->esz, bi = 12.0 + 3.0 * e, bp = 12.0 + 13.0 * e;
  // gmax_mode 2 has no gradient pass: its fix-up reads the cell mask (one bit per cell) and
  // touches only the marked cells
  bytes[0] = p->spec ? 4.0 * (double)p->fixwords : e * nupd;
  bytes[1] = bi * (double)p->n_int_cells;
  bytes[2] = bp * npml + bi * pml_launch_int;
  bytes[3] = (double)p->n_int_cells;
  bytes[4] = (double)p->n_pml_cells;
  return QP3D_OK;
}
