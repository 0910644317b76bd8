// Many shots in one launch per stage: the gradient-max pass, the interior update and the PML
// update (x and z faces, y faces) of every shot of a batch go out as four launches per time
// step whose work items carry the shot's index. Each shot keeps its own plan (grid, model, source, PML profiles,
// wavefields, record); all shots advance in lockstep. The step sequence can be captured once
// in a hipGraph, and an autotuner picks each stage's tile shape by timing the candidates.
// Persistent mode: every step of a run of every shot in one launch (k_persist, qp3d_persist.hpp),
// the shots advancing independently through per-shot item counters.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdio>
#include <vector>

#include "qp3d_launch.hpp"
#include "qp3d_persist.hpp"
#include "qp3d_plan_impl.hpp"

struct qp3d_batch {
  std::vector<qp3d_plan *> plans;
  std::vector<unsigned> gen;  // each plan's upload count when the batch was built
  int nt = 0, storage = 0, device = 0;
  bool use_graph = false;
  hipStream_t stream = nullptr;
  // gradient max, interior update, PML update on x and z faces, on y faces (QP3D_TILES index)
  int tile[4] = {6, 0, 5, 1};
  // each shot's own launches (its plan's work lists and tiles) one shot after another within a
  // step, instead of one launch per stage for all shots; the autotuner sets it when faster
  bool per_shot = false;
  // work items of the shared launches: each shot cut as its own plan cuts it, for the whole
  // machine (false), or for its share of the machine (true: fewer, longer items); the
  // autotuner picks
  bool share = false;
  // the batch kernels take G from the gradient pass: plans with gmax_mode 2 run with it off
  // while they belong to the batch, and get it back when the batch is destroyed
  std::vector<char> spec_saved;
  // stage work lists; wl[4]: y-face items of shots without compact y-face arrays (pml > 32),
  // launched with wl[3] (update kind 1 against kind 2)
  WorkList wl[5];
  void *args_g[2] = {}, *args_u[2] = {};  // per-parity argument blocks of every shot
  // the same for the persistent launch, without the amplitude bound (no bound pass there)
  void *args_gp[2] = {}, *args_up[2] = {};
  // amplitude bound of the gradient pass: the stage launches record each shot's block maxima
  // and k_gbound_batch seeds each shot's G; amp_ok: the maxima cover every shot's u (false after
  // persistent or per-shot steps, which do not record them)
  bool amp_any = false;
  bool amp_ok = false;  // until the first reset
  // autotuner: every stage does its full work (no tile skipped, no map), so that the timings
  // compare the shapes and modes at a fully active wavefield, not at rest
  bool tuning = false;
  bool hist_ended = false;  // the shots' rest histories were ended since the last reset
  int *first = nullptr;  // [nshots] an item of each shot in the gradient list (the bound pass's fallback)
  // quiet and rest maps of every shot (k_quiet_batch after the gradient pass; the updates read
  // them on steps whose maxima cover u)
  QuietArgs *qargs = nullptr;
  unsigned qblocks = 0;  // the largest shot's block count
  int *kdev = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t ev[2] = {};
  int steps_done = 0;
  bool timing = false;  // autotuner: events around stage `timed_stage`
  int timed_stage = 0;
  double timed_ms = 0;
  // persistent mode: qp3d_batch_run is one k_persist launch over its steps
  bool persist = false;
  int pset = 0;                  // tile set (PSET_TILES)
  int4 *pitems = nullptr;        // one step's items in queue order
  int pips = 0;                  // items per step
  unsigned *pnitems = nullptr;   // [2][nshots] items per shot and step: gradient, update
  unsigned *pctr = nullptr;      // queue, error word, [2][nshots] items done
  int pblocks = 0;               // resident workgroups launched
  bool pshare = false;           // items cut for each shot's share of the machine
  bool pbuilt = false;
  double wait_us = 5e6;          // bound of one dependency wait
};

constexpr int GBOUND_BLOCKS_BATCH = 32;  // workgroups of the bound pass per shot

// Work list of one stage over all shots, items in batch format (stage 3 fills wl[3] and wl[4]).
static int build_stage(qp3d_batch *b, int stage)
{
  int occ[3] = {1, 1, 1};  // stage 3: update kinds 1 and 2
  with_storage(b->storage, [&](auto sp) {
    using S = decltype(sp);
    if (stage == 0) occ[0] = occupancy_gradmax_batch<S>(b->tile[0]);
    else if (stage == 1) occ[0] = occupancy_update_batch<S, 0>(b->tile[1]);
    else if (stage == 2) occ[0] = occupancy_update_batch<S, 1>(b->tile[2]);
    else {
      occ[1] = occupancy_update_batch<S, 1>(b->tile[3]);
      occ[2] = occupancy_update_batch<S, 2>(b->tile[3]);
    }
  });
  const int n = (int)b->plans.size();
  std::vector<int4> all, all1;
  for (int i = 0; i < n; i++) {
    qp3d_plan *p = b->plans[i];
    const int o = stage == 3 ? occ[p->yw ? 2 : 1] : occ[0];
    const int cap = std::max(1, o * qp3d_cu_count(p) / (b->share ? n : 1));
    std::vector<int4> it, ip, ipy;
    if (stage == 0)  // 32-plane items where the amplitude bound can skip them
      qp3d_items_uniform(p, b->tile[0], cap, it, p->ampg.b && p->amp_on ? AMP_ZCHUNK : 0);
    else
      qp3d_items_update(p, b->tile[1], b->tile[2], cap, cap, it, ip, ipy, b->tile[3], cap);
    const std::vector<int4> &v = stage == 3 ? ipy : stage == 2 ? ip : it;
    std::vector<int4> &dst = stage == 3 && !p->yw ? all1 : all;
    for (int4 w : v) dst.push_back(make_int4(w.x, w.y, w.z | (w.w << 16), i));
  }
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  if (stage == 0) {
    // each shot's first item in the uploaded (XCD) order: empty padding items name shot 0
    std::vector<int4> ord;
    (void)qp3d_order_items(all, true, ord);
    std::vector<int> first(n, 0);
    std::vector<char> seen(n, 0);
    for (size_t j = 0; j < ord.size(); j++) {
      const int4 w = ord[j];
      const int sh = w.w;
      if (sh >= 0 && sh < n && !seen[sh] && (w.z & 0xffff) < (int)((unsigned)w.z >> 16)) {
        seen[sh] = 1;
        first[sh] = (int)j;
      }
    }
    if (!b->first) HIP_TRY(hipMalloc((void **)&b->first, sizeof(int) * (size_t)n));
    HIP_TRY(hipMemcpy(b->first, first.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
  }
  int rc = qp3d_upload_items(b->wl[stage], all, true);
  if (!rc && stage == 3) rc = qp3d_upload_items(b->wl[4], all1, true);
  return rc;
}

template <class S>
static int upload_args(qp3d_batch *b)
{
  const size_t n = b->plans.size();
  b->amp_any = false;
  for (int par = 0; par < 2; par++) {
    std::vector<GradArgs<S>> g(n), gp(n);
    std::vector<UpdArgs<S>> u(n), up(n);
    for (size_t i = 0; i < n; i++) {
      qp3d_plan *p = b->plans[i];
      make_args<S>(p, par, Step{nullptr, 0}, g[i], u[i]);
      g[i].gcount = nullptr;
      g[i].gseed_out = nullptr;
      u[i].amp.q = nullptr;
      const bool amp = p->ampg.b && p->amp_on;
      if (amp) {
        // the shot's block maxima, key and seed: tiles under the bound are skipped; its maps
        g[i].skip = 1;
        u[i].amp.q = p->ampg.q;
        b->amp_any = true;
      } else {
        g[i].amp = AmpGrid{};
        g[i].argkey = nullptr;
        g[i].skip = 0;
        g[i].gseed = nullptr;
        u[i].amp = AmpGrid{};
      }
      gp[i] = g[i];
      up[i] = u[i];
      gp[i].amp = AmpGrid{};
      gp[i].argkey = nullptr;
      gp[i].skip = 0;
      gp[i].gseed = nullptr;
      up[i].amp = AmpGrid{};
    }
    HIP_TRY(hipMalloc(&b->args_g[par], sizeof(GradArgs<S>) * n));
    HIP_TRY(hipMalloc(&b->args_u[par], sizeof(UpdArgs<S>) * n));
    HIP_TRY(hipMalloc(&b->args_gp[par], sizeof(GradArgs<S>) * n));
    HIP_TRY(hipMalloc(&b->args_up[par], sizeof(UpdArgs<S>) * n));
    HIP_TRY(hipMemcpy(b->args_g[par], g.data(), sizeof(GradArgs<S>) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->args_u[par], u.data(), sizeof(UpdArgs<S>) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->args_gp[par], gp.data(), sizeof(GradArgs<S>) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->args_up[par], up.data(), sizeof(UpdArgs<S>) * n, hipMemcpyHostToDevice));
  }
  std::vector<QuietArgs> qa(n);
  b->qblocks = 0;
  for (size_t i = 0; i < n; i++) {
    qp3d_plan *p = b->plans[i];
    QuietArgs &q = qa[i];
    q.ag = p->ampg.b && p->amp_on ? p->ampg : AmpGrid{};
    q.g = p->geo;
    q.c = p->coef;
    q.gmax = p->gmax;
    q.hist = p->rflag;
    q.sb = SrcBox{p->src_on ? 1 : 0, p->sb[0], p->sb[1], p->sb[2], p->sb[3], p->sb[4], p->sb[5]};
    if (q.ag.b) b->qblocks = std::max(b->qblocks, q.ag.nblk);
  }
  HIP_TRY(hipMalloc((void **)&b->qargs, sizeof(QuietArgs) * n));
  HIP_TRY(hipMemcpy(b->qargs, qa.data(), sizeof(QuietArgs) * n, hipMemcpyHostToDevice));
  return QP3D_OK;
}

// The persistent queue of one slot (see qp3d_persist.hpp): the gradient tiles of the even shots,
// the update tiles (interior, x/z faces, y faces, rim source) of the odd shots for the step
// before the slot, the update tiles of the even shots, the gradient tiles of the odd shots; with
// the set's tile shapes and each shot's items cut for the whole machine or for its share of it
// (b->pshare).
static int build_persist(qp3d_batch *b)
{
  int occ = 1;
  with_storage(b->storage, [&](auto sp) {
    using S = decltype(sp);
    switch (b->pset) {
      case 0: occ = occupancy_persist<S, 0>(); break;
      case 1: occ = occupancy_persist<S, 1>(); break;
      default: occ = occupancy_persist<S, 2>(); break;
    }
  });
  const int n = (int)b->plans.size();
  const int ncu = qp3d_cu_count(b->plans[0]);
  const int *t = PSET_TILES[b->pset];
  const int cap = std::max(1, occ * ncu / (b->pshare ? n : 1));
  std::vector<int4> grad[2], upd[2];  // by group (shot parity)
  std::vector<unsigned> nitems(2 * (size_t)n, 0u);
  auto put = [](std::vector<int4> &dst, const std::vector<int4> &src, int shot, int kind, int prev) {
    unsigned c = 0;
    for (const int4 &w : src) {
      if (w.z >= w.w) continue;
      dst.push_back(make_int4(w.x, w.y, w.z | (w.w << 16), shot | (kind << 16) | (prev << 24)));
      c++;
    }
    return c;
  };
  for (int i = 0; i < n; i++) {
    const qp3d_plan *p = b->plans[i];
    const int g = i & 1;  // group B (odd shots) runs its update one slot late
    std::vector<int4> ig, ii, ip, iy;
    qp3d_items_uniform(p, t[0], cap, ig);
    qp3d_items_update(p, t[1], t[2], cap, cap, ii, ip, iy, t[3], cap);
    nitems[i] = put(grad[g], ig, i, PK_GRAD, 0);
    unsigned c = put(upd[g], ii, i, PK_INT, g);
    c += put(upd[g], ip, i, PK_PML, g);
    c += put(upd[g], iy, i, p->yw ? PK_PMLY : PK_PMLYF, g);
    if (p->src_rim) {
      upd[g].push_back(make_int4(0, 0, 0, i | (PK_RIM << 16) | (g << 24)));
      c++;
    }
    nitems[n + i] = c;
  }
  std::vector<int4> q;
  for (const std::vector<int4> *v : {&grad[0], &upd[1], &upd[0], &grad[1]}) q.insert(q.end(), v->begin(), v->end());
  (void)hipFree(b->pitems);
  b->pitems = nullptr;
  b->pips = (int)q.size();
  HIP_TRY(hipMalloc((void **)&b->pitems, sizeof(int4) * q.size()));
  HIP_TRY(hipMemcpy(b->pitems, q.data(), sizeof(int4) * q.size(), hipMemcpyHostToDevice));
  if (!b->pnitems) HIP_TRY(hipMalloc((void **)&b->pnitems, sizeof(unsigned) * 2 * n));
  HIP_TRY(hipMemcpy(b->pnitems, nitems.data(), sizeof(unsigned) * 2 * n, hipMemcpyHostToDevice));
  if (!b->pctr) {
    HIP_TRY(hipMalloc((void **)&b->pctr, sizeof(unsigned) * (2 + 2 * (size_t)n)));
    HIP_TRY(hipMemset(b->pctr, 0, sizeof(unsigned) * (2 + 2 * (size_t)n)));
  }
  b->pblocks = occ * ncu;
  b->pbuilt = true;
  return QP3D_OK;
}

// Steps [k0, k1) of every shot in one persistent launch (after the counters are zeroed).
template <class S>
static int persist_run_s(qp3d_batch *b, int k0, int k1)
{
  const int n = (int)b->plans.size();
  hipLaunchKernelGGL(k_persist_init, dim3(1), dim3(256), 0, b->stream, b->pctr, 2 + 2 * n);
  HIP_TRY(hipGetLastError());
  int khz = 100000;  // wall_clock64 rate
  (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device);
  PersistArgs<S> a;
  for (int par = 0; par < 2; par++) {
    a.ag[par] = (const GradArgs<S> *)b->args_gp[par];
    a.au[par] = (const UpdArgs<S> *)b->args_up[par];
  }
  a.items = b->pitems;
  a.ips = b->pips;
  a.nshots = n;
  a.k0 = k0;
  a.k1 = k1;
  a.nitems = b->pnitems;
  a.ctr = b->pctr;
  a.wait_ticks = (long long)(b->wait_us * 1e-3 * khz);
  switch (b->pset) {
    case 0: launch_persist<S, 0>(b->pblocks, b->stream, a); break;
    case 1: launch_persist<S, 1>(b->pblocks, b->stream, a); break;
    default: launch_persist<S, 2>(b->pblocks, b->stream, a); break;
  }
  HIP_TRY(hipGetLastError());
  return QP3D_OK;
}

template <class S>
static int batch_step_s(qp3d_batch *b, int k, Step st)
{
  using T = typename S::T;
  const int par = k & 1;
  if (b->per_shot) {
    b->amp_ok = false;  // the plans' own steps here do not record the block maxima
    for (qp3d_plan *p : b->plans) {
      const int rc = qp3d_forward_step_on(p, k, st, b->stream);
      if (rc) return rc;
    }
    return QP3D_OK;
  }
  // this step's maps: only where the block maxima cover u (as the gradient pass's skipping);
  // a step without them ends the shots' rest histories (never inside a capture, which starts
  // with the maxima covering u)
  const int maps = b->amp_any && b->amp_ok && !b->tuning ? 1 : 0;
  if (b->amp_any && !maps && !b->hist_ended) {
    for (qp3d_plan *p : b->plans)
      if (p->rflag) HIP_TRY(hipMemsetAsync(p->rflag, 0, sizeof(unsigned), b->stream));
    b->hist_ended = true;
  }
  for (int s = 0; s < 4; s++) {
    const bool tm = b->timing && s == b->timed_stage;
    if (tm) HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    const UpdArgs<S> *A = (const UpdArgs<S> *)b->args_u[par];
    if (b->wl[s].n) {
      if (s == 0) {
        const GradArgs<S> *G = (const GradArgs<S> *)b->args_g[par];
        // the bound pass of every shot (seed of G_k; clears the maxima step k+1 fills), then the
        // pass itself, skipping only while every shot's maxima cover its u
        if (b->amp_any) {
          launch_gbound_batch<S>(b->tile[0], GBOUND_BLOCKS_BATCH, (int)b->plans.size(), b->wl[0].d, b->wl[0].n,
                                 b->first, b->stream, G, st);
          HIP_TRY(hipGetLastError());
        }
        launch_gradmax_batch<S>(b->tile[0], b->wl[0].n, b->wl[0].d, b->stream, G, st, b->amp_ok && !b->tuning ? 1 : 0);
        if (maps) {
          // G_k of every shot is complete: its quiet and rest maps for the updates
          HIP_TRY(hipGetLastError());
          hipLaunchKernelGGL(k_quiet_batch, dim3((b->qblocks + 255) / 256, (unsigned)b->plans.size()), dim3(256), 0,
                             b->stream, (const QuietArgs *)b->qargs, st);
        }
      } else if (s == 1)
        launch_update_batch<S, 0>(b->tile[1], b->wl[1].n, b->wl[1].d, b->stream, A, st, maps);
      else if (s == 2)
        launch_update_batch<S, 1>(b->tile[2], b->wl[2].n, b->wl[2].d, b->stream, A, st, maps);
      else
        launch_update_batch<S, 2>(b->tile[3], b->wl[3].n, b->wl[3].d, b->stream, A, st, maps);
      HIP_TRY(hipGetLastError());
    }
    if (s == 3 && b->wl[4].n) {
      launch_update_batch<S, 1>(b->tile[3], b->wl[4].n, b->wl[4].d, b->stream, A, st, maps);
      HIP_TRY(hipGetLastError());
    }
    if (tm) {
      HIP_TRY(hipEventRecord(b->ev[1], b->stream));
      HIP_TRY(hipEventSynchronize(b->ev[1]));
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
      b->timed_ms += ms;
    }
  }
  // source cells on the frozen rim of any shot (u_next = u_cur + s w)
  for (qp3d_plan *p : b->plans) {
    if (!p->src_rim) continue;
    const int prev = k & 1, cur = prev ^ 1;
    const long nb = (long)(p->sb[1] - p->sb[0] + 1) * (p->sb[3] - p->sb[2] + 1) * (p->sb[5] - p->sb[4] + 1);
    const int blocks = (int)std::min<long>((nb + 255) / 256, 4096);
    hipLaunchKernelGGL(k_rim_source<S>, dim3(blocks), dim3(256), 0, b->stream, (const T *)p->u[cur], (T *)p->u[prev],
                       p->geo, p->sb[0], p->sb[1], p->sb[2], p->sb[3], p->sb[4], p->sb[5], (const float *)p->sbox,
                       (const float *)p->wav, st, b->amp_any && p->amp_on ? p->ampg : AmpGrid{});
    HIP_TRY(hipGetLastError());
  }
  b->amp_ok = true;  // u(k+1) of every shot is covered by the maxima these launches recorded
  return QP3D_OK;
}

static int batch_step(qp3d_batch *b, int k, Step st)
{
  int rc = QP3D_OK;
  with_storage(b->storage, [&](auto sp) { rc = batch_step_s<decltype(sp)>(b, k, st); });
  return rc;
}

static void batch_free(qp3d_batch *b)
{
  if (!b) return;
  for (size_t i = 0; i < b->spec_saved.size(); i++) {
    b->plans[i]->spec = b->spec_saved[i] != 0;
    b->plans[i]->in_batch = 0;
  }
  for (WorkList &w : b->wl) (void)hipFree(w.d);
  for (int i = 0; i < 2; i++) {
    (void)hipFree(b->args_g[i]);
    (void)hipFree(b->args_u[i]);
    (void)hipFree(b->args_gp[i]);
    (void)hipFree(b->args_up[i]);
  }
  (void)hipFree(b->first);
  (void)hipFree(b->qargs);
  (void)hipFree(b->kdev);
  (void)hipFree(b->pitems);
  (void)hipFree(b->pnitems);
  (void)hipFree(b->pctr);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

extern "C" int qp3d_batch_create(qp3d_plan *const *plans, int nshots, int use_graph, qp3d_batch **out)
{
  if (!plans || !out || nshots < 1) {
    qp3d_set_error("batch_create: need at least one plan");
    return QP3D_ERR_ARG;
  }
  *out = nullptr;
  const qp3d_config &c0 = plans[0]->cfg;
  for (int i = 0; i < nshots; i++) {
    const qp3d_plan *p = plans[i];
    if (!p || !p->uploaded) {
      qp3d_set_error("batch_create: plan %d is null or has no model uploaded", i);
      return QP3D_ERR_STATE;
    }
    if (p->cfg.nt != c0.nt || p->cfg.storage != c0.storage || p->cfg.device != c0.device || p->cfg.nz >= 65536) {
      qp3d_set_error("batch_create: plan %d differs in nt, storage or device from plan 0", i);
      return QP3D_ERR_ARG;
    }
    // one plan holds one shot's wavefields, record and maxima: it can be driven by one batch
    // once, never twice in a batch or by two live batches
    if (p->in_batch) {
      qp3d_set_error("batch_create: plan %d already belongs to a live batch", i);
      return QP3D_ERR_STATE;
    }
    for (int j = 0; j < i; j++)
      if (plans[j] == p) {
        qp3d_set_error("batch_create: plan %d appears twice (also as plan %d)", i, j);
        return QP3D_ERR_ARG;
      }
  }
  HIP_TRY(hipSetDevice(c0.device));
  qp3d_batch *b = new qp3d_batch();
  b->plans.assign(plans, plans + nshots);
  for (int i = 0; i < nshots; i++) b->gen.push_back(plans[i]->upload_gen);
  for (int i = 0; i < nshots; i++) {
    b->spec_saved.push_back(plans[i]->spec ? 1 : 0);
    plans[i]->spec = false;
    plans[i]->in_batch = 1;
  }
  b->nt = c0.nt;
  b->storage = c0.storage;
  b->device = c0.device;
  b->use_graph = use_graph != 0;
  auto fail = [&](int rc) {
    batch_free(b);
    return rc;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess ||
      hipMalloc((void **)&b->kdev, sizeof(int)) != hipSuccess) {
    qp3d_set_error("batch_create: stream, events or step index allocation failed");
    return fail(QP3D_ERR_HIP);
  }
  int rc = QP3D_OK;
  with_storage(b->storage, [&](auto sp) { rc = upload_args<decltype(sp)>(b); });
  if (rc) return fail(rc);
  for (int s = 0; s < 4; s++)
    if ((rc = build_stage(b, s))) return fail(rc);
  // the plans' own streams may still hold uploads
  for (qp3d_plan *p : b->plans) HIP_TRY(hipStreamSynchronize(p->stream));
  *out = b;
  return QP3D_OK;
}

extern "C" int qp3d_batch_destroy(qp3d_batch *b)
{
  if (!b) return QP3D_OK;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  batch_free(b);
  return QP3D_OK;
}

// Zeroes every shot's wavefields, PML state, maxima, flag and record on the batch stream.
extern "C" int qp3d_batch_reset(qp3d_batch *b)
{
  if (!b) return QP3D_ERR_ARG;
  HIP_TRY(hipSetDevice(b->device));
  for (qp3d_plan *p : b->plans) {
    HIP_TRY(hipMemsetAsync(p->state, 0, p->state_bytes, b->stream));
    HIP_TRY(hipMemsetAsync(p->gmax, 0, sizeof(unsigned) * (size_t)b->nt, b->stream));
    HIP_TRY(hipMemsetAsync(p->flag, 0, sizeof(int), b->stream));
    if (p->rec) HIP_TRY(hipMemsetAsync(p->rec, 0, sizeof(float) * (size_t)b->nt * p->hx * p->hy, b->stream));
    if (p->ampg.b) {
      HIP_TRY(hipMemsetAsync(p->ampg.b, 0, sizeof(unsigned) * 9 * (size_t)p->ampg.nblk, b->stream));
      HIP_TRY(hipMemsetAsync(p->argkey, 0, ARG_WORDS * sizeof(unsigned long long), b->stream));
      // every block at rest, and a rest history from here
      HIP_TRY(hipMemsetAsync(p->ampg.q, 0, (size_t)p->ampg.nblk, b->stream));
      HIP_TRY(hipMemsetAsync(p->rflag, 1, sizeof(unsigned), b->stream));
    }
    p->steps_done = 0;
  }
  b->steps_done = 0;
  b->amp_ok = true;  // u = 0 everywhere, and so are the block maxima
  b->hist_ended = false;
  return QP3D_OK;
}

static int batch_capture(qp3d_batch *b)
{
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = QP3D_OK;
  // every replay starts with the block maxima covering u (qp3d_batch_run checks amp_ok)
  b->amp_ok = true;
  for (int i = 0; i < GRAPH_LEN && rc == QP3D_OK; i++) rc = batch_step(b, i, Step{b->kdev, i});
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc == QP3D_OK && e != hipSuccess) {
    qp3d_set_error("batch_run: graph capture failed: %s", hipGetErrorString(e));
    rc = QP3D_ERR_HIP;
  }
  if (rc == QP3D_OK && (e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0)) != hipSuccess) {
    qp3d_set_error("batch_run: graph instantiation failed: %s", hipGetErrorString(e));
    b->gexec = nullptr;
    rc = QP3D_ERR_HIP;
  }
  if (rc == QP3D_OK)
    b->graph = g;
  else if (g)
    (void)hipGraphDestroy(g);
  return rc;
}

// Enqueues time steps [k0, k1) of every shot (asynchronous on the batch's stream).
extern "C" int qp3d_batch_run(qp3d_batch *b, int k0, int k1)
{
  if (!b || k0 < 0 || k1 > b->nt || k0 > k1) {
    qp3d_set_error("batch_run: bad arguments");
    return QP3D_ERR_ARG;
  }
  for (size_t i = 0; i < b->plans.size(); i++)
    if (b->plans[i]->upload_gen != b->gen[i]) {
      qp3d_set_error("batch_run: plan %zu was uploaded again after the batch was built; build a new batch", i);
      return QP3D_ERR_STATE;
    }
  HIP_TRY(hipSetDevice(b->device));
  // the batch kernels leave the plans' block maxima stale (also when a run stops on an error)
  for (qp3d_plan *p : b->plans) p->amp_ok = false;
  if (b->persist && !b->per_shot) {
    if (!b->pbuilt) {
      int rc = build_persist(b);
      if (rc) return rc;
    }
    if (k1 > k0) {
      // the persistent launch records no block maxima and keeps no map: the shots' rest
      // histories end here
      b->amp_ok = false;
      if (!b->hist_ended) {
        for (qp3d_plan *p : b->plans)
          if (p->rflag) HIP_TRY(hipMemsetAsync(p->rflag, 0, sizeof(unsigned), b->stream));
        b->hist_ended = true;
      }
      int rc = QP3D_OK;
      with_storage(b->storage, [&](auto sp) { rc = persist_run_s<decltype(sp)>(b, k0, k1); });
      if (rc) return rc;
    }
    b->steps_done = k1;
    for (qp3d_plan *p : b->plans) p->steps_done = k1;
    return QP3D_OK;
  }
  const bool graph = b->use_graph && !b->timing;
  for (int k = k0; k < k1;) {
    if (graph && (k & 1) == 0 && k + GRAPH_LEN <= k1 && (b->amp_ok || !b->amp_any)) {
      if (!b->gexec) {
        int rc = batch_capture(b);
        if (rc) return rc;
      }
      hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, b->stream, b->kdev, k);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipGraphLaunch(b->gexec, b->stream));
      k += GRAPH_LEN;
    } else {
      int rc = batch_step(b, k, Step{nullptr, k});
      if (rc) return rc;
      k++;
    }
  }
  b->steps_done = k1;
  for (qp3d_plan *p : b->plans) p->steps_done = k1;
  return QP3D_OK;
}

// Waits for the batch; QP3D_ERR_NONFINITE names the first shot whose wavefield blew up.
extern "C" int qp3d_batch_sync(qp3d_batch *b)
{
  if (!b) return QP3D_ERR_ARG;
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (b->pctr) {
    unsigned err = 0;
    HIP_TRY(hipMemcpy(&err, b->pctr + 1, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (err) {
      qp3d_set_error("batch: a wait of the persistent launch passed its bound (%.0f us); the shots' state is "
                     "incomplete",
                     b->wait_us);
      return QP3D_ERR_TIMEOUT;
    }
  }
  for (size_t i = 0; i < b->plans.size(); i++) {
    int flag = 0;
    HIP_TRY(hipMemcpy(&flag, b->plans[i]->flag, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) {
      qp3d_set_error("batch: the wavefield of shot %zu became non-finite (inf/NaN)", i);
      return QP3D_ERR_NONFINITE;
    }
  }
  return QP3D_OK;
}

// Times every tile shape of each stage over `steps` time steps from rest (events around that
// stage's launch) and keeps the fastest; ms[s] receives the chosen shape's mean ms per step.
static int tune_stages(qp3d_batch *b, int steps, int *tiles, double *ms)
{
  int rc;
  for (int s = 0; s < 4; s++) {
    double best = DBL_MAX;
    int bt = b->tile[s];
    for (int t = 0; t < NTILES; t++) {
      b->tile[s] = t;
      if ((rc = build_stage(b, s))) return rc;
      if ((rc = qp3d_batch_reset(b))) return rc;
      if ((rc = qp3d_batch_run(b, 0, 2))) return rc;  // warm-up and a non-zero field
      b->timing = true;
      b->timed_stage = s;
      b->timed_ms = 0;
      rc = qp3d_batch_run(b, 2, 2 + steps);
      b->timing = false;
      if (rc) return rc;
      const double t_ms = b->timed_ms / steps;
      if (t_ms < best) {
        best = t_ms;
        bt = t;
      }
    }
    b->tile[s] = bt;
    if ((rc = build_stage(b, s))) return rc;
    tiles[s] = bt;
    ms[s] = best;
  }
  return QP3D_OK;
}

// Device time (ms) of `steps` whole time steps from rest, direct launches.
static int time_steps(qp3d_batch *b, int steps, double &ms_out)
{
  int rc;
  if ((rc = qp3d_batch_reset(b)) || (rc = qp3d_batch_run(b, 0, 2))) return rc;
  float ms = 0;
  HIP_TRY(hipEventRecord(b->ev[0], b->stream));
  if ((rc = qp3d_batch_run(b, 2, 2 + steps))) return rc;
  HIP_TRY(hipEventRecord(b->ev[1], b->stream));
  HIP_TRY(hipEventSynchronize(b->ev[1]));
  HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  ms_out = ms;
  return QP3D_OK;
}

// For each way of cutting the shared launches' work items (each shot cut for the whole machine,
// or for its share of it) times every tile shape per stage and then whole steps with the
// fastest shapes; times whole steps of each shot's own launches one after another; keeps the
// fastest of the three and resets the batch. tiles_out (4 ints, may be NULL) receives the tile
// choice (1-based, as in qp3d_config.tile_*: gradient max, interior, PML x/z faces, PML y
// faces) of the shared launches; ms_out (4 doubles, may be NULL) their stages' mean ms per step.
extern "C" int qp3d_batch_autotune(qp3d_batch *b, int steps, int *tiles_out, double *ms_out)
{
  if (!b || steps < 1 || steps + 2 > b->nt) {
    qp3d_set_error("batch_autotune: need 1 <= steps <= nt - 2");
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(b->device));
  const bool graph = b->use_graph;
  b->use_graph = false;
  b->tuning = true;
  int tiles[2][4] = {}, rc = QP3D_OK;
  double ms[2][4] = {}, t[3] = {0, 0, 0};
  b->per_shot = false;
  for (int sh = 0; sh < 2 && rc == QP3D_OK; sh++) {
    b->share = sh == 1;
    if ((rc = tune_stages(b, steps, tiles[sh], ms[sh]))) break;
    rc = time_steps(b, steps, t[sh]);
  }
  if (rc == QP3D_OK) {
    b->per_shot = true;
    rc = time_steps(b, steps, t[2]);
    b->per_shot = false;
  }
  // one persistent launch: every tile set with each cut
  double tp[NPSETS][2] = {};
  for (int set = 0; set < NPSETS && rc == QP3D_OK; set++)
    for (int sh = 0; sh < 2 && rc == QP3D_OK; sh++) {
      b->persist = true;
      b->pset = set;
      b->pshare = sh == 1;
      b->pbuilt = false;
      rc = time_steps(b, steps, tp[set][sh]);
      b->persist = false;
    }
  b->use_graph = graph;
  b->tuning = false;
  if (rc) return rc;
  const int best = t[2] < std::min(t[0], t[1]) ? 2 : (t[1] < t[0] ? 1 : 0);
  const int sh = best == 1 ? 1 : 0;  // the shared launches' form kept (also behind per-shot mode)
  b->per_shot = best == 2;
  b->share = sh == 1;
  for (int s = 0; s < 4; s++) {
    b->tile[s] = tiles[sh][s];
    if ((rc = build_stage(b, s))) return rc;
    if (tiles_out) tiles_out[s] = tiles[sh][s] + 1;
    if (ms_out) ms_out[s] = ms[sh][s];
  }
  int pbest = 0, pbsh = 0;
  for (int set = 0; set < NPSETS; set++)
    for (int h = 0; h < 2; h++)
      if (tp[set][h] < tp[pbest][pbsh]) {
        pbest = set;
        pbsh = h;
      }
  b->pset = pbest;
  b->pshare = pbsh == 1;
  b->pbuilt = false;
  if (tp[pbest][pbsh] < t[best]) {
    b->persist = true;
    b->per_shot = false;
  }
  // a graph captured before holds another mode's launches
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  return qp3d_batch_reset(b);
}

// How the shared launches cut each shot's work items: 1 for the shot's share of the machine,
// 0 for the whole machine.
extern "C" int qp3d_batch_share(const qp3d_batch *b) { return b ? (b->share ? 1 : 0) : QP3D_ERR_ARG; }

extern "C" int qp3d_batch_per_shot(const qp3d_batch *b) { return b ? (b->per_shot ? 1 : 0) : QP3D_ERR_ARG; }

// Persistent mode: -1 when off, else the tile set (PSET_TILES) of the one launch per run.
extern "C" int qp3d_batch_persistent(const qp3d_batch *b) { return b ? (b->persist ? b->pset : -1) : QP3D_ERR_ARG; }

// The persistent form the batch would use (the fastest one the last autotune timed, or the one
// set last): its tile set and cut, whether or not persistent mode is on.
extern "C" int qp3d_batch_persistent_choice(const qp3d_batch *b, int *set, int *share)
{
  if (!b || !set || !share) return QP3D_ERR_ARG;
  *set = b->pset;
  *share = b->pshare ? 1 : 0;
  return QP3D_OK;
}

// Turns persistent mode on with tile set `set` (0 .. NPSETS-1) and each shot's items cut for
// its share of the machine (share = 1) or for the whole machine (0), or off (set = -1).
extern "C" int qp3d_batch_set_persistent(qp3d_batch *b, int set, int share)
{
  if (!b || set < -1 || set >= NPSETS || share < 0 || share > 1) {
    qp3d_set_error("batch_set_persistent: bad arguments");
    return QP3D_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->persist = set >= 0;
  if (b->persist) {
    b->per_shot = false;
    b->pset = set;
    b->pshare = share == 1;
    b->pbuilt = false;
  }
  return QP3D_OK;
}

// Bound of one dependency wait of the persistent launch, in microseconds (default 5 s); a wait
// past it makes the launch leave and qp3d_batch_sync return QP3D_ERR_TIMEOUT.
extern "C" int qp3d_batch_set_wait_limit(qp3d_batch *b, double us)
{
  if (!b || !(us >= 0)) {
    qp3d_set_error("batch_set_wait_limit: bad arguments");
    return QP3D_ERR_ARG;
  }
  b->wait_us = us;
  return QP3D_OK;
}

extern "C" int qp3d_batch_set_per_shot(qp3d_batch *b, int on)
{
  if (!b || on < 0 || on > 1) {
    qp3d_set_error("batch_set_per_shot: bad arguments");
    return QP3D_ERR_ARG;
  }
  if (b->per_shot != (on == 1)) {
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->per_shot = on == 1;
  }
  if (on == 1) b->persist = false;
  return QP3D_OK;
}

// Algorithmic HBM bytes of one batched time step (every shot), as qp3d_plan_step_bytes sums.
extern "C" int qp3d_batch_step_bytes(const qp3d_batch *b, double *bytes)
{
  if (!b || !bytes) return QP3D_ERR_ARG;
  bytes[0] = bytes[1] = bytes[2] = bytes[3] = 0;
  for (const qp3d_plan *p : b->plans) {
    double v[5];
    int rc = qp3d_plan_step_bytes(p, v);
    if (rc) return rc;
    bytes[0] += v[0];
    bytes[1] += v[1] + v[2];
    bytes[2] += (double)(p->cfg.nx - 10) * (p->cfg.ny - 10) * (p->cfg.nz - 10);  // updated cells
    bytes[3] += (double)p->hx * p->hy * p->hz;                                 // interior cells
  }
  return QP3D_OK;
}
