// qp3d_modeling3d_noalloc: the drop-in shot (qp3d_modeling3d) for callers that cannot allocate
// device memory. Inputs and the record are the caller's device arrays; everything else the
// time loop needs (wavefields, PML state, the converted model, the PML table, per-step maxima,
// the non-finite flag, the source box and the work lists) is carved out of one fixed device
// array, `scratch`, that every GPU the library runs on holds until the process ends. The
// kernels are the plan's own (same tiles, same work items), so the record is the one
// qp3d_modeling3d returns for the same inputs.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "qp3d_plan_impl.hpp"

namespace {

static __device__ __align__(256) unsigned char scratch[QP3D_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
void *g_scratch[MAX_DEVICES] = {};  // device address of `scratch` on each device (set once)
std::mutex g_mutex;

// The device address of this device's copy of `scratch`: hipGetSymbolAddress once per device,
// with the device current. The array's name itself is a host-side stand-in and is never used
// as a device address.
int scratch_of(int dev, unsigned char **out)
{
  if (dev < 0 || dev >= MAX_DEVICES) {
    qp3d_set_error("modeling3d_noalloc: device %d outside [0,%d)", dev, MAX_DEVICES);
    return QP3D_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      qp3d_set_error("modeling3d_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                     hipGetErrorString(e));
      return QP3D_ERR_HIP;
    }
    g_scratch[dev] = ptr;
  }
  *out = (unsigned char *)g_scratch[dev];
  return QP3D_OK;
}

// ptr is device memory of device `dev` (memory the runtime does not know is plain host memory).
bool device_memory_of(const void *ptr, int dev)
{
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, ptr) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice && a.device == dev;
}

// Dense copy of the source box [z0,z0+bz) x [x0,x0+bx) x [y0,y0+by) of a full-grid source.
__global__ void k_gather_box(const float *src, int nx, int ny, int z0, int x0, int y0, int bz, int bx, int by,
                             float *dst)
{
  const long n = (long)bz * bx * by;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int yy = (int)(t % by), xx = (int)((t / by) % bx), zz = (int)(t / ((long)by * bx));
    dst[t] = src[((long)(z0 + zz) * nx + (x0 + xx)) * ny + (y0 + yy)];
  }
}

// Bump allocation inside the scratch array (256-byte aligned pieces).
struct Carve {
  unsigned char *base;
  size_t used = 0;
  bool fits = true;
  void *take(size_t bytes)
  {
    const size_t at = (used + 255) / 256 * 256;
    if (at + bytes > QP3D_NOALLOC_SCRATCH_BYTES) {
      fits = false;
      return nullptr;
    }
    used = at + bytes;
    return base + at;
  }
};

int put_list(Carve &cv, WorkList &w, const std::vector<int4> &logical, hipStream_t s)
{
  std::vector<int4> v;
  w.cells = qp3d_order_items(logical, false, v);
  w.n = (int)v.size();
  w.d = nullptr;
  if (!w.n) return QP3D_OK;
  w.d = (int4 *)cv.take(sizeof(int4) * v.size());
  if (!w.d) return QP3D_ERR_ARG;
  HIP_TRY(hipMemcpyAsync(w.d, v.data(), sizeof(int4) * v.size(), hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));  // v is about to go out of scope
  return QP3D_OK;
}

}  // namespace

extern "C" int qp3d_modeling3d_noalloc(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz,
                                       float dt, int pml, int snapflag, int sx, int sy, int sz, const float *vp,
                                       const float *epsilon, const float *delta, const float *source,
                                       const float *wavelet, float *record, const float *dlr, const float *ddlr,
                                       const float *dtb, const float *ddtb, const float *dfb, const float *ddfb,
                                       const float *c, const float *c2, const char *snap_file, void *stream)
{
  (void)ntsnap;
  (void)sx;
  (void)sy;
  (void)snap_file;
  if (!vp || !epsilon || !delta || !source || !wavelet || !record || !dlr || !ddlr || !dtb || !ddtb || !dfb ||
      !ddfb || !c || !c2) {
    qp3d_set_error("modeling3d_noalloc: null argument");
    return QP3D_ERR_ARG;
  }
  if (snapflag == 1) {
    qp3d_set_error("modeling3d_noalloc: snapshots (snapflag == 1) need memory this entry point does not allocate");
    return QP3D_ERR_ARG;
  }
  qp3d_config cfg;
  std::memset(&cfg, 0, sizeof cfg);
  cfg.nx = nx; cfg.ny = ny; cfg.nz = nz; cfg.pml = pml; cfg.nt = nt;
  cfg.dx = dx; cfg.dy = dy; cfg.dz = dz; cfg.dt = dt;
  cfg.sz = sz;
  cfg.storage = QP3D_STORAGE_FP32;
  int rc = qp3d_check_config(cfg, "modeling3d_noalloc");
  if (rc) return rc;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  cfg.device = dev;
  const void *arrays[14] = {vp, epsilon, delta, source, wavelet, record, dlr, ddlr, dtb, ddtb, dfb, ddfb, c, c2};
  for (const void *a : arrays)
    if (!device_memory_of(a, dev)) {
      qp3d_set_error("modeling3d_noalloc: every array must be device memory of the current device %d", dev);
      return QP3D_ERR_ARG;
    }
  unsigned char *base = nullptr;
  if ((rc = scratch_of(dev, &base))) return rc;
  hipStream_t s = (hipStream_t)stream;

  // the plan's host-side description over the scratch array (nothing here is allocated on the
  // device, and the plan is never destroyed)
  qp3d_plan P;
  qp3d_plan_describe(&P, cfg);
  P.stream = s;
  const size_t N = (size_t)P.N;
  const size_t fs = field_stride(sizeof(float) * N);
  Carve cv{base};
  // state block (the plan's layout, FP32), the model in the plan's row-pitched layout
  // ((vp dt)^2, eps, eps - del), the compact y-face model
  char *state = (char *)cv.take(P.state_bytes);
  P.vdt2 = (float *)cv.take(fs);
  P.eps = (float *)cv.take(fs);
  P.epsd = (float *)cv.take(fs);
  float *ymodel = P.yw ? (float *)cv.take(3 * field_stride(sizeof(float) * (size_t)P.ycells)) : nullptr;
  P.pmltab = (float4 *)cv.take(sizeof(float4) * 3 * (size_t)pml);
  P.gmax = (unsigned *)cv.take(sizeof(unsigned) * (size_t)(nt > 0 ? nt : 1));
  P.flag = (int *)cv.take(sizeof(int));
  int *dbox = (int *)cv.take(6 * sizeof(int));
  std::vector<int4> items[4];
  qp3d_plan_items(&P, items);
  if (cv.fits) {
    WorkList *wl[4] = {&P.wl_grad, &P.wl_int, &P.wl_pml, &P.wl_pmly};
    for (int i = 0; i < 4 && cv.fits; i++)
      if ((rc = put_list(cv, *wl[i], items[i], s)) && cv.fits) return rc;
  }
  if (!cv.fits) {
    qp3d_set_error("modeling3d_noalloc: a %dx%dx%d grid with %d steps needs more than the %zu-byte scratch "
                   "array (about 48 bytes per cell); use qp3d_modeling3d",
                   nx, ny, nz, nt, (size_t)QP3D_NOALLOC_SCRATCH_BYTES);
    return QP3D_ERR_ARG;
  }
  P.n_int_cells = P.wl_int.cells;
  P.n_pml_cells = P.wl_pml.cells + P.wl_pmly.cells;
  qp3d_bind_state(&P, state, nullptr);
  if (P.yw) qp3d_bind_ymodel(&P, ymodel);
  P.wav = const_cast<float *>(wavelet);  // read only
  P.rec = record;

  // model in the scratch array (row padding zeroed; the copies from device arrays allocate
  // nothing), then (vp dt)^2 and eps - del in place
  HIP_TRY(hipMemsetAsync(P.vdt2, 0, 3 * fs, s));
  if ((rc = qp3d_copy_field(&P, const_cast<float *>(vp), P.vdt2, sizeof(float), true, s)) ||
      (rc = qp3d_copy_field(&P, const_cast<float *>(epsilon), P.eps, sizeof(float), true, s)) ||
      (rc = qp3d_copy_field(&P, const_cast<float *>(delta), P.epsd, sizeof(float), true, s)))
    return rc;
  qp3d_launch_model(P.vdt2, P.eps, P.epsd, dt, (long)N, s);
  HIP_TRY(hipGetLastError());
  qp3d_launch_ymodel(&P, s);
  HIP_TRY(hipGetLastError());
  std::vector<float4> tab;
  if ((rc = qp3d_make_tables(&P, dlr, ddlr, dtb, ddtb, dfb, ddfb, c, c2, tab))) return rc;
  HIP_TRY(hipMemcpyAsync(P.pmltab, tab.data(), sizeof(float4) * tab.size(), hipMemcpyHostToDevice, s));
  // source box on the device
  int box[6] = {nz, -1, nx, -1, ny, -1};
  HIP_TRY(hipMemcpyAsync(dbox, box, sizeof box, hipMemcpyHostToDevice, s));
  qp3d_launch_source_box(source, nx, ny, nz, dbox, s);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(box, dbox, sizeof box, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));  // box and tab are host memory
  P.src_on = box[1] >= 0;
  if (P.src_on) {
    const int bz = box[1] - box[0] + 1, bx = box[3] - box[2] + 1, by = box[5] - box[4] + 1;
    const long nb = (long)bz * bx * by;
    P.sbox = (float *)cv.take(sizeof(float) * (size_t)nb);
    if (!P.sbox) {
      qp3d_set_error("modeling3d_noalloc: the source's bounding box (%d x %d x %d cells) does not fit the scratch "
                     "array; use qp3d_modeling3d", bz, bx, by);
      return QP3D_ERR_ARG;
    }
    hipLaunchKernelGGL(k_gather_box, dim3((unsigned)std::min<long>((nb + 255) / 256, 4096)), dim3(256), 0, s, source,
                       nx, ny, box[0], box[2], box[4], bz, bx, by, P.sbox);
    HIP_TRY(hipGetLastError());
    for (int i = 0; i < 6; i++) P.sb[i] = box[i];
    P.src_rim = box[0] < H || box[2] < H || box[4] < H || box[1] >= nz - H || box[3] >= nx - H || box[5] >= ny - H;
  }
  P.uploaded = true;

  // every call starts from rest: the scratch array keeps the last call's contents
  HIP_TRY(hipMemsetAsync(P.state, 0, P.state_bytes, s));
  HIP_TRY(hipMemsetAsync(P.gmax, 0, sizeof(unsigned) * (size_t)(nt > 0 ? nt : 1), s));
  HIP_TRY(hipMemsetAsync(P.flag, 0, sizeof(int), s));
  HIP_TRY(hipMemsetAsync(record, 0, sizeof(float) * (size_t)nt * P.hx * P.hy, s));
  for (int k = 0; k < nt; k++)
    if ((rc = qp3d_forward_step_on(&P, k, Step{nullptr, k}, s))) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  int flag = 0;
  HIP_TRY(hipMemcpy(&flag, P.flag, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) {
    qp3d_set_error("wavefield became non-finite (inf/NaN): dt violates the stability limit for this model");
    return QP3D_ERR_NONFINITE;
  }
  return QP3D_OK;
}
