// gfx950 kernels of the qP3D time step.
//
// One time step of the reference (3dPureqPModeling_single_shot.cu:277-337) launches five
// per-cell kernels and copies 16 B/cell over PCIe. Here a step is three launches of two
// stencil passes over the wavefield, hand-tiled for CDNA4:
//   k_gradmax        max over the grid of |du/dx|,|du/dy|,|du/dz| (GPU_kernel.cu:14-33 grad and
//                    the host absMaxval at 3dPureqPModeling_single_shot.cu:284-292), reduced on
//                    device;
//   k_update<PML=0>  interior cells: gradient recomputed in registers, scalar operator S
//                    (GPU_kernel.cu:87-115), 10th-order second derivatives and the leapfrog
//                    (GPU_kernel.cu:552-565), source injection (GPU_kernel.cu:720-731) and the
//                    receiver-plane write (3dPureqPModeling_single_shot.cu:333-337);
//   k_update<PML=1>  the same for the tiles that hold PML cells, with the split-field
//                    recursions (GPU_kernel.cu:566-637).
// Splitting interior and PML tiles keeps the PML state out of the interior kernel's
// registers. The copy-rotation kernel (GPU_kernel.cu:733-812) disappears: u_next and the PML
// state are written in place of the oldest slot (each cell reads that slot only at its own
// index) and the host flips slot parity each step.
//
// Tiling: a workgroup owns an xy tile (at most TX x 64 cells, y unit-stride, one 64-lane
// wavefront per x row group) and marches through a z range; its work item comes from a list
// the host builds (tiles aligned to the PML boundary, z ranges split at the PML planes,
// ordered so that each XCD's L2 sees neighbouring tiles). The current plane with a 5-cell
// halo sits in LDS (double-buffered, one barrier per plane) for the x and y derivatives; each
// lane keeps an 11-deep register queue of its column in z for the z derivatives. Every global
// load of plane z+1 (queue front, halos, model, previous wavefield, PML state) is issued
// before plane z is computed, so HBM latency overlaps the arithmetic of the plane before.
// Cell offsets are 32-bit (grids below 2^30 cells) so loads use a scalar base plus a vector
// offset.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qp3d {

constexpr int H = 5;    // stencil half-width (10th order)
constexpr int TY = 64;  // tile extent along y = wavefront width

// Tile shape: WX wavefronts stacked along x, RX x-rows per lane.
template <int WX_, int RX_>
struct Tile {
  static constexpr int WX = WX_, RX = RX_;
  static constexpr int TX = WX * RX;
  static constexpr int LX = TX + 2 * H, LY = TY + 2 * H;
  static constexpr int NT = TY * WX;
  static constexpr int NXH = (2 * H + WX - 1) / WX;  // x-halo rows per lane
  static_assert(TX * 2 * H <= NT, "one y-halo value per lane");
};

template <class P>
__device__ __forceinline__ P *at(P *p, unsigned byte_off)
{
  return reinterpret_cast<P *>(reinterpret_cast<uintptr_t>(p) + (uintptr_t)byte_off);
}

// Storage policies for wavefields and PML state. Arithmetic is always float.
struct StoreF32 {
  using T = float;
  static __device__ __forceinline__ float ld(const T *p, unsigned i) { return *at(p, i << 2); }
  static __device__ __forceinline__ float ld_stream(const T *p, unsigned i)
  {
    return __builtin_nontemporal_load(at(p, i << 2));
  }
  static __device__ __forceinline__ void st(T *p, unsigned i, float v) { __builtin_nontemporal_store(v, at(p, i << 2)); }
  static __device__ __forceinline__ float rnd(float v) { return v; }
};

__device__ __forceinline__ float ld_model(const float *p, unsigned i) { return __builtin_nontemporal_load(at(p, i << 2)); }

struct Geo {
  int nx, ny, nz, pml;
  unsigned pln;  // nx*ny
};

__device__ __forceinline__ unsigned cell(const Geo &g, int z, int x, int y)
{
  return (unsigned)z * g.pln + (unsigned)x * (unsigned)g.ny + (unsigned)y;
}

// Work item of one workgroup: x in [x0,x1), y in [y0,y1), z in [z0,z1); x1-x0 <= TX, y1-y0 <= 64.
// Packed as {x0 | x1<<16, y0 | y1<<16, z0, z1}; z0 == z1 marks padding.
__device__ __forceinline__ void unpack(int4 w, int &x0, int &x1, int &y0, int &y1, int &z0, int &z1)
{
  x0 = w.x & 0xffff; x1 = (unsigned)w.x >> 16;
  y0 = w.y & 0xffff; y1 = (unsigned)w.y >> 16;
  z0 = w.z; z1 = w.w;
}

struct Coef {
  float c1[H], c2[H];        // 10th-order rows of the FD tables (entries 31..35)
  float ix, iy, iz;          // 1/dx, 1/dy, 1/dz
  float ixx, iyy, izz;       // 1/dx^2, 1/dy^2, 1/dz^2
  float dt;
};

// Parameters of one k_update launch.
template <class S>
struct UpdArgs {
  Geo g;
  Coef c;
  const typename S::T *ucur;   // u at step k (u2)
  typename S::T *unext;        // in: u at step k-1 (u1); out: u at step k+1 (u3)
  const float *vdt2, *eps, *epsd;           // (vp*dt)^2, epsilon, epsilon-delta
  typename S::T *w1p, *w3p, *pp;            // PML slots k-1 (overwritten with the new value)
  const typename S::T *w1c, *w3c, *pc;      // PML slots k
  typename S::T *w21;                       // PML accumulator (updated in place)
  const float4 *pmltab;                     // [3][pml]: K1, K2, K3, E per face family and layer
  const unsigned *gmax;                     // G of this step (float bits)
  float wk;                                 // wavelet[k]
  int src_on;
  int sb_z0, sb_z1, sb_x0, sb_x1, sb_y0, sb_y1;  // source box (inclusive bounds)
  const float *sbox;                        // dense source values over the box
  float *rec;                               // record row of this step, [hx][hy]
  int sz;                                   // receiver plane
  const int4 *work;                         // one item per workgroup
};

template <class S>
struct GradArgs {
  Geo g;
  Coef c;
  const typename S::T *ucur;
  unsigned *gmax;   // out: atomic max of the gradient components (float bits)
  int *flag;        // out: set to 1 when a non-finite gradient is seen
  const int4 *work;
};

// Global loads of u for one plane of a tile: the lane's queue front (plane z+5) and its share
// of the x and y halos of plane z.
template <class S, class T>
struct PlaneLoad {
  float qn[T::RX];
  float hx[T::NXH];
  float hy;

  __device__ __forceinline__ void issue(const typename S::T *u, const Geo &g, int z, int x0, int y0, int tx, int ty)
  {
    const int y = y0 + ty;
    const bool yin = y < g.ny;
#pragma unroll
    for (int r = 0; r < T::RX; r++) {
      const int x = x0 + tx + T::WX * r;
      qn[r] = (yin && x < g.nx) ? S::ld(u, cell(g, z + H, x, y)) : 0.f;
    }
#pragma unroll
    for (int i = 0; i < T::NXH; i++) {
      const int h = tx + T::WX * i;
      const int X = h < H ? h : T::TX + h;
      const int gx = x0 - H + X;
      hx[i] = (h < 2 * H && yin && gx < g.nx) ? S::ld(u, cell(g, z, gx, y)) : 0.f;
    }
    const int t = tx * TY + ty;
    hy = 0.f;
    if (t < T::TX * 2 * H) {
      const int row = t / (2 * H), col = t - row * (2 * H);
      const int gy = y0 - H + (col < H ? col : TY + col), gx = x0 + row;
      if (gy < g.ny && gx < g.nx) hy = S::ld(u, cell(g, z, gx, gy));
    }
  }

  // Write plane z (centre values from the queue, halos from this load) into an LDS buffer.
  __device__ __forceinline__ void stage(float (*sm)[T::LY], const float (&q)[T::RX][2 * H + 1], int tx, int ty) const
  {
#pragma unroll
    for (int r = 0; r < T::RX; r++) sm[H + tx + T::WX * r][H + ty] = q[r][H];
#pragma unroll
    for (int i = 0; i < T::NXH; i++) {
      const int h = tx + T::WX * i;
      if (h < 2 * H) sm[h < H ? h : T::TX + h][H + ty] = hx[i];
    }
    const int t = tx * TY + ty;
    if (t < T::TX * 2 * H) {
      const int row = t / (2 * H), col = t - row * (2 * H);
      sm[H + row][col < H ? col : TY + col] = hy;
    }
  }
};

template <class S, class T>
__device__ __forceinline__ void prime_queue(float (&q)[T::RX][2 * H + 1], const typename S::T *u, const Geo &g, int z0,
                                            int x0, int y, int tx)
{
#pragma unroll
  for (int r = 0; r < T::RX; r++) {
    const int x = x0 + tx + T::WX * r;
#pragma unroll
    for (int j = 0; j < 2 * H; j++) q[r][j] = (y < g.ny && x < g.nx) ? S::ld(u, cell(g, z0 - H + j, x, y)) : 0.f;
  }
}

template <class T>
__device__ __forceinline__ void shift_queue(float (&q)[T::RX][2 * H + 1], const float (&qn)[T::RX])
{
#pragma unroll
  for (int r = 0; r < T::RX; r++) {
#pragma unroll
    for (int j = 0; j < 2 * H; j++) q[r][j] = q[r][j + 1];
    q[r][2 * H] = qn[r];
  }
}

template <class S, class T>
__global__ __launch_bounds__(T::NT) void k_gradmax(GradArgs<S> a)
{
  __shared__ float sm[2][T::LX][T::LY];
  __shared__ float red[T::WX];
  const Geo &g = a.g;
  const int ty = threadIdx.x, tx = threadIdx.y;
  int x0, x1, y0, y1, z0, z1;
  unpack(a.work[blockIdx.x], x0, x1, y0, y1, z0, z1);
  if (z0 >= z1) return;
  const int y = y0 + ty;
  const bool ycmp = y < y1;
  float q[T::RX][2 * H + 1];
  prime_queue<S, T>(q, a.ucur, g, z0, x0, y, tx);
  PlaneLoad<S, T> L;
  L.issue(a.ucur, g, z0, x0, y0, tx, ty);
#pragma unroll
  for (int r = 0; r < T::RX; r++) q[r][2 * H] = L.qn[r];
  L.stage(sm[0], q, tx, ty);
  __syncthreads();
  float m = 0.f;
  int bad = 0;
  int b = 0;
  for (int z = z0; z < z1; z++) {
    const bool more = z + 1 < z1;
    if (more) L.issue(a.ucur, g, z + 1, x0, y0, tx, ty);
    float (*s)[T::LY] = sm[b];
#pragma unroll
    for (int r = 0; r < T::RX; r++) {
      const int X = H + tx + T::WX * r, Y = H + ty;
      const int x = x0 + tx + T::WX * r;
      float ux = 0.f, uy = 0.f, uz = 0.f;
#pragma unroll
      for (int k = 1; k <= H; k++) {
        ux = fmaf(a.c.c1[k - 1], s[X + k][Y] - s[X - k][Y], ux);
        uy = fmaf(a.c.c1[k - 1], s[X][Y + k] - s[X][Y - k], uy);
        uz = fmaf(a.c.c1[k - 1], q[r][H + k] - q[r][H - k], uz);
      }
      if (ycmp && x < x1) {
        const float ax = fabsf(ux * a.c.ix), ay = fabsf(uy * a.c.iy), az = fabsf(uz * a.c.iz);
        m = (ax >= m) ? ax : m;
        m = (ay >= m) ? ay : m;
        m = (az >= m) ? az : m;
        bad |= !(ax <= 3.4e38f) | !(ay <= 3.4e38f) | !(az <= 3.4e38f);
      }
    }
    if (more) {
      shift_queue<T>(q, L.qn);
      L.stage(sm[b ^ 1], q, tx, ty);
      __syncthreads();
      b ^= 1;
    }
  }
  // workgroup max -> one atomic per workgroup
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const float o = __shfl_xor(m, off, 64);
    m = (o >= m) ? o : m;
  }
  if (ty == 0) red[tx] = m;
  __syncthreads();
  if (tx == 0 && ty == 0) {
    float v = red[0];
#pragma unroll
    for (int w = 1; w < T::WX; w++) v = (red[w] >= v) ? red[w] : v;
    atomicMax(a.gmax, __float_as_uint(v));
  }
  if (bad) atomicOr(a.flag, 1);
}

// Per-cell inputs of k_update other than u_cur: model, and either u_prev (interior cells)
// or the seven PML state values (PML cells).
template <class S, class T, bool PML>
struct CellLoad {
  static constexpr int NS = PML ? 7 : 1;
  float v2[T::RX], e[T::RX], ed[T::RX];
  float st[T::RX][NS];  // interior: st[0] = u_prev; PML: w1p w1c pp pc w3p w3c w21

  __device__ __forceinline__ void issue(const UpdArgs<S> &a, int z, int x0, int x1, int y, bool ycmp, int tx)
  {
    const Geo &g = a.g;
#pragma unroll
    for (int r = 0; r < T::RX; r++) {
      const int x = x0 + tx + T::WX * r;
      const unsigned i = cell(g, z, x, y);
      const bool cmp = ycmp && x < x1;
      v2[r] = cmp ? ld_model(a.vdt2, i) : 0.f;
      e[r] = cmp ? ld_model(a.eps, i) : 0.f;
      ed[r] = cmp ? ld_model(a.epsd, i) : 0.f;
      if constexpr (!PML) {
        st[r][0] = cmp ? S::ld_stream(a.unext, i) : 0.f;
      } else {
        const bool inner = x >= g.pml && x < g.nx - g.pml && y >= g.pml && y < g.ny - g.pml && z >= g.pml &&
                           z < g.nz - g.pml;
        if (cmp && inner) {
          st[r][0] = S::ld_stream(a.unext, i);
#pragma unroll
          for (int j = 1; j < NS; j++) st[r][j] = 0.f;
        } else if (cmp) {
          st[r][0] = S::ld_stream(a.w1p, i);
          st[r][1] = S::ld_stream(a.w1c, i);
          st[r][2] = S::ld_stream(a.pp, i);
          st[r][3] = S::ld_stream(a.pc, i);
          st[r][4] = S::ld_stream(a.w3p, i);
          st[r][5] = S::ld_stream(a.w3c, i);
          st[r][6] = S::ld_stream(a.w21, i);
        } else {
#pragma unroll
          for (int j = 0; j < NS; j++) st[r][j] = 0.f;
        }
      }
    }
  }
};

template <class S, class T, bool PML>
__global__ __launch_bounds__(T::NT) void k_update(UpdArgs<S> a)
{
  __shared__ float sm[2][T::LX][T::LY];
  const Geo &g = a.g;
  const int ty = threadIdx.x, tx = threadIdx.y;
  int x0, x1, y0, y1, z0, z1;
  unpack(a.work[blockIdx.x], x0, x1, y0, y1, z0, z1);
  if (z0 >= z1) return;
  const int y = y0 + ty;
  const bool ycmp = y < y1;
  const int nx = g.nx, ny = g.ny, nz = g.nz, pml = g.pml;
  const bool yint = y >= pml && y < ny - pml;

  // S threshold: |grad u| > 0.01*G compared in double (GPU_kernel.cu:99); for a float
  // magnitude this equals mode > rd(0.01*G).
  const float G = __uint_as_float(*a.gmax);
  const float thr = __double2float_rd(1.0e-02 * (double)G);

  float q[T::RX][2 * H + 1];
  prime_queue<S, T>(q, a.ucur, g, z0, x0, y, tx);
  PlaneLoad<S, T> L;
  L.issue(a.ucur, g, z0, x0, y0, tx, ty);
  CellLoad<S, T, PML> C, Cn;
  C.issue(a, z0, x0, x1, y, ycmp, tx);
#pragma unroll
  for (int r = 0; r < T::RX; r++) q[r][2 * H] = L.qn[r];
  L.stage(sm[0], q, tx, ty);
  __syncthreads();
  int b = 0;
  for (int z = z0; z < z1; z++) {
    const bool more = z + 1 < z1;
    if (more) {
      L.issue(a.ucur, g, z + 1, x0, y0, tx, ty);
      Cn.issue(a, z + 1, x0, x1, y, ycmp, tx);
    }
    float (*s)[T::LY] = sm[b];
    const bool zint = z >= pml && z < nz - pml;
#pragma unroll
    for (int r = 0; r < T::RX; r++) {
      const int X = H + tx + T::WX * r, Y = H + ty;
      const int x = x0 + tx + T::WX * r;
      if (!(ycmp && x < x1)) continue;
      const unsigned i = cell(g, z, x, y);
      const float c0 = q[r][H];
      const float c2 = c0 + c0;
      float Uxx = 0.f, Uyy = 0.f, Uzz = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
#pragma unroll
      for (int k = 1; k <= H; k++) {
        const float ax = s[X + k][Y], bx = s[X - k][Y];
        const float ay = s[X][Y + k], by = s[X][Y - k];
        const float az = q[r][H + k], bz = q[r][H - k];
        Uxx = fmaf(a.c.c2[k - 1], (ax - c2) + bx, Uxx);
        Uyy = fmaf(a.c.c2[k - 1], (ay - c2) + by, Uyy);
        Uzz = fmaf(a.c.c2[k - 1], (az - c2) + bz, Uzz);
        ux = fmaf(a.c.c1[k - 1], ax - bx, ux);
        uy = fmaf(a.c.c1[k - 1], ay - by, uy);
        uz = fmaf(a.c.c1[k - 1], az - bz, uz);
      }
      Uxx *= a.c.ixx; Uyy *= a.c.iyy; Uzz *= a.c.izz;
      ux *= a.c.ix; uy *= a.c.iy; uz *= a.c.iz;
      const float e = C.e[r], ed = C.ed[r], v2 = C.v2[r];
      // scalar operator (GPU_kernel.cu:93-111)
      float Sv = 1.f;
      const float mode = sqrtf(ux * ux + uy * uy + uz * uz);
      if (mode > thr) {
        const float inv = 1.f / mode;
        const float nxv = ux * inv, nyv = uy * inv, nzv = uz * inv;
        const float qa = nxv * nxv + nyv * nyv;
        const float den = 1.f + 2.f * e * qa;
        Sv = 0.5f * (1.f + sqrtf(1.f - (8.f * ed * qa * nzv * nzv) / (den * den)));
      }
      const float e1 = 1.f + 2.f * e;
      float un;
      const bool xint = x >= pml && x < nx - pml;
      if (!PML || (xint && yint && zint)) {
        // GPU_kernel.cu:558-565
        un = (2.f * c0 - C.st[r][0]) + ((e1 * (Uxx + Uyy) + Uzz) * Sv) * v2;
      } else if constexpr (PML) {
        // GPU_kernel.cu:566-637; x faces own edges and corners, then y, then z.
        int fam, j;
        float sgn;
        if (!xint) {
          fam = 0; j = x < pml ? pml - x : x - nx + pml; sgn = x < pml ? 1.f : -1.f;
        } else if (!yint) {
          fam = 1; j = y < pml ? pml - y : y - ny + pml; sgn = y < pml ? 1.f : -1.f;
        } else {
          fam = 2; j = z < pml ? pml - z : z - nz + pml; sgn = z < pml ? 1.f : -1.f;
        }
        const float4 t = a.pmltab[fam * pml + j];
        const float B = Sv * v2;
        float Cn_, Unn, Tg, gn;
        if (fam == 0) { Cn_ = e1 * B; Unn = Uxx; Tg = Uzz + e1 * Uyy; gn = ux; }
        else if (fam == 1) { Cn_ = e1 * B; Unn = Uyy; Tg = Uzz + e1 * Uxx; gn = uy; }
        else { Cn_ = B; Unn = Uzz; Tg = e1 * (Uxx + Uyy); gn = uz; }
        const float w1p = C.st[r][0], w1c = C.st[r][1], pp = C.st[r][2], pc = C.st[r][3];
        const float w3p = C.st[r][4], w3c = C.st[r][5], w21 = C.st[r][6];
        const float w1n = S::rnd(Cn_ * Unn - t.x * w1c + t.y * w1p);
        const float pn = S::rnd(-t.x * pc + t.y * pp + sgn * (Cn_ * t.w * gn));
        const float w21n = S::rnd(a.c.dt * pn + w21 * t.z);
        const float w3n = S::rnd(B * Tg + 2.f * w3c - w3p);
        S::st(a.w1p, i, w1n);
        S::st(a.pp, i, pn);
        S::st(a.w21, i, w21n);
        S::st(a.w3p, i, w3n);
        un = w1n + w21n + w3n;
      }
      if (a.src_on && z >= a.sb_z0 && z <= a.sb_z1 && x >= a.sb_x0 && x <= a.sb_x1 && y >= a.sb_y0 &&
          y <= a.sb_y1) {
        const int bx = a.sb_x1 - a.sb_x0 + 1, by = a.sb_y1 - a.sb_y0 + 1;
        const float sv = a.sbox[((long)(z - a.sb_z0) * bx + (x - a.sb_x0)) * by + (y - a.sb_y0)];
        un = fmaf(sv, a.wk, un);  // u3 += source*wavelet, contracted as the CUDA build does
      }
      un = S::rnd(un);
      S::st(a.unext, i, un);
      if (z == a.sz && xint && yint) a.rec[(x - pml) * (ny - 2 * pml) + (y - pml)] = un;
    }
    if (more) {
      shift_queue<T>(q, L.qn);
      L.stage(sm[b ^ 1], q, tx, ty);
      C = Cn;
      __syncthreads();
      b ^= 1;
    }
  }
}

// Source cells outside [5,n-5)^3 are never updated by the stencil; the reference's u3 keeps
// accumulating source*wavelet there (GPU_kernel.cu:720-731 with the copy rotation of
// GPU_kernel.cu:749-750), i.e. u_next = u_cur + s*w.
template <class S>
__global__ void k_rim_source(const typename S::T *ucur, typename S::T *unext, Geo g, int z0, int z1, int x0, int x1,
                             int y0, int y1, const float *sbox, float wk)
{
  const int bx = x1 - x0 + 1, by = y1 - y0 + 1, bz = z1 - z0 + 1;
  const long n = (long)bx * by * bz;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int yy = (int)(t % by), xx = (int)((t / by) % bx), zz = (int)(t / ((long)by * bx));
    const int x = x0 + xx, y = y0 + yy, z = z0 + zz;
    const bool r5 = x >= H && x < g.nx - H && y >= H && y < g.ny - H && z >= H && z < g.nz - H;
    if (r5) continue;
    const unsigned i = cell(g, z, x, y);
    const float v = fmaf(sbox[t], wk, S::ld(ucur, i));
    S::st(unext, i, S::rnd(v));
  }
}

// Interior cube [pml, n-pml)^3 of u into a dense [hz][hx][hy] float buffer
// (3dPureqPModeling_single_shot.cu:347-350).
template <class S>
__global__ void k_snapshot(const typename S::T *u, float *out, Geo g)
{
  const int hx = g.nx - 2 * g.pml, hy = g.ny - 2 * g.pml;
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  const int x = blockIdx.y, z = blockIdx.z;
  if (y >= hy) return;
  out[((long)z * hx + x) * hy + y] = S::ld(u, cell(g, z + g.pml, x + g.pml, y + g.pml));
}

// Tile shapes the plan can launch (index = qp3d_config.tile_update / tile_grad - 1).
#define QP3D_TILES(X) X(8, 2) X(16, 1) X(4, 4) X(8, 4) X(16, 2) X(4, 2)
constexpr int NTILES = 6;

}  // namespace qp3d
