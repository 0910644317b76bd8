// gfx950 kernels of the qP3D time step.
//
// One time step of the reference (3dPureqPModeling_single_shot.cu:277-337) launches five
// per-cell kernels and copies 16 B/cell over PCIe. Here a step is two stencil passes over
// the wavefield, both hand-tiled for CDNA4:
//   k_gradmax  max over the grid of |du/dx|,|du/dy|,|du/dz| (GPU_kernel.cu:14-33 grad and the
//              host absMaxval at 3dPureqPModeling_single_shot.cu:284-292), reduced on device;
//   k_update   gradient (recomputed in registers), scalar operator S (GPU_kernel.cu:87-115),
//              10th-order second derivatives, interior leapfrog or split-field PML update
//              (GPU_kernel.cu:534-639), source injection (GPU_kernel.cu:720-731) and the
//              receiver-plane write (3dPureqPModeling_single_shot.cu:333-337), in one pass.
// The copy-rotation kernel (GPU_kernel.cu:733-812) disappears: u_next and the PML state
// are written in place of the oldest slot (each cell reads that slot only at its own index)
// and the host flips slot parity each step.
//
// Tiling: a workgroup owns a TY x TX column of the xy plane (y unit-stride, one 64-lane
// wavefront per x row) and marches through a chunk of z planes. The current plane with a
// 5-cell halo sits in LDS for the x and y derivatives; each lane keeps an 11-deep register
// queue of its column in z for the z derivatives, so every u value is fetched from HBM
// once per pass plus halo reuse that L2/MALL serves.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qp3d {

constexpr int H = 5;          // stencil half-width (10th order)
constexpr int TY = 64;        // tile extent along y = wavefront width
constexpr int WX = 4;         // wavefronts per workgroup, stacked along x
constexpr int RX = 4;         // x rows per lane
constexpr int TX = WX * RX;   // tile extent along x
constexpr int LX = TX + 2 * H;
constexpr int LY = TY + 2 * H;
constexpr int NTHREADS = TY * WX;

// Storage policies for wavefields and PML state. Arithmetic is always float.
struct StoreF32 {
  using T = float;
  static __device__ __forceinline__ float ld(const T *p, long i) { return p[i]; }
  static __device__ __forceinline__ void st(T *p, long i, float v) { p[i] = v; }
  static __device__ __forceinline__ float rnd(float v) { return v; }
};

struct Geo {
  int nx, ny, nz, pml;
  long pln;               // nx*ny
  int tiles_y, tiles_x;   // xy tiles over [5,n-5)
  int zchunk;             // z planes per workgroup
  int nlog;               // tiles_y*tiles_x*z chunks; the grid is this rounded up to a multiple of 8
};

struct Coef {
  float cxx[H], cyy[H], czz[H];  // c2[31+m-1] / (d*d)
  float cx[H], cy[H], cz[H];     // c1[31+m-1] / d
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
};

template <class S>
struct GradArgs {
  Geo g;
  Coef c;
  const typename S::T *ucur;
  unsigned *gmax;   // out: atomic max of the gradient components (float bits)
  int *flag;        // out: set to 1 when a non-finite gradient is seen
};

// Map the launch's workgroup id to (tile, zchunk). Workgroups are dispatched round-robin
// over the 8 XCDs; consecutive ids therefore land on different L2s. The remap hands each XCD
// a contiguous range of z-chunk-major tile ids so that the halo rows a neighbour needs are
// read through the same L2.
// Returns false for the padding workgroups past g.nlog.
__device__ __forceinline__ bool tile_of(const Geo &g, int &ty0, int &tx0, int &z0)
{
  const int nblk = gridDim.x;  // multiple of 8
  const int b = blockIdx.x;
  const int l = (b & 7) * (nblk >> 3) + (b >> 3);
  if (l >= g.nlog) return false;
  const int ntile = g.tiles_x * g.tiles_y;
  const int zc = l / ntile;
  const int t = l - zc * ntile;
  const int txi = t / g.tiles_y;
  const int tyi = t - txi * g.tiles_y;
  ty0 = H + tyi * TY;
  tx0 = H + txi * TX;
  z0 = H + zc * g.zchunk;
  return true;
}

// Stage the current plane of u (queue centre + x/y halos) into LDS.
template <class S>
__device__ __forceinline__ void stage_plane(float (*sm)[LY], const typename S::T *u, const Geo &g, float (&q)[RX][2 * H + 1],
                                            int z, int x0, int y0, int tx, int ty)
{
  const int y = y0 + ty;
  const bool yin = y < g.ny;
#pragma unroll
  for (int r = 0; r < RX; r++) sm[H + tx + WX * r][H + ty] = q[r][H];
  const long zb = (long)z * g.pln;
  // x halo: 2H rows of TY lanes
#pragma unroll
  for (int h = tx; h < 2 * H; h += WX) {
    const int X = h < H ? h : TX + h;
    const int gx = x0 - H + X;
    float v = 0.f;
    if (yin && gx < g.nx) v = S::ld(u, zb + (long)gx * g.ny + y);
    sm[X][H + ty] = v;
  }
  // y halo: TX rows of 2H columns
  const int t = tx * TY + ty;
  if (t < TX * 2 * H) {
    const int row = t / (2 * H), col = t - row * (2 * H);
    const int Y = col < H ? col : TY + col;
    const int gy = y0 - H + Y, gx = x0 + row;
    float v = 0.f;
    if (gy < g.ny && gx < g.nx) v = S::ld(u, zb + (long)gx * g.ny + gy);
    sm[H + row][Y] = v;
  }
}

template <class S>
__global__ __launch_bounds__(NTHREADS) void k_gradmax(GradArgs<S> a)
{
  __shared__ float sm[LX][LY];
  __shared__ float red[WX];
  const Geo &g = a.g;
  const int ty = threadIdx.x, tx = threadIdx.y;
  int y0, x0, z0;
  if (!tile_of(g, y0, x0, z0)) return;
  const int z1 = min(z0 + g.zchunk, g.nz - H);
  const int y = y0 + ty;
  const bool yin = y < g.ny, ycmp = y < g.ny - H;
  float q[RX][2 * H + 1];
#pragma unroll
  for (int r = 0; r < RX; r++) {
    const int x = x0 + tx + WX * r;
#pragma unroll
    for (int j = 0; j < 2 * H; j++)
      q[r][j] = (yin && x < g.nx && z0 < z1) ? S::ld(a.ucur, (long)(z0 - H + j) * g.pln + (long)x * g.ny + y) : 0.f;
  }
  float m = 0.f;
  int bad = 0;
  for (int z = z0; z < z1; z++) {
#pragma unroll
    for (int r = 0; r < RX; r++) {
      const int x = x0 + tx + WX * r;
      q[r][2 * H] = (yin && x < g.nx) ? S::ld(a.ucur, (long)(z + H) * g.pln + (long)x * g.ny + y) : 0.f;
    }
    __syncthreads();
    stage_plane<S>(sm, a.ucur, g, q, z, x0, y0, tx, ty);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RX; r++) {
      const int X = H + tx + WX * r, Y = H + ty;
      const int x = x0 + tx + WX * r;
      float ux = 0.f, uy = 0.f, uz = 0.f;
#pragma unroll
      for (int k = 1; k <= H; k++) {
        ux = fmaf(a.c.cx[k - 1], sm[X + k][Y] - sm[X - k][Y], ux);
        uy = fmaf(a.c.cy[k - 1], sm[X][Y + k] - sm[X][Y - k], uy);
        uz = fmaf(a.c.cz[k - 1], q[r][H + k] - q[r][H - k], uz);
      }
      if (ycmp && x < g.nx - H) {
        const float ax = fabsf(ux), ay = fabsf(uy), az = fabsf(uz);
        m = (ax >= m) ? ax : m;
        m = (ay >= m) ? ay : m;
        m = (az >= m) ? az : m;
        bad |= !(ax <= 3.4e38f) | !(ay <= 3.4e38f) | !(az <= 3.4e38f);
      }
    }
#pragma unroll
    for (int r = 0; r < RX; r++)
#pragma unroll
      for (int j = 0; j < 2 * H; j++) q[r][j] = q[r][j + 1];
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
    float b = red[0];
#pragma unroll
    for (int w = 1; w < WX; w++) b = (red[w] >= b) ? red[w] : b;
    atomicMax(a.gmax, __float_as_uint(b));
  }
  if (bad) atomicOr(a.flag, 1);
}

template <class S>
__global__ __launch_bounds__(NTHREADS) void k_update(UpdArgs<S> a)
{
  __shared__ float sm[LX][LY];
  const Geo &g = a.g;
  const int ty = threadIdx.x, tx = threadIdx.y;
  int y0, x0, z0;
  if (!tile_of(g, y0, x0, z0)) return;
  const int z1 = min(z0 + g.zchunk, g.nz - H);
  const int y = y0 + ty;
  const bool yin = y < g.ny, ycmp = y < g.ny - H;
  const int nx = g.nx, ny = g.ny, nz = g.nz, pml = g.pml;

  // S threshold: |grad u| > 0.01*G compared in double (GPU_kernel.cu:99); for a float
  // magnitude this equals mode > rd(0.01*G).
  const float G = __uint_as_float(*a.gmax);
  const float thr = __double2float_rd(1.0e-02 * (double)G);
  const bool yint = y >= pml && y < ny - pml;

  float q[RX][2 * H + 1];
#pragma unroll
  for (int r = 0; r < RX; r++) {
    const int x = x0 + tx + WX * r;
#pragma unroll
    for (int j = 0; j < 2 * H; j++)
      q[r][j] = (yin && x < nx && z0 < z1) ? S::ld(a.ucur, (long)(z0 - H + j) * g.pln + (long)x * ny + y) : 0.f;
  }
  for (int z = z0; z < z1; z++) {
#pragma unroll
    for (int r = 0; r < RX; r++) {
      const int x = x0 + tx + WX * r;
      q[r][2 * H] = (yin && x < nx) ? S::ld(a.ucur, (long)(z + H) * g.pln + (long)x * ny + y) : 0.f;
    }
    __syncthreads();
    stage_plane<S>(sm, a.ucur, g, q, z, x0, y0, tx, ty);
    __syncthreads();
    const bool zint = z >= pml && z < nz - pml;
#pragma unroll
    for (int r = 0; r < RX; r++) {
      const int X = H + tx + WX * r, Y = H + ty;
      const int x = x0 + tx + WX * r;
      if (!(ycmp && x < nx - H)) continue;
      const long i = (long)z * g.pln + (long)x * ny + y;
      const float c0 = q[r][H];
      const float c2 = c0 + c0;
      float Uxx = 0.f, Uyy = 0.f, Uzz = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
#pragma unroll
      for (int k = 1; k <= H; k++) {
        const float ax = sm[X + k][Y], bx = sm[X - k][Y];
        const float ay = sm[X][Y + k], by = sm[X][Y - k];
        const float az = q[r][H + k], bz = q[r][H - k];
        Uxx = fmaf(a.c.cxx[k - 1], (ax - c2) + bx, Uxx);
        Uyy = fmaf(a.c.cyy[k - 1], (ay - c2) + by, Uyy);
        Uzz = fmaf(a.c.czz[k - 1], (az - c2) + bz, Uzz);
        ux = fmaf(a.c.cx[k - 1], ax - bx, ux);
        uy = fmaf(a.c.cy[k - 1], ay - by, uy);
        uz = fmaf(a.c.cz[k - 1], az - bz, uz);
      }
      const float e = a.eps[i];
      const float ed = a.epsd[i];
      const float v2 = a.vdt2[i];
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
      if (xint && yint && zint) {
        // GPU_kernel.cu:558-565
        const float up = S::ld(a.unext, i);
        un = (2.f * c0 - up) + ((e1 * (Uxx + Uyy) + Uzz) * Sv) * v2;
      } else {
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
        float Cn, Unn, Tg, gn;
        if (fam == 0) { Cn = e1 * B; Unn = Uxx; Tg = Uzz + e1 * Uyy; gn = ux; }
        else if (fam == 1) { Cn = e1 * B; Unn = Uyy; Tg = Uzz + e1 * Uxx; gn = uy; }
        else { Cn = B; Unn = Uzz; Tg = e1 * (Uxx + Uyy); gn = uz; }
        const float w1p = S::ld(a.w1p, i), w1c = S::ld(a.w1c, i);
        const float pp = S::ld(a.pp, i), pc = S::ld(a.pc, i);
        const float w3p = S::ld(a.w3p, i), w3c = S::ld(a.w3c, i);
        const float w21 = S::ld(a.w21, i);
        const float w1n = S::rnd(Cn * Unn - t.x * w1c + t.y * w1p);
        const float pn = S::rnd(-t.x * pc + t.y * pp + sgn * (Cn * t.w * gn));
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
        const float s = a.sbox[((long)(z - a.sb_z0) * bx + (x - a.sb_x0)) * by + (y - a.sb_y0)];
        un = fmaf(s, a.wk, un);  // u3 += source*wavelet, contracted as the CUDA build does
      }
      un = S::rnd(un);
      S::st(a.unext, i, un);
      if (z == a.sz && xint && yint) a.rec[(long)(x - pml) * (ny - 2 * pml) + (y - pml)] = un;
    }
#pragma unroll
    for (int r = 0; r < RX; r++)
#pragma unroll
      for (int j = 0; j < 2 * H; j++) q[r][j] = q[r][j + 1];
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
    const long i = (long)z * g.pln + (long)x * g.ny + y;
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
  out[((long)z * hx + x) * hy + y] =
      S::ld(u, (long)(z + g.pml) * g.pln + (long)(x + g.pml) * g.ny + (y + g.pml));
}

}  // namespace qp3d
