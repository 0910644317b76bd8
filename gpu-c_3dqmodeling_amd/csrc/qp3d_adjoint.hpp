// Adjoint-state kernels: the exact transpose of the linearised forward time step of
// qp3d_kernels.hpp, for the gradient of J = 1/2 sum_k |record_k - observed_k|^2 with respect
// to the model ((vp dt)^2, epsilon, epsilon - delta).
//
// Forward step k (per updated cell, g = grad u_k, S = S(eps, eps-del, g/|g|) above the
// 0.01*G threshold and 1 below it):
//   interior  u_{k+1} = 2 u_k - u_{k-1} + v2 S ((1+2eps)(Uxx+Uyy) + Uzz)
//   PML       w1' = Cn Unn - K1 w1 + K2 w1o,  p' = -K1 p + K2 po + sgn Cn E gn,
//             w21' = dt p' + K3 w21,  w3' = B Tg + 2 w3 - w3o,  u_{k+1} = w1' + w21' + w3'
//             (B = S v2; Cn = (1+2eps) B on x/y faces, B on z faces).
// The threshold G and the frozen rim are constants of the linearisation (G is a maximum, whose
// derivative exists almost nowhere; the rim never depends on the model).
//
// One adjoint step is two launches:
//   k_adj_prod   per cell: forward stencils of u_k (as in k_update), the adjoint of the
//                cell's own equations (PML adjoint states in place, the leapfrog terms of the
//                adjoint of u), the model-gradient terms, and six product fields whose
//                transposed stencils carry the adjoint back to u_k: Pxx, Pyy, Pzz (second
//                derivatives) and Qx, Qy, Qz (first derivatives, including the path through
//                S's dependence on the gradient direction);
//   k_adj_apply  per cell: adds the transposed stencils of the six product fields to the
//                adjoint of u_k.
// Adjoint state uses the forward's slot parity: at step k slot k&1 holds the adjoint of the
// newest level (k+1) and receives the partial adjoint of level k-1; slot (k&1)^1 holds the
// partial adjoint of level k and is completed by this step.
#pragma once
#include "qp3d_kernels.hpp"

namespace qp3d {

// Touched map of an adjoint sweep, one byte per block of 8 planes x 8 rows x 32 cells (z, x, y),
// z fastest: t[(bx * nby + by) * nbz + bz] != 0 when the adjoint state of the block (the adjoint
// of u, the PML adjoint states, the product fields) may have been nonzero at some step of the
// sweep. Before the sweep every adjoint array is zero and the receiver plane's blocks are
// marked; k_adj_apply marks the blocks where the completed adjoint of u it writes is nonzero;
// nothing clears a mark during the sweep. The outputs of k_adj_prod at a cell vanish with the
// cell's adjoint of u (and residual) and PML adjoint states, which are zero outside the marked
// blocks, so it leaves those planes out (writing zeros over zeros); k_adj_apply adds zero to the
// cells more than 5 cells from every marked block (the product fields are zero outside them) and
// leaves those planes out too. Both narrow an item's z range only: the adjoint wave enters
// through the whole receiver plane and travels in depth. t == nullptr: off.
struct AdjMap {
  unsigned char *t;
  int nbz, nbx, nby;
};

// Narrows [z0, z1) of a work item to the planes within HALO cells (along z) of a marked block
// among the item's x and y blocks widened by HALO cells (a folded tile: both of its y strips);
// z1 = z0 when there is none. Wave 0 reads the map and hands the range to the other wavefronts
// through LDS: k_adj_apply marks blocks while other workgroups read the map, and every
// wavefront of a workgroup must run the same planes (the march has a barrier per plane). A mark
// read late only widens the range; the marks an item needs were set by earlier launches.
template <int HALO>
__device__ __forceinline__ void adj_trim(const AdjMap &m, const Geo &g, int x0, int x1, int y0, int y1, bool fold,
                                         int &z0, int &z1, int *zr)
{
  if (threadIdx.y == 0) {
    const int xb0 = max(0, x0 - HALO) >> AB_X, xb1 = min(g.nx - 1, x1 - 1 + HALO) >> AB_X;
    const int yh = g.ny - g.pml;  // a folded tile's second strip starts here
    const int ya0 = max(0, y0 - HALO) >> AB_Y, ya1 = min(g.ny - 1, y1 - 1 + HALO) >> AB_Y;
    const int yc0 = max(0, yh - HALO) >> AB_Y, yc1 = min(g.ny - 1, yh + (y1 - y0) - 1 + HALO) >> AB_Y;
    const int zb0 = max(0, z0 - HALO) >> AB_Z, zb1 = min(g.nz - 1, z1 - 1 + HALO) >> AB_Z;
    int first = 0x7fffffff, last = -1;
    for (int zb = zb0; zb <= zb1; zb += TY) {  // uniform
      const int b = zb + (int)__lane_id();
      bool hit = false;
      if (b <= zb1) {
        for (int st = 0; st < (fold ? 2 : 1); st++) {
          const int yb0 = st ? yc0 : ya0, yb1 = st ? yc1 : ya1;
          for (int xb = xb0; xb <= xb1; xb++)
            for (int yb = yb0; yb <= yb1; yb++)
              hit |= m.t[((unsigned)xb * (unsigned)m.nby + (unsigned)yb) * (unsigned)m.nbz + (unsigned)b] != 0;
        }
      }
      const unsigned long long v = __ballot(hit);
      if (v) {
        if (first == 0x7fffffff) first = zb + __builtin_ctzll(v);
        last = zb + 63 - __builtin_clzll(v);
      }
    }
    if (__lane_id() == 0) {
      zr[0] = last < 0 ? z0 : max(z0, (first << AB_Z) - HALO);
      zr[1] = last < 0 ? z0 : min(z1, ((last + 1) << AB_Z) + HALO);
    }
  }
  __syncthreads();
  z0 = zr[0];
  z1 = zr[1];
}

struct AdjArgs {
  Geo g;
  Coef c;
  AdjMap m;
  const float *u;                         // forward u_k
  const float *vdt2, *eps, *epsd;         // model
  const float4 *pmltab;
  const unsigned *gmax;                   // G_k (float bits)
  float *an, *ac;                         // adjoint of u: newest level / current level
  float *o1n, *o1c, *pin, *pic, *o3n, *o3c, *o21;  // PML adjoint states (newest / current)
  float *pxx, *pyy, *pzz, *qx, *qy, *qz;  // product fields
  float *gv2, *ge, *ged;                  // gradient accumulators w.r.t. v2, eps, eps-del
  const float *resid;                     // adjoint source of u_{k+1} on the receiver plane [hx][hy]
  int sz;
  const int4 *work;
};

__device__ __forceinline__ float &at(float *p, size_t zo, unsigned b) { return *bat(p + zo, b); }

// Loads of one plane z of a field that has no z queue: the lane's own rows plus its share of
// the x halo (XH) and the y halo (YH), for fields whose stencil runs along x or y.
template <class T, bool XH = true, bool YH = true>
struct PlaneLoadC {
  float c[T::RX];
  float hx[T::NXH];
  float hy;

  __device__ __forceinline__ void issue(const float *f, const Geo &g, const Lane<T> &L, int z)
  {
    const float *fz = f + (size_t)z * g.pln;
#pragma unroll
    for (int r = 0; r < T::RX; r++) c[r] = *bat(fz, L.qoff[r]);
    if (XH) {
#pragma unroll
      for (int i = 0; i < T::NXH; i++)
        if (__builtin_amdgcn_readfirstlane(L.hx_on[i])) hx[i] = *bat(fz, L.hxo[i]);
    }
    if (YH && L.hy_on) hy = *bat(fz, L.hyo);
  }

  __device__ __forceinline__ void stage(float (*sm)[T::LY], const Lane<T> &L) const
  {
#pragma unroll
    for (int r = 0; r < T::RX; r++) sm[H + L.row(r)][L.Y] = c[r];
    if (XH) {
#pragma unroll
      for (int i = 0; i < T::NXH; i++) {
        const int h = L.tx + T::WX * i;
        if (__builtin_amdgcn_readfirstlane(L.hx_on[i])) sm[h < H ? h : T::TX + h][L.Y] = hx[i];
      }
    }
    if (YH && L.hy_on) {
      const int t = L.tx * TY + L.ty;
      const int row = t / (2 * H), col = t - row * (2 * H);
      sm[H + row][Lane<T>::halo_col(col, L.fold)] = hy;
    }
  }
};

template <class T>
__device__ __forceinline__ void shift_queue(float (&q)[T::RX][QN])
{
#pragma unroll
  for (int r = 0; r < T::RX; r++)
#pragma unroll
    for (int j = 0; j < 2 * H; j++) q[r][j] = q[r][j + 1];
}

// Per-cell inputs of k_adj_prod for one plane, issued a plane ahead of their use: the model,
// the adjoint of u (newest level; current level for interior cells), the PML adjoint states
// and the gradient sums.
template <class T, bool PML>
struct AdjCellLoad {
  static constexpr int NS = PML ? 7 : 0;
  float v2[T::RX], e[T::RX], ed[T::RX], an[T::RX], ac[T::RX];
  float st[T::RX][PML ? 7 : 1];  // o1n o1c pin pic o3n o3c o21
  float gs[T::RX][3];            // gv2 ge ged

  __device__ __forceinline__ void issue(const AdjArgs &a, const Lane<T> &L, int z)
  {
    const size_t zo = (size_t)z * a.g.pln;
#pragma unroll
    for (int r = 0; r < T::RX; r++) {
      const unsigned o = L.off[r];
      v2[r] = *bat(a.vdt2 + zo, o);
      e[r] = *bat(a.eps + zo, o);
      ed[r] = *bat(a.epsd + zo, o);
      an[r] = *bat(a.an + zo, o);
      if constexpr (!PML) ac[r] = *bat(a.ac + zo, o);
      if constexpr (PML) {
        const float *src[7] = {a.o1n, a.o1c, a.pin, a.pic, a.o3n, a.o3c, a.o21};
#pragma unroll
        for (int j = 0; j < 7; j++) st[r][j] = *bat(src[j] + zo, o);
      }
      gs[r][0] = *bat(a.gv2 + zo, o);
      gs[r][1] = *bat(a.ge + zo, o);
      gs[r][2] = *bat(a.ged + zo, o);
    }
  }
};

template <class T, bool PML>
__global__ __launch_bounds__(T::NT) void k_adj_prod(AdjArgs a)
{
  __shared__ float sm[2][T::LX + 1][T::LY];
  const Geo &g = a.g;
  int x0, x1, y0, y1, z0, z1;
  bool fold;
  unpack(a.work[blockIdx.x], x0, x1, y0, y1, z0, z1, fold);
  if (z0 >= z1) return;
  __shared__ int zr[2];
  if (a.m.t) {
    adj_trim<0>(a.m, g, x0, x1, y0, y1, fold, z0, z1, zr);
    if (z0 >= z1) return;
  }
  Lane<T> L(g, x0, x1, y0, y1, fold);
  const int y = L.y;
  const int nx = g.nx, ny = g.ny, nz = g.nz, pml = g.pml;
  const int hy = ny - 2 * pml;
  bool rows_on[T::RX];
#pragma unroll
  for (int r = 0; r < T::RX; r++) rows_on[r] = L.yon && x0 + L.row(r) < x1;
  const float G = __uint_as_float(*a.gmax);
  const float thr = __double2float_rd(1.0e-02 * (double)G);

  int fam = 0;
  float4 trow[T::RX];
  float sgn = 1.f;
  if constexpr (PML) {
    fam = (x1 <= pml || x0 >= nx - pml) ? 0 : ((y1 <= pml || y0 >= ny - pml) ? 1 : 2);
    if (fam == 0) {
      sgn = x1 <= pml ? 1.f : -1.f;
#pragma unroll
      for (int r = 0; r < T::RX; r++) {
        const int x = min(x0 + L.row(r), x1 - 1);
        trow[r] = a.pmltab[x < pml ? pml - x : x - nx + pml];
      }
    } else if (fam == 1) {
      const int yc = L.yc;
      sgn = yc < pml ? 1.f : -1.f;
      const float4 t = a.pmltab[pml + (yc < pml ? pml - yc : yc - ny + pml)];
#pragma unroll
      for (int r = 0; r < T::RX; r++) trow[r] = t;
    } else {
      sgn = z1 <= pml ? 1.f : -1.f;
    }
  }

  float q[T::RX][QN];
  prime_queue<StoreF32, T>(q, a.u, g, L, z0);
  PlaneLoad<StoreF32, T> P;
  P.issue(a.u, g, L, z0);
  float centre[T::RX];
#pragma unroll
  for (int r = 0; r < T::RX; r++) {
    q[r][2 * H] = P.qn[r];
    centre[r] = q[r][H];
  }
  P.stage(sm[0], centre, L);
  AdjCellLoad<T, PML> C, Cn;
  C.issue(a, L, z0);
  __syncthreads();
  int b = 0;
  for (int z = z0; z < z1; z++) {
    const bool more = z + 1 < z1;
    L.refresh();
    if (more) {
      P.issue(a.u, g, L, z + 1);
      Cn.issue(a, L, z + 1);
    }
    float (*s)[T::LY] = sm[b];
    float4 tz = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (PML) {
      if (fam == 2) tz = a.pmltab[2 * pml + (z < pml ? pml - z : z - nz + pml)];
    }
    const size_t zo = (size_t)z * g.pln;
    const bool rec_plane = a.resid && z == a.sz;
#pragma unroll
    for (int r = 0; r < T::RX; r++) {
      if (!rows_on[r]) continue;
      const int X = H + L.row(r), Y = L.Y;
      const int x = x0 + L.row(r);
      const float c0 = q[r][H];
      float Lxx = 0.f, Lyy = 0.f, Lzz = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
      const float c02 = 2.f * c0;
#pragma unroll
      for (int k = 1; k <= H; k++) {
        const float ax = s[X + k][Y], bx = s[X - k][Y];
        const float ay = s[X][Y + k], by = s[X][Y - k];
        const float az = q[r][H + k], bz = q[r][H - k];
        Lxx = d2tap(a.c.c2[k - 1], ax, bx, c02, Lxx);
        Lyy = d2tap(a.c.c2[k - 1], ay, by, c02, Lyy);
        Lzz = d2tap(a.c.c2[k - 1], az, bz, c02, Lzz);
        ux = fmaf(a.c.c1[k - 1], ax - bx, ux);
        uy = fmaf(a.c.c1[k - 1], ay - by, uy);
        uz = fmaf(a.c.c1[k - 1], az - bz, uz);
      }
      const float Uxx = Lxx * a.c.ixx;
      const float Uyy = Lyy * a.c.iyy;
      const float Uzz = Lzz * a.c.izz;
      ux *= a.c.ix; uy *= a.c.iy; uz *= a.c.iz;
      const unsigned o = L.off[r];
      const float e = C.e[r], ed = C.ed[r], v2 = C.v2[r];

      // S and its derivatives (the forward's arithmetic for S itself)
      float Sv = 1.f, dSde = 0.f, dSded = 0.f, dSdqa = 0.f, dSdb = 0.f;
      float nxv = 0.f, nyv = 0.f, nzv = 0.f, inv = 0.f, mode = 0.f;
      bool above = false;
      {
        mode = __builtin_amdgcn_sqrtf(fmaf(ux, ux, fmaf(uy, uy, uz * uz)));
        if (mode > thr) {
          above = true;
          inv = __builtin_amdgcn_rcpf(mode + 2.22e-17f);
          nxv = ux * inv; nyv = uy * inv; nzv = uz * inv;
          const float qa = fmaf(nxv, nxv, nyv * nyv), bq = nzv * nzv;
          const float den = fmaf(2.f * e, qa, 1.f);
          const float rd2 = __builtin_amdgcn_rcpf(den * den), rd3 = rd2 / den;
          const float X8 = 8.f * ed * qa * bq * rd2;
          const float R = __builtin_amdgcn_sqrtf(1.f - X8);
          Sv = 0.5f + 0.5f * R;
          const float dSdX = -0.25f / fmaxf(R, 1e-6f);
          dSde = dSdX * (-32.f * ed * qa * qa * bq * rd3);
          dSded = dSdX * (8.f * qa * bq * rd2);
          dSdqa = dSdX * (8.f * ed * bq * (den - 4.f * e * qa) * rd3);
          dSdb = dSdX * (8.f * ed * qa * rd2);
        }
      }
      const float e1 = fmaf(2.f, e, 1.f);
      float mu = C.an[r];
      if (rec_plane && x >= pml && x < nx - pml && y >= pml && y < ny - pml)
        mu += a.resid[(x - pml) * hy + (y - pml)];

      float Pxx, Pyy, Pzz, Qx = 0.f, Qy = 0.f, Qz = 0.f, dS, gv2, ge;
      if constexpr (!PML) {
        const float Pt = fmaf(e1, Uxx + Uyy, Uzz);
        const float w = mu * v2 * Sv;
        Pxx = w * e1; Pyy = Pxx; Pzz = w;
        dS = mu * v2 * Pt;
        gv2 = mu * Sv * Pt;
        ge = 2.f * w * (Uxx + Uyy);
        at(a.ac, zo, o) = C.ac[r] + 2.f * mu;
        at(a.an, zo, o) = -mu;
      } else {
        const float4 t = fam == 2 ? tz : trow[r];
        const float B = Sv * v2;
        const float w1b = C.st[r][0] + mu;
        const float w21b = C.st[r][6] + mu;
        const float w3b = C.st[r][4] + mu;
        const float pb = fmaf(a.c.dt, w21b, C.st[r][2]);
        at(a.o1c, zo, o) = fmaf(-t.x, w1b, C.st[r][1]);
        at(a.o1n, zo, o) = t.y * w1b;
        at(a.pic, zo, o) = fmaf(-t.x, pb, C.st[r][3]);
        at(a.pin, zo, o) = t.y * pb;
        at(a.o21, zo, o) = t.z * w21b;
        at(a.o3c, zo, o) = fmaf(2.f, w3b, C.st[r][5]);
        at(a.o3n, zo, o) = -w3b;
        at(a.an, zo, o) = 0.f;  // u_{k-1} is not read by PML cells
        float dCn, dB, de1;
        if (fam == 0) {
          const float Cn = e1 * B;
          Pxx = Cn * w1b; Pyy = e1 * B * w3b; Pzz = B * w3b; Qx = sgn * Cn * t.w * pb;
          dCn = fmaf(w1b, Uxx, pb * sgn * t.w * ux);
          dB = w3b * fmaf(e1, Uyy, Uzz) + e1 * dCn;
          de1 = w3b * B * Uyy + B * dCn;
        } else if (fam == 1) {
          const float Cn = e1 * B;
          Pyy = Cn * w1b; Pxx = e1 * B * w3b; Pzz = B * w3b; Qy = sgn * Cn * t.w * pb;
          dCn = fmaf(w1b, Uyy, pb * sgn * t.w * uy);
          dB = w3b * fmaf(e1, Uxx, Uzz) + e1 * dCn;
          de1 = w3b * B * Uxx + B * dCn;
        } else {
          Pzz = B * w1b; Pxx = e1 * B * w3b; Pyy = Pxx; Qz = sgn * B * t.w * pb;
          dCn = fmaf(w1b, Uzz, pb * sgn * t.w * uz);
          dB = w3b * e1 * (Uxx + Uyy) + dCn;
          de1 = w3b * B * (Uxx + Uyy);
        }
        dS = v2 * dB;
        gv2 = Sv * dB;
        ge = 2.f * de1;
      }
      float ged = 0.f;
      if (above) {
        ge = fmaf(dS, dSde, ge);
        ged = dS * dSded;
        // adjoint of n = g/(|g|+r), then of g
        const float vx = dS * dSdqa * 2.f * nxv, vy = dS * dSdqa * 2.f * nyv, vz = dS * dSdb * 2.f * nzv;
        const float cdot = (vx * nxv + vy * nyv + vz * nzv) * inv / mode;
        Qx += vx * inv - cdot * ux;
        Qy += vy * inv - cdot * uy;
        Qz += vz * inv - cdot * uz;
      }
      at(a.pxx, zo, o) = Pxx; at(a.pyy, zo, o) = Pyy; at(a.pzz, zo, o) = Pzz;
      at(a.qx, zo, o) = Qx; at(a.qy, zo, o) = Qy; at(a.qz, zo, o) = Qz;
      at(a.gv2, zo, o) = C.gs[r][0] + gv2;
      at(a.ge, zo, o) = C.gs[r][1] + ge;
      at(a.ged, zo, o) = C.gs[r][2] + ged;
    }
    if (more) {
      C = Cn;
      shift_queue<T>(q);
#pragma unroll
      for (int r = 0; r < T::RX; r++) {
        q[r][2 * H] = P.qn[r];
        centre[r] = q[r][H];
      }
      P.stage(sm[b ^ 1], centre, L);
      __syncthreads();
      b ^= 1;
    }
  }
}

struct ApplyArgs {
  Geo g;
  Coef c;
  const float *pxx, *pyy, *pzz, *qx, *qy, *qz;
  float *ac;
  AdjMap m;  // marked where the completed adjoint of u is nonzero
  const int4 *work;
};

// ac += Uxx^T Pxx + Uyy^T Pyy + Uzz^T Pzz + Dx^T Qx + Dy^T Qy + Dz^T Qz over [5,n-5)^3. The
// second-derivative stencils are symmetric; the first-derivative ones change sign. Product
// fields are zero outside [5,n-5)^3.
template <class T>
__global__ __launch_bounds__(T::NT) void k_adj_apply(ApplyArgs a)
{
  // Pxx, Qx, Pyy, Qy of one plane: a single buffer (a second barrier per plane) halves the
  // workgroup's LDS, so that more workgroups stay resident on a CU
  __shared__ float sm[1][4][T::LX + 1][T::LY];
  const Geo &g = a.g;
  int x0, x1, y0, y1, z0, z1;
  bool fold;
  unpack(a.work[blockIdx.x], x0, x1, y0, y1, z0, z1, fold);
  if (z0 >= z1) return;
  __shared__ int zr[2];
  const bool map_on = a.m.t != nullptr;
  if (map_on) {
    adj_trim<H>(a.m, g, x0, x1, y0, y1, fold, z0, z1, zr);
    if (z0 >= z1) return;
  }
  Lane<T> L(g, x0, x1, y0, y1, fold);
  bool rows_on[T::RX];
#pragma unroll
  for (int r = 0; r < T::RX; r++) rows_on[r] = L.yon && x0 + L.row(r) < x1;
  bool nzl = false;  // a nonzero adjoint of u written in the current z block (map on)
  // marks z block zb for the tile's x and y blocks (both strips of a folded tile)
  auto mark = [&](int zb) {
    const int xb0 = x0 >> AB_X, xb1 = (x1 - 1) >> AB_X;
    const int yh = g.ny - g.pml;
    for (int st = 0; st < (fold ? 2 : 1); st++) {
      const int yb0 = (st ? yh : y0) >> AB_Y, yb1 = ((st ? yh : y0) + (y1 - y0) - 1) >> AB_Y;
      for (int xb = xb0; xb <= xb1; xb++)
        for (int yb = yb0; yb <= yb1; yb++)
          a.m.t[((unsigned)xb * (unsigned)a.m.nby + (unsigned)yb) * (unsigned)a.m.nbz + (unsigned)zb] = 1;
    }
  };
  float qzz[T::RX][QN], qqz[T::RX][QN];
  prime_queue<StoreF32, T>(qzz, a.pzz, g, L, z0);
  prime_queue<StoreF32, T>(qqz, a.qz, g, L, z0);
  // Pxx and Qx need only their x halos, Pyy and Qy only their y halos.
  PlaneLoadC<T, true, false> Pxx, Pqx;
  PlaneLoadC<T, false, true> Pyy, Pqy;
  float fzz[T::RX], fqz[T::RX];
  float acn[T::RX], acv[T::RX];  // the adjoint of u at the cells, loaded a plane ahead
  auto issue = [&](int z) {
    Pxx.issue(a.pxx, g, L, z);
    Pqx.issue(a.qx, g, L, z);
    Pyy.issue(a.pyy, g, L, z);
    Pqy.issue(a.qy, g, L, z);
    const float *uzz = a.pzz + (size_t)(z + H) * g.pln, *uqz = a.qz + (size_t)(z + H) * g.pln;
    const size_t zo = (size_t)z * g.pln;
#pragma unroll
    for (int r = 0; r < T::RX; r++) {
      fzz[r] = *bat(uzz, L.qoff[r]);
      fqz[r] = *bat(uqz, L.qoff[r]);
      acn[r] = at(a.ac, zo, L.off[r]);
    }
  };
  auto stage = [&](int buf) {
    Pxx.stage(sm[buf][0], L);
    Pqx.stage(sm[buf][1], L);
    Pyy.stage(sm[buf][2], L);
    Pqy.stage(sm[buf][3], L);
  };
  issue(z0);
#pragma unroll
  for (int r = 0; r < T::RX; r++) {
    qzz[r][2 * H] = fzz[r];
    qqz[r][2 * H] = fqz[r];
    acv[r] = acn[r];
  }
  stage(0);
  __syncthreads();
  // The march is unrolled by 22 (twice the queue length): queue slots and the LDS buffer of
  // every plane are static (the value of plane z-5+j sits in slot (j + it) % 11).
  int zb = z0;
  auto plane = [&](auto itc) {
    constexpr int it = decltype(itc)::value;
    const int z = zb + it;
    const bool more = z + 1 < z1;
    L.refresh();
    if (more) issue(z + 1);
    const size_t zo = (size_t)z * g.pln;
    float (*sxx)[T::LY] = sm[0][0], (*sqx)[T::LY] = sm[0][1], (*syy)[T::LY] = sm[0][2], (*sqy)[T::LY] = sm[0][3];
#pragma unroll
    for (int r = 0; r < T::RX; r++) {
      if (!rows_on[r]) continue;
      const int X = H + L.row(r), Y = L.Y;
      float lxx = 0.f, lyy = 0.f, lzz = 0.f, dx = 0.f, dy = 0.f, dz = 0.f;
      const float px2 = 2.f * sxx[X][Y], py2 = 2.f * syy[X][Y], pz2 = 2.f * qzz[r][(H + it) % QN];
#pragma unroll
      for (int k = 1; k <= H; k++) {
        lxx = d2tap(a.c.c2[k - 1], sxx[X + k][Y], sxx[X - k][Y], px2, lxx);
        lyy = d2tap(a.c.c2[k - 1], syy[X][Y + k], syy[X][Y - k], py2, lyy);
        lzz = d2tap(a.c.c2[k - 1], qzz[r][(H + k + it) % QN], qzz[r][(H - k + it) % QN], pz2, lzz);
        dx = fmaf(a.c.c1[k - 1], sqx[X - k][Y] - sqx[X + k][Y], dx);
        dy = fmaf(a.c.c1[k - 1], sqy[X][Y - k] - sqy[X][Y + k], dy);
        dz = fmaf(a.c.c1[k - 1], qqz[r][(H - k + it) % QN] - qqz[r][(H + k + it) % QN], dz);
      }
      float acc = lxx * a.c.ixx;
      acc = fmaf(lyy, a.c.iyy, acc);
      acc = fmaf(lzz, a.c.izz, acc);
      acc = fmaf(dx, a.c.ix, acc);
      acc = fmaf(dy, a.c.iy, acc);
      acc = fmaf(dz, a.c.iz, acc);
      const float lam = acv[r] + acc;
      at(a.ac, zo, L.off[r]) = lam;
      nzl |= lam != 0.f;
    }
    if (map_on && ((z & ((1 << AB_Z) - 1)) == (1 << AB_Z) - 1 || !more)) {  // uniform
      if (__ballot(nzl) && __lane_id() == 0) mark(z >> AB_Z);
      nzl = false;
    }
    if (more) {
#pragma unroll
      for (int r = 0; r < T::RX; r++) {
        qzz[r][it % QN] = fzz[r];
        qqz[r][it % QN] = fqz[r];
        acv[r] = acn[r];
      }
      __syncthreads();  // every wavefront is done with plane z
      stage(0);
      __syncthreads();
    }
  };
  for (; zb < z1; zb += 2 * QN) Unroll<0, 2 * QN>::run(plane, z1 - zb);
}

// gsum += gacc over n cells (FP64 accumulation over time of the per-segment FP32 sums).
__global__ void k_fold(double *gsum, const float *gacc, long n)
{
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    gsum[i] += (double)gacc[i];
}

// Gradients w.r.t. (vp, eps, del) from those w.r.t. (v2 = (vp dt)^2, eps, eps-del), in place:
// g[0..n) v2 -> vp, g[n..2n) eps -> eps (+ eps-del part), g[2n..3n) eps-del -> del.
__global__ void k_grad_finish(double *g, const float *vdt2, float dt, long n)
{
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double vdt = sqrt((double)vdt2[i]);
    g[i] *= 2.0 * vdt * dt;
    const double ged = g[2 * n + i];
    g[n + i] += ged;
    g[2 * n + i] = -ged;
  }
}

// FP32 <-> BF16 copies of a state block (checkpoints stored in BF16).
__global__ void k_pack_bf16(unsigned short *dst, const float *src, long n)
{
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = StoreBF16::enc(src[i]);
}
__global__ void k_unpack_bf16(float *dst, const unsigned short *src, long n)
{
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = StoreBF16::dec(src[i]);
}

}  // namespace qp3d
