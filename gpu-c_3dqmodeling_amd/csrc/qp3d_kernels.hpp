// gfx950 kernels of the qP3D time step.
//
// One time step of the reference (3dPureqPModeling_single_shot.cu:277-337) launches five
// per-cell kernels and copies 16 B/cell over PCIe. Here a step is four launches of two
// stencil passes over the wavefield (the PML update goes out as two launches: x and z faces,
// y faces), hand-tiled for CDNA4:
//   k_gradmax        max over the grid of |du/dx|,|du/dy|,|du/dz| (GPU_kernel.cu:14-33 grad and
//                    the host absMaxval at 3dPureqPModeling_single_shot.cu:284-292), reduced on
//                    device;
//   k_update<PML=0>  interior cells: gradient recomputed in registers, scalar operator S
//                    (GPU_kernel.cu:87-115), 10th-order second derivatives and the leapfrog
//                    (GPU_kernel.cu:552-565), source injection (GPU_kernel.cu:720-731) and the
//                    receiver-plane write (3dPureqPModeling_single_shot.cu:333-337);
//   k_update<PML=1>  the same for the PML tiles, with the split-field recursions
//                    (GPU_kernel.cu:566-637);
//   k_update<PML=2>  the same for the folded y-face tiles, whose model and PML state live in
//                    compact arrays (YCompact).
// k_sfield forms the scalar-operator field of a snapshot step on request (the reference's
// writer for it is commented out, 3dPureqPModeling_single_shot.cu:381-409).
// Optionally (SPEC, gmax_mode 2) there is no gradient pass: k_update<..., SPEC> takes a
// threshold predicted from the previous steps' G, reduces G itself and marks the cells near the
// predicted threshold; then k_spec_mark (only when the true threshold left the marked band)
// and k_spec_fix correct the cells whose decision the prediction got wrong.
// Tiles are aligned to the PML boundary, so an interior tile holds only interior cells and a
// PML tile only cells of one face family (x faces own edges and corners, then y, then z),
// which makes the family uniform over a workgroup and keeps the PML state out of the interior
// kernel's registers. The copy-rotation kernel (GPU_kernel.cu:733-812) disappears: u_next and
// the PML state are written in place of the oldest slot (each cell reads that slot only at its
// own index) and the host flips slot parity each step.
//
// Tiling: a workgroup owns an xy tile (at most TX x 64 cells, y unit-stride, one 64-lane
// wavefront per x row group) and marches through a z range; its work item comes from a list
// the host builds (ordered so that each XCD's L2 sees neighbouring tiles). The current plane
// with a 5-cell halo sits in LDS (double-buffered, one barrier per plane) for the x and y
// derivatives; each lane keeps an 11-deep register queue of its column in z for the z
// derivatives. Every global load of plane z+1 (queue front, halos, model, previous wavefield,
// PML state) is issued before plane z is computed, so HBM latency overlaps the arithmetic of
// the plane before. Loads of lanes outside the tile read a clamped cell and are never used.
// A lane's offsets within a plane are fixed for the whole march and the plane
// base is a scalar, so each load is one instruction with no vector address arithmetic
// (grids stay below 2^30 cells). Lanes start on a 128-byte line where the tile allows it, and
// in the gradient pass each lane owns a run of consecutive rows (see Lane).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

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

// The double-buffered LDS plane of a tile (plane z with its 5-cell halos, LX + 1 rows: the
// last holds a spare slot), owned by the kernel that runs the tile.
template <class T>
using TileLds = float (*)[T::LX + 1][T::LY];

// p advanced by a byte offset (a 32-bit offset from a scalar base keeps loads in the
// scalar-base + vector-offset form).
template <class P>
__device__ __forceinline__ const P *bat(const P *p, unsigned b)
{
  return reinterpret_cast<const P *>(reinterpret_cast<const char *>(p) + b);
}
template <class P>
__device__ __forceinline__ P *bat(P *p, unsigned b)
{
  return reinterpret_cast<P *>(reinterpret_cast<char *>(p) + b);
}

// Storage policies for wavefields and PML state. Arithmetic is always float. Offsets are byte
// offsets of the cell in a 4-byte-per-cell array.
struct StoreF32 {
  using T = float;
  static __device__ __forceinline__ float ld(const T *p, unsigned b) { return *bat(p, b); }
  static __device__ __forceinline__ float ld_stream(const T *p, unsigned b) { return __builtin_nontemporal_load(bat(p, b)); }
  static __device__ __forceinline__ void st(T *p, unsigned b, float v) { __builtin_nontemporal_store(v, bat(p, b)); }
  static __device__ __forceinline__ float rnd(float v) { return v; }
};

// bfloat16 storage (round to nearest even), float arithmetic. Offsets stay 4-byte-cell
// offsets; the element offset is half of it.
struct StoreBF16 {
  using T = unsigned short;
  static __device__ __forceinline__ unsigned short enc(float v) { return __builtin_bit_cast(unsigned short, (__bf16)v); }
  static __device__ __forceinline__ float dec(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
  static __device__ __forceinline__ float ld(const T *p, unsigned b) { return dec(*bat(p, b >> 1)); }
  static __device__ __forceinline__ float ld_stream(const T *p, unsigned b) { return dec(__builtin_nontemporal_load(bat(p, b >> 1))); }
  static __device__ __forceinline__ void st(T *p, unsigned b, float v) { __builtin_nontemporal_store(enc(v), bat(p, b >> 1)); }
  static __device__ __forceinline__ float rnd(float v) { return dec(enc(v)); }
};

// IEEE half storage, float arithmetic. The fields are stored times a power-of-two scale that
// the plan folds into the source (the equations are linear in the wavefield and PML state) and
// takes out of the record, snapshots and maxima it returns.
struct StoreF16 {
  using T = unsigned short;
  static __device__ __forceinline__ unsigned short enc(float v) { return __builtin_bit_cast(unsigned short, (_Float16)v); }
  static __device__ __forceinline__ float dec(unsigned short h) { return (float)__builtin_bit_cast(_Float16, h); }
  static __device__ __forceinline__ float ld(const T *p, unsigned b) { return dec(*bat(p, b >> 1)); }
  static __device__ __forceinline__ float ld_stream(const T *p, unsigned b) { return dec(__builtin_nontemporal_load(bat(p, b >> 1))); }
  static __device__ __forceinline__ void st(T *p, unsigned b, float v) { __builtin_nontemporal_store(enc(v), bat(p, b >> 1)); }
  static __device__ __forceinline__ float rnd(float v) { return dec(enc(v)); }
};

__device__ __forceinline__ float ld_model(const float *p, unsigned b) { return __builtin_nontemporal_load(bat(p, b)); }

// Grid geometry. Fields are [nz][nx][ys] on the device: each x row holds ny cells and is padded
// to ys, a multiple of 32 cells, so that every row starts on a 128-byte line (a row of 64 lanes
// then touches two lines instead of three; rows off the line boundary cost up to 37% of the
// update's time). The padding cells are never read or written by the stencils.
struct Geo {
  int nx, ny, nz, pml;
  int ys;        // row stride in cells (row_pitch(ny))
  unsigned pln;  // nx * ys
};

__host__ __device__ constexpr int row_pitch(int ny) { return (ny + 31) / 32 * 32; }

// Work item of one workgroup: x in [x0,x1), y in [y0,y1), z in [z0,z1); x1-x0 <= TX, y1-y0 <= 64.
// Packed as {x0 | x1<<16, y0 | y1<<16, z0, z1}; z0 == z1 marks padding.
// Bit 31 of the y word marks a folded y-face tile (both y faces of the PML, see Lane).
__device__ __forceinline__ void unpack(int4 w, int &x0, int &x1, int &y0, int &y1, int &z0, int &z1, bool &fold)
{
  x0 = w.x & 0xffff; x1 = (unsigned)w.x >> 16;
  y0 = w.y & 0xffff; y1 = ((unsigned)w.y >> 16) & 0x7fff;
  fold = w.y < 0;
  z0 = w.z; z1 = w.w;
}

struct Coef {
  float c1[H], c2[H];        // 10th-order rows of the FD tables (entries 31..35)
  float ix, iy, iz;          // 1/dx, 1/dy, 1/dz
  float ixx, iyy, izz;       // 1/dx^2, 1/dy^2, 1/dz^2
  float dt;
  float g1x[H], g1y[H], g1z[H];  // c1 / dx, c1 / dy, c1 / dz (gradient-max pass)
};

// Time-step index of a launch: *kdev + koff when kdev is set (launches replayed from a
// captured graph read the step from device memory), koff otherwise.
struct Step {
  const int *kdev;
  int koff;
  __device__ __forceinline__ int get() const { return (kdev ? *kdev : 0) + koff; }
};

// Amplitude bound of the gradient pass. The update kernels of step k record max |u_{k+1}| per
// block of 8 planes x 8 rows x 32 cells (z, x, y; float bits, atomic max) in buffer (k+1) % 3 of
// three. A gradient component at a cell is sum_m g_m (u(+m) - u(-m)), so its magnitude is at
// most 2 sum_m |g_m| times max |u| within 5 cells: a gradient tile whose blocks (with that
// halo) bound every component below a value of G_k already computed exactly cannot hold the
// maximum, and the pass skips it without loading u. G_k stays exact. b == nullptr: off.
//
// Counters of instrumented steps: CNT_SLOTS slots of 16 words, one 128-byte line each, a
// workgroup adding into slot blockIdx.x % CNT_SLOTS (one address for every workgroup of a launch
// serialises its atomics): words 0, 1 gradient tiles evaluated, skipped; 2 + 3 * kind + j update
// cells in items (j = 0), left out at rest (1), without eps - del (2).
constexpr int CNT_SLOTS = 32, CNT_STRIDE = 16;
__device__ __forceinline__ unsigned long long *cnt_slot(unsigned long long *c)
{
  return c + (blockIdx.x % CNT_SLOTS) * CNT_STRIDE;
}
constexpr int AB_Z = 3, AB_X = 3, AB_Y = 5;  // log2 of the block extents
//
// The interior updates record in the same buffer slot two more maxima per block: of the
// increment |u_{k+1} - u_k| (d) and of |grad u_k| as the update computed it for S (c; it bounds
// every component). Since grad u_{k+1} = grad u_k + grad(u_{k+1} - u_k), a component of u_{k+1}
// at a cell is at most c of its block plus 2 sum_m |g_m| times the increment's maximum within 5
// cells: the increment is a small part of u (omega dt ~ 0.16), so this bound is several times
// tighter than the one from max |u|, and the pass uses the smaller of the two. The PML kernels,
// at their register limit, do not track them: a bound over blocks that meet the PML (pml_block)
// is the one from max |u|.
struct AmpGrid {
  // [3][3][nbz][nbx][nby]: max |u|, then max |u_{k+1} - u_k| and max |grad u_k| (both in slot
  // k + 1), three rotating slots each (one pointer: fewer scalar registers in the updates)
  unsigned *b;
  int nbz, nbx, nby;
  unsigned nblk;  // nbz * nbx * nby
  // [nbx][nby][nbz] map of the current step (k_quiet): bit 0 quiet (S = 1 in every cell), bit 1
  // touched (u has been nonzero within the block's neighbourhood at some step since the reset;
  // a block without it is at rest). z fastest: a wavefront reading one z block per lane loads one
  // or two cache lines.
  unsigned char *q;
  __device__ __forceinline__ unsigned *buf(int k) const { return b + (unsigned)(k % 3) * nblk; }
  __device__ __forceinline__ unsigned *dbuf(int k) const { return b + (unsigned)(3 + k % 3) * nblk; }
  __device__ __forceinline__ unsigned *cbuf(int k) const { return b + (unsigned)(6 + k % 3) * nblk; }
  // whether block (bz, bx, by) holds a PML cell (d and c not recorded there)
  __device__ __forceinline__ static bool pml_block(const Geo &g, int bz, int bx, int by)
  {
    return (bz << AB_Z) < g.pml || ((bz + 1) << AB_Z) > g.nz - g.pml || (bx << AB_X) < g.pml ||
           ((bx + 1) << AB_X) > g.nx - g.pml || (by << AB_Y) < g.pml || ((by + 1) << AB_Y) > g.ny - g.pml;
  }
  __device__ __forceinline__ unsigned qi(int bz, int bx, int by) const
  {
    return ((unsigned)bx * (unsigned)nby + (unsigned)by) * (unsigned)nbz + (unsigned)bz;
  }
};

// Per-lane running maximum of |u_next| (float bits: inf and NaN order above every finite value)
// over planes [za, zb), flushed into the block maxima of every z block those planes touch. Each
// half-wavefront holds 32 consecutive y cells (a folded tile: one y face each), so its maximum
// goes to the one or two y blocks its first and last lanes fall in, for every x block of the
// tile. The (z block, x block, half, y block) targets are spread over the lanes: one atomic
// instruction per flush (at most 4 z blocks x 3 x blocks x 4 targets = 48 lanes).
__device__ __forceinline__ unsigned half_wave_max(unsigned am)
{
  int v = (int)am;
  v = max((unsigned)v, (unsigned)__builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));
  v = max((unsigned)v, (unsigned)__builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));
  v = max((unsigned)v, (unsigned)__builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));
  v = max((unsigned)v, (unsigned)__builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));
  v = max((unsigned)v, (unsigned)__builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));
  return (unsigned)v;
}

__device__ __forceinline__ void amp_flush(const AmpGrid &ag, int kstep, int za, int zb, int x0, int x1, int ylane,
                                          unsigned &am, unsigned &ad, unsigned &ac)
{
  // half-wavefront maxima with DPP: prefix maxima within each 16-lane row (row_shr 1, 2, 4, 8),
  // then row 0's maximum into row 1 and row 2's into row 3 (row_bcast:15), so lanes 31 and 63
  // hold the maxima of lanes 0..31 and 32..63 (lanes without a source read 0, the identity)
  const unsigned vm = half_wave_max(am), vd = half_wave_max(ad), vc = half_wave_max(ac);
  const int yb = min(ylane >> AB_Y, ag.nby - 1);
  // targets t = 0..3: half 0 first / last lane's y block, half 1 first / last
  const int ty[4] = {__builtin_amdgcn_readlane(yb, 0), __builtin_amdgcn_readlane(yb, 31),
                     __builtin_amdgcn_readlane(yb, 32), __builtin_amdgcn_readlane(yb, 63)};
  const int bz0 = za >> AB_Z, nbz = ((zb - 1) >> AB_Z) - bz0 + 1;
  const int bx0 = x0 >> AB_X, nbx = ((x1 - 1) >> AB_X) - bx0 + 1;
  const int l = __lane_id();
  const int t = l & 3, c = l >> 2;  // target, (z block, x block) pair
  const int bz = bz0 + c / nbx, bx = bx0 + c % nbx;
  const int ybt = t == 0 ? ty[0] : t == 1 ? ty[1] : t == 2 ? ty[2] : ty[3];
  const unsigned wm = t < 2 ? (unsigned)__builtin_amdgcn_readlane((int)vm, 31) : (unsigned)__builtin_amdgcn_readlane((int)vm, 63);
  const unsigned wd = t < 2 ? (unsigned)__builtin_amdgcn_readlane((int)vd, 31) : (unsigned)__builtin_amdgcn_readlane((int)vd, 63);
  const unsigned wc = t < 2 ? (unsigned)__builtin_amdgcn_readlane((int)vc, 31) : (unsigned)__builtin_amdgcn_readlane((int)vc, 63);
  // the last-lane target only where it differs from the first-lane one
  const bool dup = (t == 1 && ty[1] == ty[0]) || (t == 3 && ty[3] == ty[2]);
  if (c < nbz * nbx && !dup) {
    const unsigned o = (unsigned)(kstep + 1) % 3u * ag.nblk + ((unsigned)bz * ag.nbx + (unsigned)bx) * ag.nby + (unsigned)ybt;
    if (wm != 0u) atomicMax(ag.b + o, wm);
    if (wd != 0u) atomicMax(ag.b + 3 * ag.nblk + o, wd);
    if (wc != 0u) atomicMax(ag.b + 6 * ag.nblk + o, wc);
  }
  am = 0u;
  ad = 0u;
  ac = 0u;
}

// Block maxima of the interior update through LDS: every wavefront adds its half-wavefront
// maxima of the three values into one workgroup set per y block of the tile (two lanes, LDS
// atomics), and wavefront 0 moves a set into the global block maxima of every (z block, x block,
// y block) of the stretch and the tile once a barrier has passed, then clears it. One set of
// global atomics per workgroup and stretch instead of one per wavefront.
__device__ __forceinline__ void amp_lds_add(unsigned (&sv)[3][4], int y0, int ylane, int nby, unsigned &am,
                                            unsigned &ad, unsigned &ac)
{
  const unsigned vm = half_wave_max(am), vd = half_wave_max(ad), vc = half_wave_max(ac);
  const int yb = min(ylane >> AB_Y, nby - 1) - (y0 >> AB_Y);
  const int f0 = __builtin_amdgcn_readlane(yb, 0), l0 = __builtin_amdgcn_readlane(yb, 31);
  const int f1 = __builtin_amdgcn_readlane(yb, 32), l1 = __builtin_amdgcn_readlane(yb, 63);
  const int l = __lane_id();
  if (l == 31 || l == 63) {
    const int yf = l == 31 ? f0 : f1, yl = l == 31 ? l0 : l1;
    atomicMax(&sv[0][yf], vm);
    atomicMax(&sv[1][yf], vd);
    atomicMax(&sv[2][yf], vc);
    if (yl != yf) {
      atomicMax(&sv[0][yl], vm);
      atomicMax(&sv[1][yl], vd);
      atomicMax(&sv[2][yl], vc);
    }
  }
  am = 0u;
  ad = 0u;
  ac = 0u;
}

// Wavefront 0: planes [za, zb) of the tile's x range [x0, x1) and the y blocks of its 64 lanes
// from y0.
__device__ __forceinline__ void amp_lds_flush(const AmpGrid &ag, int kstep, unsigned (&sv)[3][4], int za, int zb,
                                              int x0, int x1, int y0)
{
  const int yb0 = y0 >> AB_Y, nyb = min((y0 + TY - 1) >> AB_Y, ag.nby - 1) - yb0 + 1;
  const int bz0 = za >> AB_Z, nbz = ((zb - 1) >> AB_Z) - bz0 + 1;
  const int bx0 = x0 >> AB_X, nbx = ((x1 - 1) >> AB_X) - bx0 + 1;
  const int l = __lane_id();
  const int yi = l % nyb, t = l / nyb, xi = t % nbx, zi = t / nbx;
  const bool on = l < nbz * nbx * nyb;
  const int yr = on ? yi : 0;
  const unsigned wm = sv[0][yr], wd = sv[1][yr], wc = sv[2][yr];
  if (on) {
    const unsigned o = (unsigned)(kstep + 1) % 3u * ag.nblk +
                       ((unsigned)(bz0 + zi) * ag.nbx + (unsigned)(bx0 + xi)) * ag.nby + (unsigned)(yb0 + yi);
    if (wm != 0u) atomicMax(ag.b + o, wm);
    if (wd != 0u) atomicMax(ag.b + 3 * ag.nblk + o, wd);
    if (wc != 0u) atomicMax(ag.b + 6 * ag.nblk + o, wc);
  }
  if (l < 12) sv[l >> 2][l & 3] = 0u;  // after every lane's read above (one wavefront, in order)
}

// The y faces of the PML fold into one tile per x row group when each is at most 27 cells
// wide (pml <= 32: the two faces with their 5-cell halos fit one 64-lane wavefront).
__host__ __device__ constexpr bool y_fold(int pml) { return pml <= 32; }

// Compact arrays of the y-face family (folded faces only). Each y-face row is a 27-cell piece of
// a 128-byte line in the full-grid layout, so per-cell loads of the y-face tiles would fetch
// whole lines for partial use; these arrays hold the model and the seven PML state fields of the
// y-face cells as [nz][nx - 2 pml][2 yw], yw = pml - 5 cells per face: row x - pml holds the low
// face's y = 5..pml-1 in columns [0, yw) and the high face's y = ny-pml..ny-6 in [yw, 2 yw), so a
// tile's rows are contiguous. u stays in the full-grid layout.
template <class S>
struct YCompact {
  typename S::T *w1p, *w3p, *pp;        // PML slots k-1 (overwritten with the new value)
  const typename S::T *w1c, *w3c, *pc;  // PML slots k
  typename S::T *w21;
  const float *vdt2, *eps, *epsd;
  unsigned pln;  // cells per plane: (nx - 2 pml) * 2 yw
  int yw;        // cells per face; 0: the y faces use the full-grid arrays
};

// Compact index (cells) of a cell of the y-face family: x in [pml, nx-pml), y in [5, pml) or
// [ny-pml, ny-5).
__device__ __forceinline__ unsigned ycompact_cell(int ny, int pml, int yw, unsigned pln, int z, int x, int y)
{
  const int col = y < pml ? y - H : yw + (y - (ny - pml));
  return (unsigned)z * pln + (unsigned)(x - pml) * (unsigned)(2 * yw) + (unsigned)col;
}

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
  YCompact<S> yc;                           // model and PML state of the y-face family, compact
  const float4 *pmltab;                     // [3][pml]: K1, K2, K3, E per face family and layer
  unsigned *gmax;                           // [nt] G per step (float bits); SPEC: G_k reduced here
  float *gpred;                             // SPEC: [nt] predicted G per step (fix-up writes k+1)
  int *flag;                                // SPEC: set to 1 when a non-finite gradient is seen
  unsigned *fixmask;                        // SPEC: one bit per cell, cells to check after the step
  const float *wav;                         // [nt] wavelet
  Step step;
  int src_on;
  int sb_z0, sb_z1, sb_x0, sb_x1, sb_y0, sb_y1;  // source box (inclusive bounds)
  const float *sbox;                        // dense source values over the box
  float *rec;                               // record [nt][hx][hy]
  int sz;                                   // receiver plane
  const int4 *work;                         // one item per workgroup
  AmpGrid amp;                              // block maxima of |u_next| (b == nullptr: off)
  unsigned *rest_off;  // set when this step runs without the map: the launch clears the word
                       // that tells k_quiet every step since the reset kept the map
  unsigned long long *ucount;  // counter slots + 2 (cnt_slot), instrumented steps: per update kind,
                               // cells in items, left out at rest, updated without eps - del
};

template <class S>
struct GradArgs {
  Geo g;
  Coef c;
  const typename S::T *ucur;
  unsigned *gmax;   // [nt] out: atomic max of the gradient components (float bits)
  int *flag;        // out: set to 1 when a non-finite gradient is seen
  const int4 *work;
  Step step;
  unsigned *fixmask;  // marking pass: one bit per cell whose threshold decision changes
  const float *gpred; // marking pass: predicted G per step
  AmpGrid amp;        // block maxima of |u_k| (b == nullptr: no skipping)
  int skip;           // 1: skip tiles whose amplitude bound is below the G_k found so far
  unsigned long long *argkey;  // [2]: (tile maximum bits << 32 | item) of the step's largest tile
  int nwork;          // items in `work`
  unsigned long long *gcount;  // counter slots (cnt_slot): tiles evaluated, skipped (instrumented steps)
  unsigned *gseed;    // [2] per step parity: the exact maximum the bound pass found (float bits)
  unsigned *gseed_out;  // the bound pass: where its tiles' maximum goes (else null)
};

// A lane's fixed position in its tile and its in-plane offsets.
//
// In a folded tile (the two y faces of the PML, each at most 27 cells wide) lanes 0..31 cover
// y in [y0, y0+32) and lanes 32..63 cover [yb, yb+32) with yb = ny - pml, and the LDS row holds
// the two strips side by side, each with its own 5-cell halos (columns [0,37) and [37,74)).
// Loads of u are exact up to 5 cells past the tile (the halo the stencil reads) and clamped
// beyond; per-cell loads (model, previous wavefield, PML state) of lanes outside the tile are
// clamped into the tile, so idle lanes add no memory traffic. A non-folded tile may start its
// lanes below y0 (at a 32-cell boundary); those lanes are outside the tile.
//
// Rows: a lane's rows r = 0..RX-1 sit at tile row tx + WX*r (interleaved across the wavefronts)
// or, with RUN, at RX*tx + r (a run of consecutive rows per wavefront, so the x stencil of the
// run reads only the 10 rows around it from LDS and the rest from the lane's own registers).
template <class T, bool RUN = false>
struct Lane {
  int tx, ty;            // wavefront (x group) and lane in the workgroup
  int y, Y;              // y of the lane and its LDS column
  int yc;                // y clamped into the tile (per-cell loads)
  bool yon;              // the lane's y is inside the tile
  bool fold;             // folded y-face tile (uniform)
  unsigned qoff[T::RX];  // in-plane byte offsets (4-byte cells) of the lane's rows for u
  unsigned off[T::RX];   // the same, clamped into the tile, for per-cell fields
  unsigned hxo[T::NXH];  // x-halo byte offsets
  unsigned hyo;          // y-halo byte offset
  int hyL;               // LDS slot (row * LY + column) of the lane's y-halo value; lanes
                         // without one write a spare slot past the plane (LX * LY)
  bool hx_on[T::NXH];    // wave-uniform
  bool hy_on;

  __device__ __forceinline__ int row(int r) const { return RUN ? T::RX * tx + r : tx + T::WX * r; }

  __device__ __forceinline__ Lane(const Geo &g, int x0, int x1, int y0, int y1, bool fold_) : fold(fold_)
  {
    ty = threadIdx.x;
    tx = __builtin_amdgcn_readfirstlane(threadIdx.y);
    int yq;  // y for u loads
    int ylo = y0;  // y of lane 0 (non-folded tiles)
    if (!fold) {
      // Lane 0 sits on the 32-cell (128-byte) boundary at or below y0 when the tile still fits
      // the wavefront from there, so each row a wavefront loads starts on a cache line.
      if (y1 - (y0 & ~31) <= TY) ylo = y0 & ~31;
      y = ylo + ty;
      Y = H + ty;
      yon = y >= y0 && y < y1;
      yq = min(y, y1 + H - 1);
      yc = min(max(y, y0), y1 - 1);
    } else {
      const int s = ty >> 5, j = ty & 31, w = y1 - y0;
      const int ybase = s ? g.ny - g.pml : y0;
      y = ybase + j;
      Y = H + ty + (s ? H : 0);
      yon = j < w;
      yq = ybase + min(j, w + H - 1);
      yc = ybase + min(j, w - 1);
    }
#pragma unroll
    for (int r = 0; r < T::RX; r++) {
      const int x = x0 + row(r);
      qoff[r] = 4u * ((unsigned)min(x, x1 + H - 1) * g.ys + yq);
      off[r] = 4u * ((unsigned)min(x, x1 - 1) * g.ys + yc);
    }
#pragma unroll
    for (int i = 0; i < T::NXH; i++) {
      const int h = tx + T::WX * i;
      hx_on[i] = h < 2 * H;
      const int X = h < H ? h : T::TX + h;
      hxo[i] = 4u * ((unsigned)min(x0 - H + X, x1 + H - 1) * g.ys + yq);
    }
    const int t = tx * TY + ty;
    hy_on = t < T::TX * 2 * H;
    const int row = t / (2 * H), col = t - row * (2 * H);
    int gy;
    if (!fold) gy = max(0, min(ylo - H + (col < H ? col : TY + col), y1 + H - 1));
    else gy = col < H ? y0 - H + col : g.ny - g.pml - H + (col - H);
    hyo = 4u * ((unsigned)min(x0 + row, x1 + H - 1) * g.ys + gy);
    hyL = hy_on ? (H + row) * T::LY + halo_col(col, fold) : T::LX * T::LY;
  }

  // Re-issues the 32-bit offsets inside the march (an empty asm with read-write operands, no
  // instruction). Without it the compiler hoists their zero-extension out of the loop and then
  // forms every load address in 64-bit vector registers; with it each load keeps the
  // scalar-base + 32-bit vector-offset form.
  __device__ __forceinline__ void refresh()
  {
#pragma unroll
    for (int r = 0; r < T::RX; r++) asm volatile("" : "+v"(qoff[r]), "+v"(off[r]));
#pragma unroll
    for (int i = 0; i < T::NXH; i++) asm volatile("" : "+v"(hxo[i]));
    asm volatile("" : "+v"(hyo));
  }

  // LDS column of y-halo value `col` (0..9) of this lane's load.
  __device__ __forceinline__ static int halo_col(int col, bool fold)
  {
    // folded: strip A's left halo in columns 0..4, strip B's in 37..41
    return !fold ? (col < H ? col : TY + col) : (col < H ? col : 32 + col);
  }
};

// Global loads of u for one plane of a tile: the lane's queue front (plane z+5) and its share
// of the x and y halos of plane z.
template <class S, class T>
struct PlaneLoad {
  float qn[T::RX];
  float hx[T::NXH];
  float hy;

  template <bool RUN>
  __device__ __forceinline__ void issue(const typename S::T *u, const Geo &g, const Lane<T, RUN> &L, int z)
  {
    const typename S::T *uq = u + (size_t)(z + H) * g.pln;
    const typename S::T *uz = u + (size_t)z * g.pln;
#pragma unroll
    for (int r = 0; r < T::RX; r++) qn[r] = S::ld(uq, L.qoff[r]);
#pragma unroll
    for (int i = 0; i < T::NXH; i++)
      if (__builtin_amdgcn_readfirstlane(L.hx_on[i])) hx[i] = S::ld(uz, L.hxo[i]);
    if (L.hy_on) hy = S::ld(uz, L.hyo);
  }

  // Write plane z (centre values from the queue, halos from this load) into an LDS buffer
  // (LX + 1 rows: the last holds the spare slot).
  template <bool RUN>
  __device__ __forceinline__ void stage(float (*sm)[T::LY], const float (&centre)[T::RX], const Lane<T, RUN> &L) const
  {
#pragma unroll
    for (int r = 0; r < T::RX; r++) sm[H + L.row(r)][L.Y] = centre[r];
#pragma unroll
    for (int i = 0; i < T::NXH; i++) {
      const int h = L.tx + T::WX * i;
      if (__builtin_amdgcn_readfirstlane(L.hx_on[i])) sm[h < H ? h : T::TX + h][L.Y] = hx[i];
    }
    (&sm[0][0])[L.hyL] = hy;  // lanes without a y-halo value write the spare slot
  }
};

template <class S, class T, bool RUN>
__device__ __forceinline__ void prime_queue(float (&q)[T::RX][2 * H + 1], const typename S::T *u, const Geo &g,
                                            const Lane<T, RUN> &L, int z0)
{
#pragma unroll
  for (int j = 0; j < 2 * H; j++) {
    const typename S::T *up = u + (size_t)(z0 - H + j) * g.pln;
#pragma unroll
    for (int r = 0; r < T::RX; r++) q[r][j] = S::ld(up, L.qoff[r]);
  }
}

// The z queue rotates instead of shifting: the march is unrolled by its length (11), and at
// unrolled position `it` the value of plane z-5+j sits in slot (j + it) % 11, so the only
// register traffic per plane is the new front landing in the freed slot.
constexpr int QN = 2 * H + 1;

template <int I>
using ic = std::integral_constant<int, I>;

// f(ic<I>) for I = 0.. while I < remaining, fully unrolled at compile time.
template <int I, int N>
struct Unroll {
  template <class F>
  __device__ __forceinline__ static void run(F &f, int remaining)
  {
    if (I < remaining) {
      f(ic<I>{});
      Unroll<I + 1, N>::run(f, remaining);
    }
  }
};
template <int N>
struct Unroll<N, N> {
  template <class F>
  __device__ __forceinline__ static void run(F &, int)
  {
  }
};

// Scalar operator of a cell from its gradient (GPU_kernel.cu:93-111): 1 at or below the
// threshold |grad u| <= thr; above it n = grad u / (|grad u| + 2.22e-17) and
// S = (1 + sqrt(1 - 8 (eps-del) (nx^2+ny^2) nz^2 / (1 + 2 eps (nx^2+ny^2))^2)) / 2.
// e = eps, ed = eps - del. The update kernel and the S-field snapshot share it.
__device__ __forceinline__ float cell_mode(float ux, float uy, float uz)
{
  return __builtin_amdgcn_sqrtf(fmaf(ux, ux, fmaf(uy, uy, uz * uz)));
}
__device__ __forceinline__ float scalar_op_m(float mode, float ux, float uy, float uz, float e, float ed, float thr)
{
  float Sv = 1.f;
  if (mode > thr) {
    const float inv = __builtin_amdgcn_rcpf(mode + 2.22e-17f);
    const float nxv = ux * inv, nyv = uy * inv, nzv = uz * inv;
    const float qa = fmaf(nxv, nxv, nyv * nyv);
    const float den = fmaf(2.f * e, qa, 1.f);
    const float X8 = 8.f * ed * qa * (nzv * nzv) * __builtin_amdgcn_rcpf(den * den);
    Sv = 0.5f + 0.5f * __builtin_amdgcn_sqrtf(1.f - X8);
  }
  return Sv;
}
__device__ __forceinline__ float scalar_op(float ux, float uy, float uz, float e, float ed, float thr)
{
  return scalar_op_m(cell_mode(ux, uy, uz), ux, uy, uz, e, ed, thr);
}

// One tap of a 10th-order second derivative: acc + c2_m ((u+m - 2 u0) + u-m), with the
// reference's term-wise second difference (CPU_function.c:383-404), twou0 = 2 u0 (exact). Summing
// c2_m (u+m + u-m) and subtracting 2 u0 sum(c2) instead cancels two terms three times the
// size of u0 and carries about three times the rounding noise, which the threshold of the
// scalar operator amplifies over long runs.
__device__ __forceinline__ float d2tap(float c2m, float up, float um, float twou0, float acc)
{
  return fmaf(c2m, (up - twou0) + um, acc);
}

// The threshold 0.01*G of the scalar operator: the reference compares |grad u| > 0.01*G in
// double (GPU_kernel.cu:99); for a float magnitude this equals mode > rd(0.01*G).
__device__ __forceinline__ float s_threshold(float G) { return __double2float_rd(1.0e-02 * (double)G); }

// Predicted threshold (gmax_mode 2). The update of step k runs before G_k is known: it takes
// ts = rd(0.01 Gp) with Gp = gpred[k], the prediction stored by the fix-up of step k-1
// (predict_next), marks every cell whose |grad u| lies in the band (lo, hi] around ts, and
// reduces G_k itself. The true threshold tt = rd(0.01 G_k) then changes the decision
// |grad u| > thr only for cells between ts and tt: all of them are marked when lo <= tt <= hi
// (valid), and a marking pass over the grid finds them otherwise.
constexpr float SPEC_BAND = 1.f / 128.f;
struct SpecBand {
  float ts, lo, hi;
};
__device__ __forceinline__ SpecBand spec_band(const float *gpred, int k)
{
  SpecBand b;
  b.ts = s_threshold(k >= 1 ? gpred[k] : 0.f);
  b.lo = b.ts * (1.f - SPEC_BAND);
  b.hi = b.ts * (1.f + SPEC_BAND);
  return b;
}
__device__ __forceinline__ bool spec_valid(const SpecBand &b, float tt) { return tt >= b.lo && tt <= b.hi; }
__device__ __forceinline__ void mark_cell(unsigned *mask, unsigned cell) { atomicOr(mask + (cell >> 5), 1u << (cell & 31)); }

// Workgroup maximum of m into *gmax (float bits, one atomic per workgroup); *flag |= 1 when a
// lane saw a non-finite gradient (nonfinite, or m above the float range).
template <class T>
__device__ __forceinline__ void reduce_gmax(float m, bool nonfinite, int tx, int ty, unsigned *gmax, int *flag,
                                            unsigned long long *argkey = nullptr, int item = -1,
                                            unsigned *seed = nullptr)
{
  __shared__ float red[T::WX];
  __shared__ int redbad[T::WX];
  int bad = nonfinite || !(m <= 3.4e38f);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    m = fmaxf(m, __shfl_xor(m, off, 64));
    bad |= __shfl_xor(bad, off, 64);
  }
  if (ty == 0) {
    red[tx] = m;
    redbad[tx] = bad;
  }
  __syncthreads();
  if (tx == 0 && ty == 0) {
    float v = red[0];
    int bb = redbad[0];
#pragma unroll
    for (int w = 1; w < T::WX; w++) {
      v = fmaxf(v, red[w]);
      bb |= redbad[w];
    }
    atomicMax(gmax, __float_as_uint(v));
    if (bb) atomicOr(flag, 1);
    // the tile's maximum with its item: the next step's bound pass starts from that tile
    if (argkey && item >= 0) atomicMax(argkey, ((unsigned long long)__float_as_uint(v) << 32) | (unsigned)item);
    if (seed) atomicMax(seed, __float_as_uint(v));
  }
}

__device__ __forceinline__ unsigned wave_max(unsigned m)
{
  // wavefront maximum into lane 63 (row prefix maxima, then rows 0 -> 1, 2 -> 3, 1 -> 2 and 3)
  int v = (int)m;
  v = max((unsigned)v, (unsigned)__builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));
  v = max((unsigned)v, (unsigned)__builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));
  v = max((unsigned)v, (unsigned)__builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));
  v = max((unsigned)v, (unsigned)__builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));
  v = max((unsigned)v, (unsigned)__builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));
  v = max((unsigned)v, (unsigned)__builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));
  return (unsigned)__builtin_amdgcn_readlane(v, 63);
}

// The smaller of the two bounds on a gradient component of u_k over cells whose blocks hold at
// most C (the previous step's components) and whose 5-cell neighbourhoods hold at most M (|u_k|)
// and D (|u_k - u_{k-1}|); fac = 2 sum_m |g_m|. The margins cover the rounding of both gradient
// sums and of the increment (each within about 1e-6 of the magnitudes they add). Non-finite
// maxima (float bits at or above inf's) give no bound.
__device__ __forceinline__ float grad_bound(unsigned M, unsigned D, unsigned C, float fac)
{
  if (M >= 0x7f800000u) return __builtin_inff();
  const float m = __uint_as_float(M);
  const float old = m * fac * (1.f + 1.0e-5f);
  // D and C are passed as inf bits for blocks that meet the PML (whose kernels do not record
  // them); a non-finite u already shows in M
  if (D >= 0x7f800000u || C >= 0x7f800000u) return old;
  const float d = __uint_as_float(D), c = __uint_as_float(C);
  const float inc = fmaf(d, fac, c) * (1.f + 1.0e-5f) + 1.0e-5f * fac * (m + d);
  return fminf(old, inc);
}

// Amplitude bound of a gradient tile (see AmpGrid): true when every gradient component of the
// tile is provably below the exact maximum the bound pass found for this step (the seed, fixed
// during the pass). Every wavefront decides on its own from the same loads, so the decision is
// the same in all of them without LDS or a barrier.
template <class S, class T>
__device__ __forceinline__ bool amp_skip(const GradArgs<S> &a, int x0, int x1, int y0, int y1, int z0, int z1,
                                         int kstep)
{
  const AmpGrid &ag = a.amp;
  const int zb0 = max(0, (z0 - H) >> AB_Z), zb1 = min(ag.nbz - 1, (z1 - 1 + H) >> AB_Z);
  const int xb0 = max(0, (x0 - H) >> AB_X), xb1 = min(ag.nbx - 1, (x1 - 1 + H) >> AB_X);
  const int yb0 = max(0, (y0 - H) >> AB_Y), yb1 = min(ag.nby - 1, (y1 - 1 + H) >> AB_Y);
  // the tile's own blocks (for the previous components), inside the halo range above
  const int zo0 = z0 >> AB_Z, zo1 = (z1 - 1) >> AB_Z, xo0 = x0 >> AB_X, xo1 = (x1 - 1) >> AB_X;
  const int yo0 = y0 >> AB_Y, yo1 = (y1 - 1) >> AB_Y;
  const int nxb = xb1 - xb0 + 1, nyb = yb1 - yb0 + 1;
  const int nb = (zb1 - zb0 + 1) * nxb * nyb;
  const unsigned *bb = ag.buf(kstep), *db = ag.dbuf(kstep), *cb = ag.cbuf(kstep);
  unsigned m = 0u, d = 0u, c = 0u;
  for (int i = __lane_id(); i < nb; i += TY) {
    const int yb = yb0 + i % nyb, t = i / nyb, xb = xb0 + t % nxb, zb = zb0 + t / nxb;
    const unsigned o = ((unsigned)zb * ag.nbx + (unsigned)xb) * ag.nby + (unsigned)yb;
    const bool pb = AmpGrid::pml_block(a.g, zb, xb, yb);
    m = max(m, bb[o]);
    d = max(d, pb ? 0x7f800000u : db[o]);
    if (zb >= zo0 && zb <= zo1 && xb >= xo0 && xb <= xo1 && yb >= yo0 && yb <= yo1) c = max(c, pb ? 0x7f800000u : cb[o]);
  }
  const unsigned M = wave_max(m), D = wave_max(d), Cm = wave_max(c);
  const float Gk = __uint_as_float(a.gseed[kstep & 1]);
  float fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll
  for (int k = 0; k < H; k++) {
    fx += fabsf(a.c.g1x[k]);
    fy += fabsf(a.c.g1y[k]);
    fz += fabsf(a.c.g1z[k]);
  }
  return grad_bound(M, D, Cm, 2.f * fmaxf(fx, fmaxf(fy, fz))) < Gk;
}

// One work item of the gradient-max pass (w: packed tile, kstep: time step). Each lane owns a
// run of RX consecutive rows: the x stencil of the run reads 10 rows from LDS and takes the
// run's own rows from the lane's queue, which cuts the pass's LDS reads (its bottleneck) from
// 10 + 10 to 10/RX + 10 per cell.
//
// MARK (gmax_mode 2, a step whose predicted threshold ts missed: tt outside the marked band):
// the same march computes |grad u| with the update kernel's arithmetic and marks the cells
// whose decision |grad u| > thr differs between ts and tt; the workgroup exits at once when
// the band was valid.
template <class S, class T, bool MARK = false>
__device__ __forceinline__ void gradmax_tile(const GradArgs<S> &a, int4 w, int kstep, TileLds<T> sm, int item = -1)
{
  const Geo &g = a.g;
  int x0, x1, y0, y1, z0, z1;
  bool fold;
  unpack(w, x0, x1, y0, y1, z0, z1, fold);
  if (z0 >= z1) return;
  if constexpr (!MARK) {
    if (a.skip && a.amp.b) {
      // a uniform decision in a scalar register
      const int sk = __builtin_amdgcn_readfirstlane(amp_skip<S, T>(a, x0, x1, y0, y1, z0, z1, kstep) ? 1 : 0);
      if (a.gcount && threadIdx.x == 0 && threadIdx.y == 0) atomicAdd(cnt_slot(a.gcount) + sk, 1ull);
      if (sk) return;
    }
  }
  float ts = 0.f, tt = 0.f;
  if constexpr (MARK) {
    const SpecBand sb = spec_band(a.gpred, kstep);
    tt = s_threshold(__uint_as_float(a.gmax[kstep]));
    if (spec_valid(sb, tt)) return;
    ts = sb.ts;
  }
  Lane<T, true> L(g, x0, x1, y0, y1, fold);
  const bool ycmp = L.yon;
  bool rows_on[T::RX];
#pragma unroll
  for (int r = 0; r < T::RX; r++) rows_on[r] = ycmp && x0 + L.row(r) < x1;
  float q[T::RX][QN];
  prime_queue<S, T>(q, a.ucur, g, L, z0);
  // The march is unrolled by 22 (twice the queue length), so the LDS buffer of every plane
  // (it & 1) is known at compile time.
  PlaneLoad<S, T> P;
  P.issue(a.ucur, g, L, z0);
  float centre[T::RX];
#pragma unroll
  for (int r = 0; r < T::RX; r++) {
    q[r][2 * H] = P.qn[r];
    centre[r] = q[r][H];
  }
  P.stage(sm[0], centre, L);
  __syncthreads();
  // max of |sum_m (c1_m/d) (u+m - u-m)| over the three axes; fmaxf ignores NaN like the
  // reference's '>=' scan (CPU_function.c:18), and x*0 accumulates NaN for any non-finite
  // gradient.
  float m = 0.f, nanacc = 0.f;
  int zb = z0;
  auto plane = [&](auto itc) {
    constexpr int it = decltype(itc)::value;
    {
      const int z = zb + it;
      const bool more = z + 1 < z1;
      L.refresh();
      if (more) P.issue(a.ucur, g, L, z + 1);
      float (*s)[T::LY] = sm[it & 1];
      const int XB = H + L.row(0);
      // plane-z value of run row j (j outside [0, RX): LDS); the same float LDS holds
      auto xv = [&](int j) -> float { return (j >= 0 && j < T::RX) ? q[j][(H + it) % QN] : s[XB + j][L.Y]; };
#pragma unroll
      for (int r = 0; r < T::RX; r++) {
        const int X = XB + r, Y = L.Y;
        float ux = 0.f, uy = 0.f, uz = 0.f;
        if constexpr (MARK) {
          // the update kernel's first derivatives, bit for bit: sum_m c1_m (u+m - u-m), then 1/d
#pragma unroll
          for (int k = 1; k <= H; k++) {
            ux = fmaf(a.c.c1[k - 1], xv(r + k) - xv(r - k), ux);
            uy = fmaf(a.c.c1[k - 1], s[X][Y + k] - s[X][Y - k], uy);
            uz = fmaf(a.c.c1[k - 1], q[r][(H + k + it) % QN] - q[r][(H - k + it) % QN], uz);
          }
          ux *= a.c.ix; uy *= a.c.iy; uz *= a.c.iz;
          const float mode = cell_mode(ux, uy, uz);
          if (rows_on[r] && ((mode > ts) != (mode > tt))) mark_cell(a.fixmask, (unsigned)z * g.pln + (L.off[r] >> 2));
        } else {
#pragma unroll
          for (int k = 1; k <= H; k++) {
            ux = fmaf(a.c.g1x[k - 1], xv(r + k) - xv(r - k), ux);
            uy = fmaf(a.c.g1y[k - 1], s[X][Y + k] - s[X][Y - k], uy);
            uz = fmaf(a.c.g1z[k - 1], q[r][(H + k + it) % QN] - q[r][(H - k + it) % QN], uz);
          }
          if (rows_on[r]) {
            m = fmaxf(m, fmaxf(fabsf(ux), fmaxf(fabsf(uy), fabsf(uz))));
            nanacc = fmaf(ux, 0.f, nanacc);
          }
        }
      }
      if (more) {
#pragma unroll
        for (int r = 0; r < T::RX; r++) {
          q[r][it % QN] = P.qn[r];
          centre[r] = q[r][(H + 1 + it) % QN];
        }
        P.stage(sm[(it + 1) & 1], centre, L);
        __syncthreads();
      }
    }
  };
  for (; zb < z1; zb += 2 * QN) Unroll<0, 2 * QN>::run(plane, z1 - zb);
  if constexpr (!MARK)
    reduce_gmax<T>(m, !(nanacc == 0.f), L.tx, L.ty, a.gmax + kstep, a.flag, a.argkey ? a.argkey + (kstep & 1) : nullptr,
                   item, a.gseed_out);
}

// Quiet map of step k (after its gradient pass): a block is quiet when the amplitude bound over
// its 3 x 3 x 3 block neighbourhood (every cell within 5 of it) keeps |grad u| = sqrt(ux^2 + uy^2
// + uz^2) of each of its cells below the scalar operator's threshold rd(0.01 G_k): there S = 1
// without reading eps - del. The bound on each component is the gradient pass's (grad_bound:
// the smaller of the one from max |u| and the one from the previous step's components plus the
// increment), with the update's first derivatives (sum c1 (u+m - u-m), then 1/d) and a margin
// for their rounding and the hardware square root.
//
// The same pass keeps the touched bit: set once the neighbourhood's maximum of |u_k| is nonzero
// or the block meets the source box, and never cleared until the reset. A block without it (at
// rest) has u = 0 within 5 cells at every step so far, so u_{k-1}, u_k and the PML state
// (driven only by derivatives of u at the cell) are 0 there and the step would write zeros over
// zeros: the updates skip it. The history is whole only while every step since the reset ran
// this pass (*hist nonzero; an update launched without the map clears it): otherwise every
// block counts as touched.
struct SrcBox {
  int on, z0, z1, x0, x1, y0, y1;  // inclusive bounds
};

// The quiet and rest maps of one grid at step kstep, blocks b = first, first + stride, ...
__device__ __forceinline__ void quiet_blocks(const AmpGrid &ag, const Geo &g, const Coef &c, const unsigned *gmax,
                                             int kstep, const unsigned *hist, const SrcBox &sb, unsigned first,
                                             unsigned stride)
{
  const unsigned *bb = ag.buf(kstep);
  const float thr = s_threshold(__uint_as_float(gmax[kstep]));
  const bool whole = *hist != 0u;
  float s1 = 0.f;
#pragma unroll
  for (int k = 0; k < H; k++) s1 += fabsf(c.c1[k]);
  const float fac = 2.f * s1 * fmaxf(c.ix, fmaxf(c.iy, c.iz));
  const unsigned *db = ag.dbuf(kstep), *cb = ag.cbuf(kstep);
  for (unsigned b = first; b < ag.nblk; b += stride) {
    const int by = (int)(b % ag.nby), t = (int)(b / ag.nby), bx = t % ag.nbx, bz = t / ag.nbx;
    // the 3 x 3 x 3 neighbourhood with indices clamped to the grid (a repeated block does not
    // change a maximum), unrolled so that all 54 loads are in flight at once
    unsigned m = 0u, d = 0u;
#pragma unroll
    for (int i = 0; i < 27; i++) {
      const int z = min(max(bz + i / 9 - 1, 0), ag.nbz - 1);
      const int x = min(max(bx + (i / 3) % 3 - 1, 0), ag.nbx - 1);
      const int y = min(max(by + i % 3 - 1, 0), ag.nby - 1);
      const unsigned o = ((unsigned)z * ag.nbx + x) * ag.nby + y;
      m = max(m, bb[o]);
      d = max(d, AmpGrid::pml_block(g, z, x, y) ? 0x7f800000u : db[o]);
    }
    const unsigned cown = AmpGrid::pml_block(g, bz, bx, by) ? 0x7f800000u : cb[b];
    // |grad u| <= sqrt(3) times the largest component (1e-4 for the hardware square root)
    const float gb = grad_bound(m, d, cown, fac) * 1.7320509f * (1.f + 1.0e-4f);
    const bool src = sb.on && sb.z0 < ((bz + 1) << AB_Z) && sb.z1 >= (bz << AB_Z) && sb.x0 < ((bx + 1) << AB_X) &&
                     sb.x1 >= (bx << AB_X) && sb.y0 < ((by + 1) << AB_Y) && sb.y1 >= (by << AB_Y);
    const unsigned qb = ag.qi(bz, bx, by);
    const bool touched = !whole || (ag.q[qb] & 2) != 0 || m != 0u || src;
    const bool quiet = gb < thr;
    ag.q[qb] = (unsigned char)((quiet ? 1 : 0) | (touched ? 2 : 0));
  }
}

// Clears the slot of block maxima that the updates of step k+1 fill (for steps without a bound
// pass, whose k_gbound does it otherwise).
static __global__ void k_amp_clear(AmpGrid ag, Step st)
{
  const int kstep = st.get();
  unsigned *clr = ag.buf(kstep + 2), *clrd = ag.dbuf(kstep + 2), *clrc = ag.cbuf(kstep + 2);
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < ag.nblk; i += gridDim.x * blockDim.x) {
    clr[i] = 0u;
    clrd[i] = 0u;
    clrc[i] = 0u;
  }
}

static __global__ void k_quiet(AmpGrid ag, Geo g, Coef c, const unsigned *gmax, Step st, const unsigned *hist, SrcBox sb)
{
  quiet_blocks(ag, g, c, gmax, st.get(), hist, sb, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// The maps of every shot of a batch (blockIdx.y: shot; shots without them leave at once).
struct QuietArgs {
  AmpGrid ag;  // ag.q == nullptr: no maps for this shot
  Geo g;
  Coef c;
  const unsigned *gmax;
  const unsigned *hist;
  SrcBox sb;
};

static __global__ void k_quiet_batch(const QuietArgs *A, Step st)
{
  const QuietArgs &a = A[blockIdx.y];
  if (!a.ag.q || !a.ag.b) return;
  quiet_blocks(a.ag, a.g, a.c, a.gmax, st.get(), a.hist, a.sb, blockIdx.x * blockDim.x + threadIdx.x,
               gridDim.x * blockDim.x);
}

// Batched launches (several grids in one launch): item .z packs z0 | z1 << 16 and .w is the
// index of the grid's argument block.
__device__ __forceinline__ int4 batch_item(const int4 *work, int &shot)
{
  const int4 w = work[blockIdx.x];
  shot = w.w;
  return make_int4(w.x, w.y, w.z & 0xffff, (int)((unsigned)w.z >> 16));
}

template <class S, class T>
__global__ __launch_bounds__(T::NT) void k_gradmax(GradArgs<S> a)
{
  __shared__ float sm[2][T::LX + 1][T::LY];
  gradmax_tile<S, T>(a, a.work[blockIdx.x], a.step.get(), sm, (int)blockIdx.x);
}

// First part of the gradient pass with the amplitude bound: the tile that held the largest
// gradient component at step k-1, split in z over the grid, evaluated exactly, so that the
// skipping pass starts from a G_k close to the final one. It also clears the block maxima that
// the updates of step k+1 fill ((k+2) % 3) and the key of this step's largest tile.
// The bound pass's work on one workgroup: the clearing share and one z piece of tile w.
template <class S, class T>
__device__ __forceinline__ void gbound_tile(const GradArgs<S> &a, int4 w, int kstep, TileLds<T> sm)
{
  const unsigned tid = threadIdx.y * TY + threadIdx.x;
  unsigned *clr = a.amp.buf(kstep + 2), *clrd = a.amp.dbuf(kstep + 2), *clrc = a.amp.cbuf(kstep + 2);
  for (unsigned i = blockIdx.x * T::NT + tid; i < a.amp.nblk; i += gridDim.x * T::NT) {
    clr[i] = 0u;
    clrd[i] = 0u;
    clrc[i] = 0u;
  }
  if (blockIdx.x == 0 && tid == 0) {
    a.argkey[kstep & 1] = 0ull;
    a.gseed[(kstep + 1) & 1] = 0u;  // the next step's seed; this step's was cleared by step k-1
  }
  const int len = w.w - w.z;
  const int c = (len + (int)gridDim.x - 1) / (int)gridDim.x;
  const int s0 = w.z + (int)blockIdx.x * c, s1 = min(w.w, s0 + c);
  if (s0 >= s1) return;
  w.z = s0;
  w.w = s1;
  GradArgs<S> b = a;
  b.skip = 0;
  b.argkey = nullptr;
  b.gseed_out = a.gseed + (kstep & 1);
  gradmax_tile<S, T>(b, w, kstep, sm);
}

template <class S, class T>
__global__ __launch_bounds__(T::NT) void k_gbound(GradArgs<S> a)
{
  __shared__ float sm[2][T::LX + 1][T::LY];
  const int kstep = a.step.get();
  const unsigned long long key = a.argkey[(kstep + 1) & 1];  // written by step k-1
  int item = (int)(key & 0xffffffffull);
  if (item < 0 || item >= a.nwork) item = 0;
  gbound_tile<S, T>(a, a.work[item], kstep, sm);
}

// The bound pass of every shot of a batch (blockIdx.y: shot) over the batch's shared gradient
// list, whose items carry their shot: a key that names another shot's item (or none) falls back
// to one of the shot's own. Shots without block maxima leave at once.
template <class S, class T>
__global__ __launch_bounds__(T::NT) void k_gbound_batch(const GradArgs<S> *A, const int4 *work, int nwork,
                                                        const int *first, Step st)
{
  __shared__ float sm[2][T::LX + 1][T::LY];
  const int shot = (int)blockIdx.y;
  const GradArgs<S> a = A[shot];
  if (!a.amp.b || !a.argkey) return;
  const int kstep = st.get();
  const unsigned long long key = a.argkey[(kstep + 1) & 1];
  int item = (int)(key & 0xffffffffull);
  if (item < 0 || item >= nwork || work[item].w != shot) item = first[shot];
  // a shot without an item of its own gets no seed (nothing skipped), but its slot and next
  // seed are cleared all the same (an empty z range)
  const bool own = item >= 0 && item < nwork && work[item].w == shot;
  const int4 v = own ? work[item] : make_int4(0, 0, 0, 0);
  gbound_tile<S, T>(a, make_int4(v.x, v.y, v.z & 0xffff, (int)((unsigned)v.z >> 16)), kstep, sm);
}

// Marking pass of a step whose predicted threshold missed (gmax_mode 2); see gradmax_tile.
template <class S, class T>
__global__ __launch_bounds__(T::NT) void k_spec_mark(GradArgs<S> a)
{
  __shared__ float sm[2][T::LX + 1][T::LY];
  gradmax_tile<S, T, true>(a, a.work[blockIdx.x], a.step.get(), sm);
}

// skip 0: every tile evaluated (the shots' block maxima do not cover u); otherwise each shot's
// own argument block decides (its maxima and seed from k_gbound_batch)
template <class S, class T>
__global__ __launch_bounds__(T::NT) void k_gradmax_batch(const GradArgs<S> *A, const int4 *work, Step st, int skip)
{
  int shot;
  const int4 w = batch_item(work, shot);
  GradArgs<S> a = A[shot];
  if (!skip) a.skip = 0;
  __shared__ float sm[2][T::LX + 1][T::LY];
  gradmax_tile<S, T>(a, w, st.get(), sm, (int)blockIdx.x);
}


// Update kinds (the PML template parameter of the update kernels): 0 interior tiles, 1 PML
// tiles over the full-grid arrays, 2 folded y-face tiles over the compact arrays (YCompact).

// The per-cell arrays a tile reads besides u_cur (model; u_prev or the PML state), with their
// plane stride in cells.
template <class S>
struct CellSrc {
  const float *v2, *e, *ed;
  const typename S::T *uprev;  // interior tiles: u at step k-1
  typename S::T *w1p, *w3p, *pp;
  const typename S::T *w1c, *w3c, *pc;
  typename S::T *w21;
  size_t pln;
};

// The arrays of a tile of update kind PML: the full-grid arrays, read at the lane's own offsets
// (Lane::off), or for kind 2 the compact arrays, read at the offsets this writes into yoff
// (bytes of 4-byte cells; folded tile: lanes 0..31 hold the low face, 32..63 the high face).
template <class S, class T, int PML>
__device__ __forceinline__ CellSrc<S> cell_src(const UpdArgs<S> &a, const Lane<T> &L, int x0, int x1,
                                               unsigned (&yoff)[T::RX])
{
  if constexpr (PML == 2) {
    const YCompact<S> &c = a.yc;
    const int col = (L.ty >> 5) ? c.yw + (L.yc - (a.g.ny - a.g.pml)) : L.yc - H;
#pragma unroll
    for (int r = 0; r < T::RX; r++) {
      const int x = min(x0 + L.row(r), x1 - 1);
      yoff[r] = 4u * ((unsigned)(x - a.g.pml) * (unsigned)(2 * c.yw) + (unsigned)col);
    }
    return CellSrc<S>{c.vdt2, c.eps, c.epsd, nullptr, c.w1p, c.w3p, c.pp, c.w1c, c.w3c, c.pc, c.w21, c.pln};
  } else {
    return CellSrc<S>{a.vdt2, a.eps, a.epsd, a.unext, a.w1p, a.w3p, a.pp, a.w1c, a.w3c, a.pc, a.w21, a.g.pln};
  }
}

// Per-cell inputs of k_update other than u_cur: model, and either u_prev (interior tiles) or
// the seven PML state values (PML tiles).
template <class S, class T, int PML>
struct CellLoad {
  static constexpr int NS = PML ? 7 : 1;
  float v2[T::RX], e[T::RX], ed[T::RX];
  float st[T::RX][NS];  // interior: st[0] = u_prev; PML: w1p w1c pp pc w3p w3c w21

  // ed_on false (uniform): every cell of the plane keeps S = 1, which does not read eps - del
  __device__ __forceinline__ void issue(const CellSrc<S> &c, const unsigned (&offs)[T::RX], int z, bool ed_on = true)
  {
    const size_t zo = (size_t)z * c.pln;
    const float *pv = c.v2 + zo, *pe = c.e + zo, *pd = c.ed + zo;
#pragma unroll
    for (int r = 0; r < T::RX; r++) {
      v2[r] = ld_model(pv, offs[r]);
      e[r] = ld_model(pe, offs[r]);
    }
    if (ed_on) {
#pragma unroll
      for (int r = 0; r < T::RX; r++) ed[r] = ld_model(pd, offs[r]);
    } else {
#pragma unroll
      for (int r = 0; r < T::RX; r++) ed[r] = 0.f;
    }
    if constexpr (!PML) {
      const typename S::T *pu = c.uprev + zo;
#pragma unroll
      for (int r = 0; r < T::RX; r++) st[r][0] = S::ld_stream(pu, offs[r]);
    } else {
      const typename S::T *src[7] = {c.w1p + zo, c.w1c + zo, c.pp + zo, c.pc + zo, c.w3p + zo, c.w3c + zo, c.w21 + zo};
#pragma unroll
      for (int j = 0; j < 7; j++)
#pragma unroll
        for (int r = 0; r < T::RX; r++) st[r][j] = S::ld(src[j], offs[r]);  // cached: 6% faster than non-temporal
    }
  }
};

// One work item of the update pass (w: packed tile, kstep: time step).
// SPEC (gmax_mode 2): the threshold is predicted (spec_band), cells in the band are marked in
// a.fixmask, and the tile's maximum gradient component goes into G_k.
template <class S, class T, int PML, bool SPEC = false>
__device__ __forceinline__ void update_tile(const UpdArgs<S> &a, int4 w, int kstep, TileLds<T> sm)
{
  const Geo &g = a.g;
  int x0, x1, y0, y1, z0, z1;
  bool fold;
  unpack(w, x0, x1, y0, y1, z0, z1, fold);
  if (a.rest_off && blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) *a.rest_off = 0u;
  if (z0 >= z1) return;
  Lane<T> L(g, x0, x1, y0, y1, fold);
  // per-cell arrays and the lane's offsets into them (kind 2: compact y-face arrays; the host
  // gives kind 2 folded tiles only)
  unsigned yoff[T::RX];
  const CellSrc<S> cs = cell_src<S, T, PML>(a, L, x0, x1, yoff);
  const unsigned(&coff)[T::RX] = PML == 2 ? yoff : L.off;
  const int y = L.y;
  const bool ycmp = L.yon;
  const int nx = g.nx, ny = g.ny, nz = g.nz, pml = g.pml;
  bool rows_on[T::RX];
#pragma unroll
  for (int r = 0; r < T::RX; r++) rows_on[r] = ycmp && x0 + L.row(r) < x1;

  const float wk = a.wav[kstep];
  float *const rec = a.rec + (size_t)kstep * (g.nx - 2 * g.pml) * (g.ny - 2 * g.pml);
  // threshold of the scalar operator: from G_k, or predicted (SPEC) with the band (blo, bhi]
  // of cells to mark
  float thr, blo = 0.f, bhi = 0.f;
  if constexpr (SPEC) {
    const SpecBand sb = spec_band(a.gpred, kstep);
    thr = sb.ts;
    blo = sb.lo;
    bhi = sb.hi;
  } else {
    thr = s_threshold(__uint_as_float(a.gmax[kstep]));
  }
  float gm = 0.f;          // SPEC: G_k over this tile's cells
  bool nonfinite = false;  // SPEC: a non-finite gradient component seen
  const bool amp_on = a.amp.b != nullptr;  // block maxima of |u_next| for the next gradient pass
  // running maxima (float bits) of |u_next|, and on interior tiles of |u_next - u_cur| and of
  // |grad u_cur|
  unsigned am = 0u, ad = 0u, ac = 0u;
  // source box and receiver plane tests, hoisted to the workgroup and the plane (uniform)
  const bool src_tile = a.src_on && a.sb_x0 < x1 && a.sb_x1 >= x0 &&
                        (fold || (a.sb_y0 < y1 && a.sb_y1 >= y0)) && a.sb_z0 < z1 && a.sb_z1 >= z0;
  const bool rec_tile = a.sz >= z0 && a.sz < z1;

  // PML face family of the tile (uniform) and the per-row / per-lane coefficients of x / y faces
  int fam = 0;
  float4 trow[T::RX];
  float sgn = 1.f;
  if constexpr (PML) {
    fam = PML == 2 ? 1 : (x1 <= pml || x0 >= nx - pml) ? 0 : ((y1 <= pml || y0 >= ny - pml) ? 1 : 2);
    if (fam == 0) {
      sgn = x1 <= pml ? 1.f : -1.f;
#pragma unroll
      for (int r = 0; r < T::RX; r++) {
        const int x = min(x0 + L.row(r), x1 - 1);
        trow[r] = a.pmltab[x < pml ? pml - x : x - nx + pml];
      }
    } else if (fam == 1) {
      const int yc = L.yc;  // per lane: a folded tile holds both y faces
      sgn = yc < pml ? 1.f : -1.f;
      const float4 t = a.pmltab[pml + (yc < pml ? pml - yc : yc - ny + pml)];
#pragma unroll
      for (int r = 0; r < T::RX; r++) trow[r] = t;
    } else {
      sgn = z1 <= pml ? 1.f : -1.f;
    }
  }

  // The map of this step (when on), for the tile's z blocks: bit i of qm / rm is z block
  // (z0 >> AB_Z) + i (i < 128; blocks past that count as neither). Lanes take one z block each
  // and AND the map over the tile's x and y blocks (a folded tile: over both of its y strips).
  // Quiet (qm): every cell has |grad u| below the threshold, so S = 1 and eps - del is not
  // loaded. At rest (rm): the step would write zeros over zeros, so the planes of such blocks at
  // either end of the item are left out (and the whole item when every block is at rest).
  unsigned long long qm[2] = {0ull, 0ull}, rm[2] = {0ull, 0ull};
  const int zq0 = z0 >> AB_Z, za = z0, zc = z1;
  const int nzb = ((z1 - 1) >> AB_Z) - zq0 + 1;
  if (a.amp.q) {
    const int xb0 = x0 >> AB_X, xb1 = (x1 - 1) >> AB_X;
    const int ya0 = y0 >> AB_Y, ya1 = (y1 - 1) >> AB_Y;
    const int yh = g.ny - g.pml;  // a folded tile's second strip starts here
    const int yc0 = fold ? yh >> AB_Y : ya0, yc1 = fold ? (yh + (y1 - y0) - 1) >> AB_Y : ya1;
    // every block of the tile in one batch of loads (indices clamped to the tile's last block:
    // repeats do not change an AND), so an item waits for the map once
    constexpr int NXB = T::TX / (1 << AB_X) + 1, NYB = TY / (1 << AB_Y) + 1;
    const int ns = fold ? 2 : 1;  // y strips (uniform)
#pragma unroll
    for (int h = 0; h < 2; h++) {
      if (h == 1 && nzb <= TY) break;  // uniform
      const int i = __lane_id() + TY * h;
      const int zb = min(zq0 + i, a.amp.nbz - 1);
      unsigned v = 3u;  // AND of (bit 0 quiet, bit 1 not touched)
#pragma unroll
      for (int st = 0; st < 2; st++) {
        if (st >= ns) break;
        const int yb0 = st ? yc0 : ya0, yb1 = st ? yc1 : ya1;
#pragma unroll
        for (int xi = 0; xi < NXB; xi++)
#pragma unroll
          for (int yi = 0; yi < NYB; yi++) {
            const unsigned m = a.amp.q[a.amp.qi(zb, min(xb0 + xi, xb1), min(yb0 + yi, yb1))];
            v &= (m & 1u) | (~m & 2u);
          }
      }
      qm[h] = __ballot(i < nzb && (v & 1u) != 0u);
      rm[h] = __ballot(i < nzb && (v & 2u) != 0u);
    }
    if (rm[0] | rm[1]) {
      // z blocks not at rest (blocks past the 128th always are)
      const unsigned long long v0 = nzb >= 64 ? ~0ull : (1ull << nzb) - 1ull;
      const unsigned long long v1 = nzb >= 128 ? ~0ull : nzb <= 64 ? 0ull : (1ull << (nzb - 64)) - 1ull;
      const unsigned long long n0 = ~rm[0] & v0, n1 = ~rm[1] & v1;
      if (!(n0 | n1) && nzb <= 128) {
        z1 = z0;  // the whole item is at rest
      } else {
        const int first = n0 ? __builtin_ctzll(n0) : n1 ? 64 + __builtin_ctzll(n1) : 128;
        z0 = max(z0, (zq0 + first) << AB_Z);
        if (nzb <= 128) {
          const int last = n1 ? 127 - __builtin_clzll(n1) : 63 - __builtin_clzll(n0);
          z1 = min(z1, (zq0 + last + 1) << AB_Z);
        }
      }
    }
  }
  if (a.ucount && threadIdx.x == 0 && threadIdx.y == 0) {
    // instrumented steps, one thread per workgroup: cells of the item, cells left out at rest,
    // cells without eps - del (a folded tile holds two y strips)
    const unsigned cpp = (unsigned)(x1 - x0) * (unsigned)(y1 - y0) * (fold ? 2u : 1u);
    int nq = 0;  // planes of [z0, z1) in quiet blocks
    for (int i = 0; i < min(nzb, 128); i++)
      if (((i < 64 ? qm[0] >> i : qm[1] >> (i - 64)) & 1ull) != 0ull)
        nq += max(0, min(z1, (zq0 + i + 1) << AB_Z) - max(z0, (zq0 + i) << AB_Z));
    unsigned long long *const uc = cnt_slot(a.ucount) + 3 * PML;
    atomicAdd(uc, (unsigned long long)cpp * (unsigned)(zc - za));
    atomicAdd(uc + 1, (unsigned long long)cpp * (unsigned)((zc - za) - (z1 - z0)));
    atomicAdd(uc + 2, (unsigned long long)cpp * (unsigned)nq);
  }
  if (z0 >= z1) return;  // uniform over the workgroup
  auto ed_needed = [&](int z) {
    const int i = (z >> AB_Z) - zq0;
    return i >= 128 || !(((i < 64 ? qm[0] >> i : qm[1] >> (i - 64)) & 1ull) != 0ull);
  };
  float q[T::RX][QN];
  prime_queue<S, T>(q, a.ucur, g, L, z0);
  PlaneLoad<S, T> P;
  P.issue(a.ucur, g, L, z0);
  // per-cell loads, double-buffered: the interior march (unrolled by 22) addresses the buffer
  // of plane `it` as CL[it & 1] at compile time; the rolled PML march copies the next into
  // the current
  CellLoad<S, T, PML> CL[2];
  CL[0].issue(cs, coff, z0, ed_needed(z0));
  float centre[T::RX];
#pragma unroll
  for (int r = 0; r < T::RX; r++) {
    q[r][2 * H] = P.qn[r];
    centre[r] = q[r][H];
  }
  P.stage(sm[0], centre, L);
  // interior: the workgroup's two LDS sets of block maxima (amp_lds_add), cleared before the
  // march's first barrier
  __shared__ unsigned amx[2][3][4];
  const bool w0 = __builtin_amdgcn_readfirstlane(threadIdx.y) == 0;
  if constexpr (!PML) {
    if (amp_on && w0 && __lane_id() < 24) (&amx[0][0][0])[__lane_id()] = 0u;
  }
  __syncthreads();
  int b = 0;
  int zb = z0;
  auto plane = [&](auto itc) {
    constexpr int it = decltype(itc)::value;
    {
      const int z = zb + it;
      const bool more = z + 1 < z1;
      L.refresh();
      if constexpr (PML == 2) {
#pragma unroll
        for (int r = 0; r < T::RX; r++) asm volatile("" : "+v"(yoff[r]));
      }
      CellLoad<S, T, PML> &C = CL[PML ? 0 : (it & 1)];
      CellLoad<S, T, PML> &Cn = CL[PML ? 1 : ((it + 1) & 1)];
      if (more) {
        P.issue(a.ucur, g, L, z + 1);
        Cn.issue(cs, coff, z + 1, ed_needed(z + 1));
      }
      // interior: unrolled by 22, the buffer of plane `it` is it & 1 at compile time
      float (*s)[T::LY] = sm[PML ? b : (it & 1)];
      const bool src_plane = src_tile && z >= a.sb_z0 && z <= a.sb_z1;
      const bool rec_plane = rec_tile && z == a.sz;
      float4 tz;
      if constexpr (PML) {
        if (fam == 2) tz = a.pmltab[2 * pml + (z < pml ? pml - z : z - nz + pml)];
      }
      const size_t zo = (size_t)z * g.pln;
#pragma unroll
      for (int r = 0; r < T::RX; r++) {
        if (!rows_on[r]) continue;
        const int X = H + L.row(r), Y = L.Y;
        const int x = x0 + L.row(r);
        const float c0 = q[r][(H + it) % QN];
        // second derivatives as sum_m c2_m ((u+m - 2 u0) + u-m) (d2tap), first derivatives as
        // sum_m c1_m (u+m - u-m), from the same pair of values.
        float Lxx = 0.f, Lyy = 0.f, Lzz = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
        const float c02 = 2.f * c0;
#pragma unroll
        for (int k = 1; k <= H; k++) {
          const float ax = s[X + k][Y], bx = s[X - k][Y];
          const float ay = s[X][Y + k], by = s[X][Y - k];
          const float az = q[r][(H + k + it) % QN], bz = q[r][(H - k + it) % QN];
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
        const float e = C.e[r], ed = C.ed[r], v2 = C.v2[r];
        const float mode = cell_mode(ux, uy, uz);
        const float Sv = scalar_op_m(mode, ux, uy, uz, e, ed, thr);
        const float e1 = fmaf(2.f, e, 1.f);
        const unsigned o = L.off[r];
        if constexpr (SPEC) {
          gm = fmaxf(gm, fmaxf(fabsf(ux), fmaxf(fabsf(uy), fabsf(uz))));
          nonfinite |= !(fabsf(ux) + fabsf(uy) + fabsf(uz) <= 3.4e38f);
          if (mode > blo && mode <= bhi) mark_cell(a.fixmask, (unsigned)z * g.pln + (o >> 2));
        }
        float un;
        if constexpr (!PML) {
          // GPU_kernel.cu:558-565
          un = (2.f * c0 - C.st[r][0]) + ((e1 * (Uxx + Uyy) + Uzz) * Sv) * v2;
        } else {
          // GPU_kernel.cu:566-637
          const float4 t = fam == 2 ? tz : trow[r];
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
          const size_t zs = (size_t)z * cs.pln;
          const unsigned os = coff[r];
          S::st(cs.w1p + zs, os, w1n);
          S::st(cs.pp + zs, os, pn);
          S::st(cs.w21 + zs, os, w21n);
          S::st(cs.w3p + zs, os, w3n);
          un = w1n + w21n + w3n;
        }
        if (src_plane && x >= a.sb_x0 && x <= a.sb_x1 && y >= a.sb_y0 && y <= a.sb_y1) {
          const int bx = a.sb_x1 - a.sb_x0 + 1, by = a.sb_y1 - a.sb_y0 + 1;
          const float sv = a.sbox[((long)(z - a.sb_z0) * bx + (x - a.sb_x0)) * by + (y - a.sb_y0)];
          un = fmaf(sv, wk, un);  // u3 += source*wavelet, contracted as the CUDA build does
        }
        un = S::rnd(un);
        S::st(a.unext + zo, o, un);
        // unconditional: a branch per cell costs more
        am = max(am, __float_as_uint(fabsf(un)));
        if constexpr (!PML) {
          // |grad u| (mode, computed for S) bounds every component; integer maxima of float
          // bits need no NaN canonicalisation (mode is never negative)
          ad = max(ad, __float_as_uint(fabsf(un - c0)));
          ac = max(ac, __float_as_uint(mode));
        }
        if (rec_plane && x >= pml && x < nx - pml && y >= pml && y < ny - pml)
          rec[(x - pml) * (ny - 2 * pml) + (y - pml)] = un;
      }
      if constexpr (PML != 0) {
        // the rolled PML march flushes at the end of every z block
        if (amp_on && ((z & ((1 << AB_Z) - 1)) == (1 << AB_Z) - 1 || !more)) amp_flush(a.amp, kstep, z, z + 1, x0, x1, y, am, ad, ac);
      }
      if (more) {
#pragma unroll
        for (int r = 0; r < T::RX; r++) {
          q[r][it % QN] = P.qn[r];
          centre[r] = q[r][(H + 1 + it) % QN];
        }
        P.stage(sm[PML ? (b ^ 1) : ((it + 1) & 1)], centre, L);
        if constexpr (PML) C = Cn;
        __syncthreads();
        b ^= 1;
      }
    }
  };
  if constexpr (!PML) {
    // block maxima once per unrolled stretch (its planes may cover up to four z blocks), through
    // the LDS sets (amp_lds_add): a set is moved out at the end of the next stretch, whose planes
    // hold a barrier (or get one here), and cleared before it is filled again two stretches on
    int set = 0;  // this stretch's set; the previous stretch (full length) used the other
    for (; zb < z1; zb += 2 * QN) {
      Unroll<0, 2 * QN>::run(plane, z1 - zb);
      if (amp_on) {
        if (z1 - zb < 2) __syncthreads();  // a stretch of one plane passed no barrier
        if (zb > z0 && w0) amp_lds_flush(a.amp, kstep, amx[set ^ 1], zb - 2 * QN, zb, x0, x1, y0);
        amp_lds_add(amx[set], y0, y, a.amp.nby, am, ad, ac);
        set ^= 1;
      }
    }
    if (amp_on) {
      // the last stretch: from the last multiple of the stretch length past z0 to z1
      __syncthreads();
      if (w0) amp_lds_flush(a.amp, kstep, amx[set ^ 1], zb - 2 * QN, z1, x0, x1, y0);
    }
  } else {
    // the PML kernel carries seven more state values per row; a rolled loop with a shifted
    // queue keeps it free of spills
    for (; zb < z1; zb++) {
      plane(ic<0>{});
#pragma unroll
      for (int r = 0; r < T::RX; r++) {
        const float f = q[r][0];
#pragma unroll
        for (int j = 0; j < 2 * H; j++) q[r][j] = q[r][j + 1];
        q[r][2 * H] = f;
      }
    }
  }
  if constexpr (SPEC) reduce_gmax<T>(gm, nonfinite, L.tx, L.ty, a.gmax + kstep, a.flag);
}

template <class S, class T, int PML, bool SPEC = false>
__global__ __launch_bounds__(T::NT) void k_update(UpdArgs<S> a)
{
  __shared__ float sm[2][T::LX + 1][T::LY];
  update_tile<S, T, PML, SPEC>(a, a.work[blockIdx.x], a.step.get(), sm);
}

// maps 0: the shots' quiet and rest maps are not valid for this step (not read)
template <class S, class T, int PML>
__global__ __launch_bounds__(T::NT) void k_update_batch(const UpdArgs<S> *A, const int4 *work, Step st, int maps)
{
  int shot;
  const int4 w = batch_item(work, shot);
  UpdArgs<S> a = A[shot];  // one copy: stores to the fields must not force reloads
  if (!maps) a.amp.q = nullptr;
  __shared__ float sm[2][T::LX + 1][T::LY];
  update_tile<S, T, PML>(a, w, st.get(), sm);
}

// Source cells outside [5,n-5)^3 are never updated by the stencil; the reference's u3 keeps
// accumulating source*wavelet there (GPU_kernel.cu:720-731 with the copy rotation of
// GPU_kernel.cu:749-750), i.e. u_next = u_cur + s*w.
template <class S>
__global__ void k_rim_source(const typename S::T *ucur, typename S::T *unext, Geo g, int z0, int z1, int x0, int x1,
                             int y0, int y1, const float *sbox, const float *wav, Step step, AmpGrid amp)
{
  const int kstep = step.get();
  const float wk = wav[kstep];
  const int bx = x1 - x0 + 1, by = y1 - y0 + 1, bz = z1 - z0 + 1;
  const long n = (long)bx * by * bz;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int yy = (int)(t % by), xx = (int)((t / by) % bx), zz = (int)(t / ((long)by * bx));
    const int x = x0 + xx, y = y0 + yy, z = z0 + zz;
    const bool r5 = x >= H && x < g.nx - H && y >= H && y < g.ny - H && z >= H && z < g.nz - H;
    if (r5) continue;
    const unsigned b = 4u * ((unsigned)z * g.pln + (unsigned)x * g.ys + (unsigned)y);
    const float uc = S::ld(ucur, b);
    const float v = S::rnd(fmaf(sbox[t], wk, uc));
    S::st(unext, b, v);
    if (amp.b) {
      // the rim has no gradient of its own (it is outside [5, n-5)^3), only its increment
      const unsigned o = ((unsigned)(z >> AB_Z) * amp.nbx + (unsigned)(x >> AB_X)) * amp.nby + (y >> AB_Y);
      atomicMax(amp.buf(kstep + 1) + o, __float_as_uint(fabsf(v)));
      atomicMax(amp.dbuf(kstep + 1) + o, __float_as_uint(fabsf(v - uc)));
    }
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
      S::ld(u, 4u * ((unsigned)(z + g.pml) * g.pln + (unsigned)(x + g.pml) * g.ys + (unsigned)(y + g.pml)));
}

// The scalar-operator field S of step k over the whole grid [nz][nx][ny] (1 outside
// [5,n-5)^3), from u_k and G_k with the update kernel's arithmetic, so it is the S that step
// used. The reference computes it every step into a full-grid array and has its writer
// commented out (3dPureqPModeling_single_shot.cu:381-409); here it is formed only for snapshot
// steps that ask for it. One thread per cell of the dense output; the stencil reads go through
// the caches.
template <class S>
__global__ void k_sfield(const typename S::T *u, const float *eps, const float *epsd, const unsigned *gmax, Step st,
                         Geo g, Coef c, float *out)
{
  const float thr = s_threshold(__uint_as_float(gmax[st.get()]));
  const long dpln = (long)g.nx * g.ny, n = dpln * g.nz;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int z = (int)(i / dpln);
    const int rem = (int)(i - (long)z * dpln);
    const int x = rem / g.ny, y = rem - x * g.ny;
    const unsigned f = (unsigned)z * g.pln + (unsigned)x * g.ys + (unsigned)y;  // cell in the device layout
    float Sv = 1.f;
    if (z >= H && z < g.nz - H && x >= H && x < g.nx - H && y >= H && y < g.ny - H) {
      const unsigned b = 4u * f, bx = 4u * (unsigned)g.ys, bz = 4u * g.pln;
      float ux = 0.f, uy = 0.f, uz = 0.f;
#pragma unroll
      for (int k = 1; k <= H; k++) {
        ux = fmaf(c.c1[k - 1], S::ld(u, b + k * bx) - S::ld(u, b - k * bx), ux);
        uy = fmaf(c.c1[k - 1], S::ld(u, b + 4u * k) - S::ld(u, b - 4u * k), uy);
        uz = fmaf(c.c1[k - 1], S::ld(u, b + k * bz) - S::ld(u, b - k * bz), uz);
      }
      ux *= c.ix; uy *= c.iy; uz *= c.iz;
      Sv = scalar_op(ux, uy, uz, eps[f], epsd[f], thr);
    }
    out[i] = Sv;
  }
}

// Fix-up of a step taken with the predicted threshold ts (gmax_mode 2), after every update of
// the step has run and G_k is complete. Each thread takes words of the cell mask, clears them,
// and for each marked cell recomputes the update kernel's derivatives from u_k (bit for bit:
// same terms in the same order); where |grad u| > ts and |grad u| > tt = rd(0.01 G_k) disagree,
// the scalar operator that step used (S at ts) is replaced by S at tt: the new wavefield, the
// PML state and the record move by the change of S times the terms it multiplies (every
// update is linear in S). u_k, which the step does not write, is still in place.
template <class S>
__device__ __forceinline__ void fix_cell(const UpdArgs<S> &a, unsigned i, float ts, float tt, int k)
{
  const Geo &g = a.g;
  const int z = (int)(i / g.pln);
  const unsigned rem = i - (unsigned)z * g.pln;
  const int x = (int)(rem / (unsigned)g.ys), y = (int)(rem - (unsigned)x * g.ys);
  if (z < H || z >= g.nz - H || x < H || x >= g.nx - H || y < H || y >= g.ny - H) return;
  const unsigned b = 4u * i, bx = 4u * (unsigned)g.ys, bz = 4u * g.pln;
  const float c0 = S::ld(a.ucur, b);
  const float c02 = 2.f * c0;
  float Lxx = 0.f, Lyy = 0.f, Lzz = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
#pragma unroll
  for (int m = 1; m <= H; m++) {
    const float ax = S::ld(a.ucur, b + m * bx), bx_ = S::ld(a.ucur, b - m * bx);
    const float ay = S::ld(a.ucur, b + 4u * m), by = S::ld(a.ucur, b - 4u * m);
    const float az = S::ld(a.ucur, b + m * bz), bz_ = S::ld(a.ucur, b - m * bz);
    Lxx = d2tap(a.c.c2[m - 1], ax, bx_, c02, Lxx);
    Lyy = d2tap(a.c.c2[m - 1], ay, by, c02, Lyy);
    Lzz = d2tap(a.c.c2[m - 1], az, bz_, c02, Lzz);
    ux = fmaf(a.c.c1[m - 1], ax - bx_, ux);
    uy = fmaf(a.c.c1[m - 1], ay - by, uy);
    uz = fmaf(a.c.c1[m - 1], az - bz_, uz);
  }
  ux *= a.c.ix; uy *= a.c.iy; uz *= a.c.iz;
  const float mode = cell_mode(ux, uy, uz);
  if ((mode > ts) == (mode > tt)) return;
  const float Uxx = Lxx * a.c.ixx;
  const float Uyy = Lyy * a.c.iyy;
  const float Uzz = Lzz * a.c.izz;
  const float e = a.eps[i], ed = a.epsd[i], v2 = a.vdt2[i];
  const float dS = scalar_op_m(mode, ux, uy, uz, e, ed, tt) - scalar_op_m(mode, ux, uy, uz, e, ed, ts);
  const float e1 = fmaf(2.f, e, 1.f);
  const int pml = g.pml;
  const bool xf = x < pml || x >= g.nx - pml, yf = y < pml || y >= g.ny - pml, zf = z < pml || z >= g.nz - pml;
  const float un = S::ld(a.unext, b);
  float unn;
  if (!xf && !yf && !zf) {
    unn = S::rnd(fmaf((e1 * (Uxx + Uyy) + Uzz) * dS, v2, un));
  } else {
    // face priority x > y > z, as the PML tiles
    float4 t;
    float sgn, Unn, Tg, gn, dC;
    const float dB = dS * v2;
    if (xf) {
      t = a.pmltab[x < pml ? pml - x : x - g.nx + pml];
      sgn = x < pml ? 1.f : -1.f;
      dC = e1 * dB; Unn = Uxx; Tg = Uzz + e1 * Uyy; gn = ux;
    } else if (yf) {
      t = a.pmltab[pml + (y < pml ? pml - y : y - g.ny + pml)];
      sgn = y < pml ? 1.f : -1.f;
      dC = e1 * dB; Unn = Uyy; Tg = Uzz + e1 * Uxx; gn = uy;
    } else {
      t = a.pmltab[2 * pml + (z < pml ? pml - z : z - g.nz + pml)];
      sgn = z < pml ? 1.f : -1.f;
      dC = dB; Unn = Uzz; Tg = e1 * (Uxx + Uyy); gn = uz;
    }
    // state of a y-face cell in the compact arrays when the plan keeps them
    const bool cy = !xf && yf && a.yc.yw > 0;
    typename S::T *const w1a = cy ? a.yc.w1p : a.w1p, *const pa = cy ? a.yc.pp : a.pp;
    typename S::T *const w21a = cy ? a.yc.w21 : a.w21, *const w3a = cy ? a.yc.w3p : a.w3p;
    const unsigned bs = cy ? 4u * ycompact_cell(g.ny, pml, a.yc.yw, a.yc.pln, z, x, y) : b;
    const float w1 = S::ld(w1a, bs), p = S::ld(pa, bs), w21 = S::ld(w21a, bs), w3 = S::ld(w3a, bs);
    const float w1n = S::rnd(fmaf(dC, Unn, w1));
    const float pn = S::rnd(fmaf(sgn, dC * t.w * gn, p));
    const float w21n = S::rnd(fmaf(a.c.dt, pn - p, w21));
    const float w3n = S::rnd(fmaf(dB, Tg, w3));
    S::st(w1a, bs, w1n);
    S::st(pa, bs, pn);
    S::st(w21a, bs, w21n);
    S::st(w3a, bs, w3n);
    unn = S::rnd(un + ((w1n - w1) + (w21n - w21) + (w3n - w3)));
  }
  S::st(a.unext, b, unn);
  if (z == a.sz && x >= pml && x < g.nx - pml && y >= pml && y < g.ny - pml)
    a.rec[(size_t)k * (g.nx - 2 * pml) * (g.ny - 2 * pml) + (size_t)(x - pml) * (g.ny - 2 * pml) + (y - pml)] = unn;
}

// Prediction of G_{k+1} from G_0..G_k (double precision, deterministic): an AR(6) fit of
// log G over the last 32 steps (least squares through the normal equations, centred at
// log G_k, tiny ridge), which follows the few-step periodic jitter of G as wave crests cross
// grid points; with fewer steps, or a non-positive G in the window, G_k (G_k / G_{k-1})
// clamped to [1/2, 2]. The result is clamped to [G_k / 2, 2 G_k]. All threads of the calling
// workgroup (256) take part.
constexpr int AR_P = 6, AR_W = 32;
__device__ __forceinline__ void predict_next(const unsigned *gmax, float *gpred, int k, int nt)
{
  __shared__ double lg[AR_P + AR_W];
  __shared__ double sums[35];
  if (k + 1 >= nt) return;  // uniform
  const int t = threadIdx.x;
  const int n0 = k + 1 - AR_W - AR_P;
  const float gk = __uint_as_float(gmax[k]);
  int ok = n0 >= 0 && gk > 0.f && gk <= 3.4e38f;  // uniform
  if (ok && t < AR_P + AR_W) {
    const float gv = __uint_as_float(gmax[n0 + t]);
    if (gv > 0.f && gv <= 3.4e38f) lg[t] = log((double)gv) - log((double)gk);
    else ok = 0;
  }
  const bool ar = __syncthreads_and(ok);  // every G of the window usable; lg complete
  if (ar && t < 35) {
    // t < 28: entry (a, b), a <= b, of the 7x7 normal matrix; t >= 28: right-hand side a
    int a = 0, b = 0;
    if (t < 28) {
      int r = t;
      while (r >= 7 - a) { r -= 7 - a; a++; }
      b = a + r;
    } else {
      a = t - 28;
      b = -1;
    }
    double acc = 0.0;
    for (int j = 0; j < AR_W; j++) {
      const int n = AR_P + j;  // target index in the window
      const double xa = a < AR_P ? lg[n - 1 - a] : 1.0;
      const double xb = b < 0 ? lg[n] : (b < AR_P ? lg[n - 1 - b] : 1.0);
      acc = fma(xa, xb, acc);
    }
    sums[t] = acc;
  }
  __syncthreads();
  if (t != 0) return;
  double gp = gk;
  bool done = false;
  if (ar) {
    double M[7][8];
    int idx = 0;
    for (int a = 0; a < 7; a++)
      for (int b = a; b < 7; b++, idx++) M[a][b] = M[b][a] = sums[idx];
    double tr = 0.0;
    for (int a = 0; a < 7; a++) {
      M[a][7] = sums[28 + a];
      tr += M[a][a];
    }
    for (int a = 0; a < 7; a++) M[a][a] += 1e-9 * tr / 7.0;
    bool ok = tr > 0.0;
    for (int c = 0; c < 7 && ok; c++) {  // Gaussian elimination, partial pivoting
      int piv = c;
      for (int r = c + 1; r < 7; r++)
        if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
      if (!(fabs(M[piv][c]) > 1e-300)) { ok = false; break; }
      if (piv != c)
        for (int j = c; j < 8; j++) { const double x = M[c][j]; M[c][j] = M[piv][j]; M[piv][j] = x; }
      for (int r = c + 1; r < 7; r++) {
        const double f = M[r][c] / M[c][c];
        for (int j = c; j < 8; j++) M[r][j] -= f * M[c][j];
      }
    }
    if (ok) {
      double cf[7];
      for (int r = 6; r >= 0; r--) {
        double v = M[r][7];
        for (int j = r + 1; j < 7; j++) v -= M[r][j] * cf[j];
        cf[r] = v / M[r][r];
      }
      double L = cf[AR_P];
      for (int m = 0; m < AR_P; m++) L += cf[m] * lg[AR_P + AR_W - 1 - m];
      if (L == L && fabs(L) < 50.0) {
        gp = (double)gk * exp(L);
        done = true;
      }
    }
  }
  if (!done && k >= 1) {
    const float g1 = __uint_as_float(gmax[k - 1]);
    if (g1 > 0.f && gk > 0.f && gk <= 3.4e38f) gp = (double)gk * fmin(fmax((double)gk / (double)g1, 0.5), 2.0);
  }
  if (gk > 0.f && gk <= 3.4e38f) gp = fmin(fmax(gp, 0.5 * (double)gk), 2.0 * (double)gk);
  gpred[k + 1] = (float)gp;
}

template <class S>
__global__ __launch_bounds__(256) void k_spec_fix(UpdArgs<S> a, long nwords, int nt)
{
  const int k = a.step.get();
  if (blockIdx.x == 0) predict_next(a.gmax, a.gpred, k, nt);  // for the next step's updates
  const float ts = spec_band(a.gpred, k).ts, tt = s_threshold(__uint_as_float(a.gmax[k]));
  // four mask words per thread and iteration (one 16-byte load)
  const long nq = (nwords + 3) / 4;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nq; q += (long)gridDim.x * blockDim.x) {
    const long w0 = 4 * q;
    uint4 v;
    if (w0 + 4 <= nwords) {
      v = *reinterpret_cast<const uint4 *>(a.fixmask + w0);
    } else {
      v.x = a.fixmask[w0];
      v.y = w0 + 1 < nwords ? a.fixmask[w0 + 1] : 0u;
      v.z = w0 + 2 < nwords ? a.fixmask[w0 + 2] : 0u;
      v.w = 0u;
    }
    if (!(v.x | v.y | v.z | v.w)) continue;
    const unsigned wv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      unsigned bits = wv[i];
      if (!bits) continue;
      a.fixmask[w0 + i] = 0u;
      while (bits) {
        const int j = __builtin_ctz(bits);
        bits &= bits - 1u;
        fix_cell<S>(a, (unsigned)(w0 + i) * 32u + (unsigned)j, ts, tt, k);
      }
    }
  }
}

// Sets the device step index read by graph-replayed launches.
static __global__ void k_set_step(int *kdev, int k)
{
  if (threadIdx.x == 0) *kdev = k;
}

// Tile shapes the plan can launch (index = qp3d_config.tile_update / tile_grad - 1).
#define QP3D_TILES(X) X(8, 2) X(16, 1) X(4, 2) X(8, 1) X(12, 1) X(16, 2) X(4, 4)
constexpr int NTILES = 7;

}  // namespace qp3d
