// One persistent launch for many time steps of many shots (batched shots).
//
// The launch holds as many workgroups as the GPU keeps resident. Each workgroup takes work items
// from one queue in a fixed order (a global counter). A shot's items depend only on the same
// shot's previous phase: its update tiles of step k (interior, PML x/z faces, y faces, and the
// rim-source cells of a shot whose source reaches the frozen 5-cell rim) need G_k, i.e. all of
// its gradient tiles of step k; its gradient tiles of step k+1 need u_{k+1}, i.e. all of its
// update tiles of step k (which also ends every read of the slot that step k+1 overwrites). Two
// counters per shot, bumped once per finished item, carry these dependencies, so shots advance
// independently: no barrier across shots.
//
// The queue is skewed so that no phase waits on the one right before it: the shots fall in two
// groups (even and odd index), and queue slot q holds the gradient tiles of group A at step q,
// the update tiles of group B at step q-1, the update tiles of A at step q and the gradient tiles
// of B at step q; a run of steps [k0, k1) takes k1-k0+1 slots, items whose step falls outside
// the run being skipped. Every dependency then has a whole phase of the other group's work
// before it in the queue.
//
// Progress: an item waits only for items that come before it in the queue, and every such item
// was taken by a workgroup that is running, so every wait ends while the launch is healthy. A
// wait is bounded all the same (wall clock): past the bound the workgroup sets the error word
// and leaves, and every other workgroup leaves at its next wait or item.
//
// Visibility across workgroups (and XCDs, whose L2 caches are not coherent with each other):
// after an item every wave waits for its own stores to complete (they have reached its XCD's
// L2), then, after the barrier, wavefront 0 alone releases at device scope (one write-back of
// the L2) and signals; the waiting wavefront acquires at device scope once its counter reached
// the target (stale lines dropped before the item's loads). A fence per wave and item, each
// with a write-back and an invalidate, measured 2-4x slower.
#pragma once
#include "qp3d_kernels.hpp"

namespace qp3d {

// Item kinds of the persistent queue (bits 16..23 of an item's .w word; bits 0..15: shot; bit 24:
// the item belongs to the step before its queue slot).
enum PersistKind : int {
  PK_GRAD = 0,   // gradient-max tile
  PK_INT = 1,    // interior update tile
  PK_PML = 2,    // PML tile, x or z faces (full-grid state)
  PK_PMLY = 3,   // folded y-face tile over the compact arrays
  PK_PMLYF = 4,  // y-face tile over the full-grid state (pml > 32)
  PK_RIM = 5,    // source cells on the frozen rim
};

template <class S>
struct PersistArgs {
  const GradArgs<S> *ag[2];  // per slot parity (step & 1), one block per shot
  const UpdArgs<S> *au[2];
  const int4 *items;         // one queue slot's items, in queue order
  int ips;                   // items per slot
  int nshots;
  int k0, k1;                // time steps [k0, k1)
  const unsigned *nitems;    // [2][nshots]: gradient items, update items of one shot per step
  unsigned *ctr;             // [0]: queue; [1]: error word; [2 ..]: [2][nshots] items done
  long long wait_ticks;      // bound of one wait, in wall-clock ticks (wall_clock64)
};

// Waits until *p >= need, at most `ticks`; false (error word set, or the bound passed: then the
// error word is set here) means leave. Run by a whole wavefront: every lane loads the same word
// and the loop condition is the first lane's value, so the loop is uniform (a loop that only
// some lanes run, next to the workgroup barriers, lets the compiler park lanes outside it).
__device__ __forceinline__ bool persist_wait(const unsigned *p, unsigned need, unsigned *err, long long ticks)
{
  auto ld = [](const unsigned *q) {
    return __builtin_amdgcn_readfirstlane(__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  };
  if (ld(p) >= need) return true;
  const long long t0 = wall_clock64();
  for (;;) {
    __builtin_amdgcn_s_sleep(2);
    if (ld(p) >= need) return true;
    if (ld(err)) return false;
    if (wall_clock64() - t0 > ticks) {
      atomicOr(err, __lane_id() == 0 ? 1u : 0u);
      return false;
    }
  }
}

// Rim-source cells of one shot (k_rim_source's work, by one workgroup): u_next = u_cur + s w
// on the cells of the source box outside [5, n-5)^3.
template <class S>
__device__ __forceinline__ void rim_source_item(const UpdArgs<S> &a, int kstep, int tid, int nthreads)
{
  const Geo &g = a.g;
  const float wk = a.wav[kstep];
  const int bx = a.sb_x1 - a.sb_x0 + 1, by = a.sb_y1 - a.sb_y0 + 1, bz = a.sb_z1 - a.sb_z0 + 1;
  const long n = (long)bx * by * bz;
  for (long t = tid; t < n; t += nthreads) {
    const int yy = (int)(t % by), xx = (int)((t / by) % bx), zz = (int)(t / ((long)by * bx));
    const int x = a.sb_x0 + xx, y = a.sb_y0 + yy, z = a.sb_z0 + zz;
    const bool r5 = x >= H && x < g.nx - H && y >= H && y < g.ny - H && z >= H && z < g.nz - H;
    if (r5) continue;
    const unsigned b = 4u * ((unsigned)z * g.pln + (unsigned)x * g.ys + (unsigned)y);
    S::st(a.unext, b, S::rnd(fmaf(a.sbox[t], wk, S::ld(a.ucur, b))));
  }
}

template <class T>
constexpr int lds_floats() { return 2 * (T::LX + 1) * T::LY; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

// TG: gradient tiles (rows in runs), TI: interior, TP: PML x/z faces, TYF: y faces; all with
// the same number of wavefronts (block = 64 x WX).
template <class S, class TG, class TI, class TP, class TYF>
__global__ __launch_bounds__(TG::NT, 4) void k_persist(PersistArgs<S> pa)
{
  static_assert(TG::WX == TI::WX && TG::WX == TP::WX && TG::WX == TYF::WX, "one block shape");
  constexpr int NF = cmax(cmax(lds_floats<TG>(), lds_floats<TI>()), cmax(lds_floats<TP>(), lds_floats<TYF>()));
  __shared__ float lds[NF];
  __shared__ unsigned s_item;
  __shared__ int s_ok;
  const int tid = threadIdx.y * TY + threadIdx.x;
  // Wavefront 0 takes items, waits and signals. Every branch of the loop is uniform over a
  // wavefront (the atomics run on all 64 lanes, lane 0 adding 1 and the others 0), so all
  // waves meet every barrier once per item.
  const bool w0 = __builtin_amdgcn_readfirstlane(threadIdx.y) == 0;
  const bool lane0 = __lane_id() == 0;
  unsigned *const queue = pa.ctr, *const err = pa.ctr + 1, *const done = pa.ctr + 2;
  const unsigned total = (unsigned)(pa.k1 - pa.k0 + 1) * (unsigned)pa.ips;
  for (;;) {
    if (w0) {
      unsigned v = total;
      if (!__builtin_amdgcn_readfirstlane(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
        v = __builtin_amdgcn_readfirstlane(atomicAdd(queue, lane0 ? 1u : 0u));
      s_item = v;
    }
    __syncthreads();
    // the item is uniform over the workgroup: keep it (and so the argument blocks) in scalar
    // registers
    const unsigned i = __builtin_amdgcn_readfirstlane(s_item);
    if (i >= total) break;
    const unsigned kr = i / (unsigned)pa.ips, j = i - kr * (unsigned)pa.ips;
    const int4 itv = pa.items[j];
    const int4 it = make_int4(__builtin_amdgcn_readfirstlane(itv.x), __builtin_amdgcn_readfirstlane(itv.y),
                              __builtin_amdgcn_readfirstlane(itv.z), __builtin_amdgcn_readfirstlane(itv.w));
    const int shot = it.w & 0xffff, kind = (it.w >> 16) & 0xff;
    const int ph = kind == PK_GRAD ? 0 : 1;
    const int k = pa.k0 + (int)kr - ((it.w >> 24) & 1), par = k & 1;
    const bool valid = k >= pa.k0 && k < pa.k1;  // uniform
    if (w0) {
      // gradient items of step k: all update items of the steps before; update items: all
      // gradient items up to and including step k
      const unsigned kk = (unsigned)(k - pa.k0);
      const unsigned need = ph == 0 ? kk * pa.nitems[pa.nshots + shot] : (kk + 1) * pa.nitems[shot];
      const unsigned *c = done + (ph == 0 ? pa.nshots + shot : shot);
      const bool ok = !valid || persist_wait(c, need, err, pa.wait_ticks);
      if (ok && valid) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the producers' stores, not stale lines
      s_ok = ok;
    }
    __syncthreads();
    if (!__builtin_amdgcn_readfirstlane(s_ok)) break;
    if (!valid) continue;
    const int4 w = make_int4(it.x, it.y, it.z & 0xffff, (int)((unsigned)it.z >> 16));
    switch (kind) {
      case PK_GRAD: {
        const GradArgs<S> a = pa.ag[par][shot];
        gradmax_tile<S, TG>(a, w, k, reinterpret_cast<TileLds<TG>>(lds));
        break;
      }
      case PK_INT: {
        const UpdArgs<S> a = pa.au[par][shot];
        update_tile<S, TI, 0>(a, w, k, reinterpret_cast<TileLds<TI>>(lds));
        break;
      }
      case PK_PML: {
        const UpdArgs<S> a = pa.au[par][shot];
        update_tile<S, TP, 1>(a, w, k, reinterpret_cast<TileLds<TP>>(lds));
        break;
      }
      case PK_PMLY: {
        const UpdArgs<S> a = pa.au[par][shot];
        update_tile<S, TYF, 2>(a, w, k, reinterpret_cast<TileLds<TYF>>(lds));
        break;
      }
      case PK_PMLYF: {
        const UpdArgs<S> a = pa.au[par][shot];
        update_tile<S, TYF, 1>(a, w, k, reinterpret_cast<TileLds<TYF>>(lds));
        break;
      }
      default: {
        const UpdArgs<S> a = pa.au[par][shot];
        rim_source_item<S>(a, k, tid, TG::NT);
        break;
      }
    }
    // this wave's stores complete (vmcnt counts stores on gfx9-family parts), then one release
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (w0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      atomicAdd(done + (ph == 0 ? shot : pa.nshots + shot), lane0 ? 1u : 0u);
    }
  }
}

// Zeroes the queue, the error word and the done counters (a fill kernel rather than a memset,
// so that the pair can be captured in a graph).
static __global__ void k_persist_init(unsigned *ctr, int n)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ctr[i] = 0u;
}

}  // namespace qp3d
