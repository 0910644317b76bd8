// Instantiations of the update kernels (plain, predicted-threshold, batched) for every tile
// shape, for the storage policy QP3D_STORE and the tile kind QP3D_PML (0: interior tiles,
// 1: PML tiles, 2: folded y-face tiles over compact arrays); the Makefile builds this file once
// per combination.
#include "qp3d_launch.hpp"

#ifndef QP3D_STORE
#error "QP3D_STORE (StoreF32, StoreBF16 or StoreF16) must be defined"
#endif
#ifndef QP3D_PML
#error "QP3D_PML (0, 1 or 2) must be defined"
#endif

namespace qp3d {

namespace {
int occ_of(const void *fn, int threads)
{
  int occ = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, threads, 0) != hipSuccess) occ = 1;
  return occ > 0 ? occ : 1;
}
}  // namespace

template <class S, int PML>
void launch_update(int tile, int blocks, hipStream_t st, const UpdArgs<S> &a, bool spec)
{
  int i = 0;
#define QP3D_L(wx, rx)                                                                                 \
  if (tile == i++) {                                                                                   \
    if (spec)                                                                                          \
      hipLaunchKernelGGL((k_update<S, Tile<wx, rx>, PML, true>), dim3(blocks), dim3(TY, wx), 0, st, a); \
    else                                                                                               \
      hipLaunchKernelGGL((k_update<S, Tile<wx, rx>, PML>), dim3(blocks), dim3(TY, wx), 0, st, a);      \
    return;                                                                                            \
  }
  QP3D_TILES(QP3D_L)
#undef QP3D_L
}

template <class S, int PML>
int occupancy_update(int tile, bool spec)
{
  int i = 0;
#define QP3D_O(wx, rx)                                                                                     \
  if (tile == i++)                                                                                         \
    return occ_of(spec ? reinterpret_cast<const void *>(&k_update<S, Tile<wx, rx>, PML, true>)             \
                       : reinterpret_cast<const void *>(&k_update<S, Tile<wx, rx>, PML>),                  \
                  TY * wx);
  QP3D_TILES(QP3D_O)
#undef QP3D_O
  return 1;
}

template <class S, int PML>
void launch_update_batch(int tile, int blocks, const int4 *work, hipStream_t st, const UpdArgs<S> *A, Step s,
                         int maps)
{
  int i = 0;
#define QP3D_L(wx, rx)                                                                                             \
  if (tile == i++) {                                                                                               \
    hipLaunchKernelGGL((k_update_batch<S, Tile<wx, rx>, PML>), dim3(blocks), dim3(TY, wx), 0, st, A, work, s, maps); \
    return;                                                                                                        \
  }
  QP3D_TILES(QP3D_L)
#undef QP3D_L
}

template <class S, int PML>
int occupancy_update_batch(int tile)
{
  int i = 0;
#define QP3D_O(wx, rx)                                                                                       \
  if (tile == i++) return occ_of(reinterpret_cast<const void *>(&k_update_batch<S, Tile<wx, rx>, PML>), TY * wx);
  QP3D_TILES(QP3D_O)
#undef QP3D_O
  return 1;
}

template void launch_update<QP3D_STORE, QP3D_PML>(int, int, hipStream_t, const UpdArgs<QP3D_STORE> &, bool);
template int occupancy_update<QP3D_STORE, QP3D_PML>(int, bool);
template void launch_update_batch<QP3D_STORE, QP3D_PML>(int, int, const int4 *, hipStream_t,
                                                        const UpdArgs<QP3D_STORE> *, Step, int);
template int occupancy_update_batch<QP3D_STORE, QP3D_PML>(int);

}  // namespace qp3d
