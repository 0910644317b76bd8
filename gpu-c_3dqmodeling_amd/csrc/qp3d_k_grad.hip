// Instantiations of the gradient-max kernels (plain, marking, batched) for every tile shape,
// for the storage policy QP3D_STORE (the Makefile builds this file once per policy).
#include "qp3d_launch.hpp"

#ifndef QP3D_STORE
#error "QP3D_STORE (StoreF32, StoreBF16 or StoreF16) must be defined"
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

template <class S>
void launch_gradmax(int tile, int blocks, hipStream_t st, const GradArgs<S> &a)
{
  int i = 0;
#define QP3D_L(wx, rx)                                                                              \
  if (tile == i++) {                                                                                \
    hipLaunchKernelGGL((k_gradmax<S, Tile<wx, rx>>), dim3(blocks), dim3(TY, wx), 0, st, a);         \
    return;                                                                                         \
  }
  QP3D_TILES(QP3D_L)
#undef QP3D_L
}

template <class S>
void launch_mark(int tile, int blocks, hipStream_t st, const GradArgs<S> &a)
{
  int i = 0;
#define QP3D_L(wx, rx)                                                                              \
  if (tile == i++) {                                                                                \
    hipLaunchKernelGGL((k_spec_mark<S, Tile<wx, rx>>), dim3(blocks), dim3(TY, wx), 0, st, a);       \
    return;                                                                                         \
  }
  QP3D_TILES(QP3D_L)
#undef QP3D_L
}

template <class S>
void launch_gbound(int tile, int blocks, hipStream_t st, const GradArgs<S> &a)
{
  int i = 0;
#define QP3D_L(wx, rx)                                                                              \
  if (tile == i++) {                                                                                \
    hipLaunchKernelGGL((k_gbound<S, Tile<wx, rx>>), dim3(blocks), dim3(TY, wx), 0, st, a);          \
    return;                                                                                         \
  }
  QP3D_TILES(QP3D_L)
#undef QP3D_L
}

template <class S>
int occupancy_gradmax(int tile)
{
  int i = 0;
#define QP3D_O(wx, rx) \
  if (tile == i++) return occ_of(reinterpret_cast<const void *>(&k_gradmax<S, Tile<wx, rx>>), TY * wx);
  QP3D_TILES(QP3D_O)
#undef QP3D_O
  return 1;
}

template <class S>
void launch_gradmax_batch(int tile, int blocks, const int4 *work, hipStream_t st, const GradArgs<S> *A, Step s,
                          int skip)
{
  int i = 0;
#define QP3D_L(wx, rx)                                                                                       \
  if (tile == i++) {                                                                                         \
    hipLaunchKernelGGL((k_gradmax_batch<S, Tile<wx, rx>>), dim3(blocks), dim3(TY, wx), 0, st, A, work, s, skip); \
    return;                                                                                                  \
  }
  QP3D_TILES(QP3D_L)
#undef QP3D_L
}

template <class S>
void launch_gbound_batch(int tile, int blocks, int nshots, const int4 *work, int nwork, const int *first,
                         hipStream_t st, const GradArgs<S> *A, Step s)
{
  int i = 0;
#define QP3D_L(wx, rx)                                                                                   \
  if (tile == i++) {                                                                                     \
    hipLaunchKernelGGL((k_gbound_batch<S, Tile<wx, rx>>), dim3(blocks, nshots), dim3(TY, wx), 0, st, A, work, \
                       nwork, first, s);                                                                 \
    return;                                                                                              \
  }
  QP3D_TILES(QP3D_L)
#undef QP3D_L
}

template <class S>
int occupancy_gradmax_batch(int tile)
{
  int i = 0;
#define QP3D_O(wx, rx) \
  if (tile == i++) return occ_of(reinterpret_cast<const void *>(&k_gradmax_batch<S, Tile<wx, rx>>), TY * wx);
  QP3D_TILES(QP3D_O)
#undef QP3D_O
  return 1;
}

template void launch_gradmax<QP3D_STORE>(int, int, hipStream_t, const GradArgs<QP3D_STORE> &);
template void launch_mark<QP3D_STORE>(int, int, hipStream_t, const GradArgs<QP3D_STORE> &);
template void launch_gbound<QP3D_STORE>(int, int, hipStream_t, const GradArgs<QP3D_STORE> &);
template int occupancy_gradmax<QP3D_STORE>(int);
template void launch_gradmax_batch<QP3D_STORE>(int, int, const int4 *, hipStream_t, const GradArgs<QP3D_STORE> *, Step,
                                                int);
template void launch_gbound_batch<QP3D_STORE>(int, int, int, const int4 *, int, const int *, hipStream_t,
                                              const GradArgs<QP3D_STORE> *, Step);
template int occupancy_gradmax_batch<QP3D_STORE>(int);

}  // namespace qp3d
