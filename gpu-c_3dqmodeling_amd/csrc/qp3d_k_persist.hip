// Instantiation of the persistent batch kernel (k_persist) for the storage policy QP3D_STORE and
// the tile set QP3D_PSET (PSET_TILES in qp3d_launch.hpp); the Makefile builds this file once per
// combination.
#include "qp3d_launch.hpp"
#include "qp3d_persist.hpp"

#ifndef QP3D_STORE
#error "QP3D_STORE (StoreF32, StoreBF16 or StoreF16) must be defined"
#endif
#ifndef QP3D_PSET
#error "QP3D_PSET (0, 1 or 2) must be defined"
#endif

namespace qp3d {

namespace {
// gradient, interior, PML x/z and y-face tile shapes of each set (PSET_TILES)
template <int SET> struct PSet;
template <> struct PSet<0> { using G = Tile<8, 2>; using I = Tile<8, 2>; using P = Tile<8, 1>; using Y = Tile<8, 1>; };
template <> struct PSet<1> { using G = Tile<12, 1>; using I = Tile<12, 1>; using P = Tile<12, 1>; using Y = Tile<12, 1>; };
template <> struct PSet<2> { using G = Tile<16, 1>; using I = Tile<16, 1>; using P = Tile<16, 1>; using Y = Tile<16, 1>; };
}  // namespace

template <class S, int SET>
void launch_persist(int blocks, hipStream_t st, const PersistArgs<S> &a)
{
  using P = PSet<SET>;
  hipLaunchKernelGGL((k_persist<S, typename P::G, typename P::I, typename P::P, typename P::Y>), dim3(blocks),
                     dim3(TY, P::G::WX), 0, st, a);
}

template <class S, int SET>
int occupancy_persist()
{
  using P = PSet<SET>;
  int occ = 1;
  const void *fn = reinterpret_cast<const void *>(&k_persist<S, typename P::G, typename P::I, typename P::P, typename P::Y>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, TY * P::G::WX, 0) != hipSuccess) occ = 1;
  return occ > 0 ? occ : 1;
}

template void launch_persist<QP3D_STORE, QP3D_PSET>(int, hipStream_t, const PersistArgs<QP3D_STORE> &);
template int occupancy_persist<QP3D_STORE, QP3D_PSET>();

}  // namespace qp3d
