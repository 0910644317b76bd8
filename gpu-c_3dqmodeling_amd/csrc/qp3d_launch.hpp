// Launchers of the stencil kernels, one per (kernel, storage policy), dispatching a runtime
// tile index (QP3D_TILES order) to the kernel instantiated for that tile shape; `blocks` is the
// length of the work list the arguments point to. The kernels are instantiated in their own
// translation units (qp3d_k_grad.hip and qp3d_k_upd.hip, built once per storage policy and
// update kind), so that the library compiles in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "qp3d_kernels.hpp"

namespace qp3d {

// gradient-max pass (k_gradmax) and its marking variant (k_spec_mark)
template <class S>
void launch_gradmax(int tile, int blocks, hipStream_t st, const GradArgs<S> &a);
template <class S>
void launch_mark(int tile, int blocks, hipStream_t st, const GradArgs<S> &a);
// first part of the gradient pass with the amplitude bound (k_gbound), `blocks` workgroups
template <class S>
void launch_gbound(int tile, int blocks, hipStream_t st, const GradArgs<S> &a);
// resident workgroups per CU of k_gradmax for a tile shape
template <class S>
int occupancy_gradmax(int tile);

// update pass (k_update<PML>, PML = update kind 0, 1 or 2; spec: the predicted-threshold form)
template <class S, int PML>
void launch_update(int tile, int blocks, hipStream_t st, const UpdArgs<S> &a, bool spec);
template <class S, int PML>
int occupancy_update(int tile, bool spec);

// batched forms (argument blocks in device memory, items carry the grid index)
template <class S>
void launch_gradmax_batch(int tile, int blocks, const int4 *work, hipStream_t st, const GradArgs<S> *A, Step s,
                          int skip);
// bound pass of every shot of a batch: `blocks` workgroups per shot
template <class S>
void launch_gbound_batch(int tile, int blocks, int nshots, const int4 *work, int nwork, const int *first,
                         hipStream_t st, const GradArgs<S> *A, Step s);
template <class S, int PML>
void launch_update_batch(int tile, int blocks, const int4 *work, hipStream_t st, const UpdArgs<S> *A, Step s,
                         int maps);
template <class S>
int occupancy_gradmax_batch(int tile);
template <class S, int PML>
int occupancy_update_batch(int tile);

// persistent batch kernel (qp3d_persist.hpp) of tile set SET: its launch (blocks resident
// workgroups) and its resident workgroups per CU
template <class S> struct PersistArgs;
template <class S, int SET>
void launch_persist(int blocks, hipStream_t st, const PersistArgs<S> &a);
template <class S, int SET>
int occupancy_persist();
// tile shapes (QP3D_TILES index) of each persistent tile set: gradient pass, interior update,
// PML x/z faces, y faces; one block shape (64 x WX) per set
constexpr int NPSETS = 3;
constexpr int PSET_TILES[NPSETS][4] = {{0, 0, 3, 3}, {4, 4, 4, 4}, {1, 1, 1, 1}};

}  // namespace qp3d
