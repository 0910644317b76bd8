// C++-linkage modeling3d (the symbol of 3dPureqPModeling_single_shot.h:4-6) over the FP32
// oracle, so that the reference's own driver Main.c, compiled unchanged with g++, runs its shot
// loop on the CPU restatement: oracle/_ref/ref_main_oracle. Its twin oracle/_ref/ref_main_qp3d
// links the same Main.o against libqp3d.so. The reference swaps its CPU and GPU bodies of
// this symbol the same way, at link time (Makefile:8, compile_GPU.sh:11).
// TEST INFRASTRUCTURE ONLY: the GPU drop-in test compares the two drivers' output files.
#include <cstdlib>

extern "C" int oracle_modeling3d_f32(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz,
                                     float dt, int pml, int snapflag, int sz, const float *vp, const float *eps,
                                     const float *del, const float *source, const float *wavelet, float *record,
                                     const float *dlr, const float *ddlr, const float *dtb, const float *ddtb,
                                     const float *dfb, const float *ddfb, const float *c, const float *c2,
                                     const char *snap_file, int nthreads);

void modeling3d(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt, int pml,
                int snapflag, int sx, int sy, int sz, float *vp, float *epsilon, float *delta, float *source,
                float *wavelet, float *record, float *dlr, float *ddlr, float *dtb, float *ddtb, float *dfb,
                float *ddfb, float *c, float *c2, const char *snap_file)
{
  (void)sx;
  (void)sy;
  const char *t = std::getenv("OMP_NUM_THREADS");
  const int nthreads = t && std::atoi(t) > 0 ? std::atoi(t) : 8;
  oracle_modeling3d_f32(nx, ny, nz, nt, ntsnap, dx, dy, dz, dt, pml, snapflag, sz, vp, epsilon, delta, source,
                        wavelet, record, dlr, ddlr, dtb, ddtb, dfb, ddfb, c, c2, snap_file, nthreads);
}
