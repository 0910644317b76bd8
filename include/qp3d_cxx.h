/*
 * The reference's solver entry point with C++ linkage, for callers compiled as C++.
 *
 * The reference declares modeling3d without extern "C" (3dPureqPModeling_single_shot.h:4-6) and
 * builds its driver with g++ (Makefile:3,11-13; compile_GPU.sh:4-11), so its Main.o refers to
 * the C++-mangled symbol. libqp3d.so exports that symbol too (host/qp3d_dropin_cxx.cpp, which
 * forwards to qp3d_modeling3d), so the reference's unchanged Main.o, built against its own
 * header, links against libqp3d.so in place of modeling3D.o / the CUDA objects.
 *
 * This header only restates that declaration. Do not include it in a translation unit that
 * also includes qp3d.h: C++ does not allow one function name and signature with both C and
 * C++ linkage in one scope.
 */
#ifndef QP3D_CXX_H
#define QP3D_CXX_H
#ifdef __cplusplus

/* Replaces void modeling3d(...) of 3dPureqPModeling_single_shot.h:4-6 (GPU body
 * 3dPureqPModeling_single_shot.cu:26-447, CPU twin modeling3D.c:17-153) for C++ callers: same
 * argument list, meaning and host-memory buffers as the C entry point of qp3d.h. Failures are
 * printed to stderr and kept in qp3d_last_error(); on a non-finite wavefield the record is
 * still written (as the reference's record would hold the NaNs). */
void modeling3d(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt, int pml,
                int snapflag, int sx, int sy, int sz, float *vp, float *epsilon, float *delta, float *source,
                float *wavelet, float *record, float *dlr, float *ddlr, float *dtb, float *ddtb, float *dfb,
                float *ddfb, float *c, float *c2, const char *snap_file);

#endif /* __cplusplus */
#endif /* QP3D_CXX_H */
