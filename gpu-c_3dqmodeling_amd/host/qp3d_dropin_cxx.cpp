// The C++-linkage modeling3d (declared in include/qp3d_cxx.h): the symbol a driver compiled
// with g++ against the reference's header (3dPureqPModeling_single_shot.h:4-6) refers to.
// It forwards to the status-returning C entry point.
#include <cstdio>

#include "qp3d_cxx.h"

extern "C" {
// from qp3d.h (not included here: its modeling3d has C linkage)
int qp3d_modeling3d(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt, int pml,
                    int snapflag, int sx, int sy, int sz, const float *vp, const float *epsilon,
                    const float *delta, const float *source, const float *wavelet, float *record,
                    const float *dlr, const float *ddlr, const float *dtb, const float *ddtb, const float *dfb,
                    const float *ddfb, const float *c, const float *c2, const char *snap_file);
const char *qp3d_last_error(void);
}

void modeling3d(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt, int pml,
                int snapflag, int sx, int sy, int sz, float *vp, float *epsilon, float *delta, float *source,
                float *wavelet, float *record, float *dlr, float *ddlr, float *dtb, float *ddtb, float *dfb,
                float *ddfb, float *c, float *c2, const char *snap_file)
{
  const int rc = qp3d_modeling3d(nx, ny, nz, nt, ntsnap, dx, dy, dz, dt, pml, snapflag, sx, sy, sz, vp, epsilon,
                                 delta, source, wavelet, record, dlr, ddlr, dtb, ddtb, dfb, ddfb, c, c2, snap_file);
  if (rc) std::fprintf(stderr, "modeling3d: %s\n", qp3d_last_error());
}
