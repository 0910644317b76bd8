/*
 * The drop-in solver entry points: modeling3d() with the reference's signature
 * (3dPureqPModeling_single_shot.h:4-6) and its status-returning twin, built on the plan API.
 * One call = one shot: allocate, upload, run nt steps on the GPU, copy the receiver record
 * back, write the snapshot files, free (the reference's call also allocates and frees per
 * shot, 3dPureqPModeling_single_shot.cu:100-186,416-441).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qp3d.h"

static int write_floats(const char *name, const float *a, size_t n)
{
  FILE *fp = fopen(name, "wb");
  if (!fp) {
    qp3d_set_error("cannot open %s for writing", name);
    return QP3D_ERR_IO;
  }
  size_t w = fwrite(a, sizeof(float), n, fp);
  fclose(fp);
  if (w != n) {
    qp3d_set_error("short write to %s", name);
    return QP3D_ERR_IO;
  }
  return QP3D_OK;
}

/* 3dPureqPModeling_single_shot.cu:340-377 / modeling3D.c:89-124: the interior cube and its
 * three centre slices, xz at y'=(hy-1)/2, yz at x'=(hx-1)/2, xy at z'=(hz-1)/2. With tag "S"
 * the same for the S field over the padded grid, named as the reference's commented-out writer
 * names them (3dPureqPModeling_single_shot.cu:381-409): <prefix>_S<k>.dat, _Sxz, _Syz, _Sxy. */
static int write_snapshot_files(const char *prefix, const char *tag, int k, const float *snap, int hx, int hy, int hz)
{
  char name[4096];
  float *xz = (float *)malloc(sizeof(float) * (size_t)hz * hx);
  float *yz = (float *)malloc(sizeof(float) * (size_t)hz * hy);
  float *xy = (float *)malloc(sizeof(float) * (size_t)hx * hy);
  int rc = QP3D_OK;
  if (!xz || !yz || !xy) {
    qp3d_set_error("snapshot: out of host memory");
    rc = QP3D_ERR_ALLOC;
    goto done;
  }
  for (int z = 0; z < hz; z++)
    for (int x = 0; x < hx; x++) xz[(size_t)z * hx + x] = snap[((size_t)z * hx + x) * hy + (hy - 1) / 2];
  for (int z = 0; z < hz; z++)
    for (int y = 0; y < hy; y++) yz[(size_t)z * hy + y] = snap[((size_t)z * hx + (hx - 1) / 2) * hy + y];
  for (int x = 0; x < hx; x++)
    for (int y = 0; y < hy; y++) xy[(size_t)x * hy + y] = snap[((size_t)((hz - 1) / 2) * hx + x) * hy + y];
  snprintf(name, sizeof name, "%s%s%s%d.dat", prefix, *tag ? "_" : "", tag, k);
  if ((rc = write_floats(name, snap, (size_t)hx * hy * hz))) goto done;
  snprintf(name, sizeof name, "%s_%sxz%d.dat", prefix, tag, k);
  if ((rc = write_floats(name, xz, (size_t)hz * hx))) goto done;
  snprintf(name, sizeof name, "%s_%syz%d.dat", prefix, tag, k);
  if ((rc = write_floats(name, yz, (size_t)hz * hy))) goto done;
  snprintf(name, sizeof name, "%s_%sxy%d.dat", prefix, tag, k);
  rc = write_floats(name, xy, (size_t)hx * hy);
done:
  free(xz);
  free(yz);
  free(xy);
  return rc;
}

int qp3d_modeling3d(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt, int pml,
                    int snapflag, int sx, int sy, int sz, const float *vp, const float *epsilon, const float *delta,
                    const float *source, const float *wavelet, float *record, const float *dlr, const float *ddlr,
                    const float *dtb, const float *ddtb, const float *dfb, const float *ddfb, const float *c,
                    const float *c2, const char *snap_file)
{
  (void)sx;
  (void)sy;
  if (!record) {
    qp3d_set_error("modeling3d: record is null");
    return QP3D_ERR_ARG;
  }
  if (snapflag == 1 && !snap_file) {
    qp3d_set_error("modeling3d: snapflag == 1 needs snap_file");
    return QP3D_ERR_ARG;
  }
  qp3d_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.nx = nx; cfg.ny = ny; cfg.nz = nz; cfg.pml = pml; cfg.nt = nt;
  cfg.dx = dx; cfg.dy = dy; cfg.dz = dz; cfg.dt = dt;
  cfg.sz = sz;
  cfg.storage = QP3D_STORAGE_FP32;
  cfg.snapflag = snapflag == 1 ? 1 : 0;
  cfg.ntsnap = ntsnap;
  /* QP3D_SNAP_S=1 also writes the S field at every snapshot step (the reference has that
     writer commented out, so it is off by default) */
  const char *ss = getenv("QP3D_SNAP_S");
  cfg.snap_s = cfg.snapflag && ss && atoi(ss) == 1;
  /* QP3D_GMAX_MODE=2 takes G from the update kernels under a predicted threshold (the same
     results up to rounding; see qp3d_config.gmax_mode); unset or 0: the default */
  const char *gm = getenv("QP3D_GMAX_MODE");
  cfg.gmax_mode = gm ? atoi(gm) : 0;
  qp3d_plan *p = NULL;
  int nonfinite = 0;
  int rc = qp3d_plan_create(&cfg, &p);
  if (rc) return rc;
  if ((rc = qp3d_plan_upload(p, vp, epsilon, delta, source, wavelet, dlr, ddlr, dtb, ddtb, dfb, ddfb, c, c2)) ||
      (rc = qp3d_plan_reset(p)) || (rc = qp3d_plan_run(p, 0, nt)))
    goto out;
  /* A non-finite wavefield (CFL violation) still returns its record and snapshots, as the
     reference's would hold the inf/NaN values; the status says what happened. */
  if ((rc = qp3d_plan_sync(p)) == QP3D_ERR_NONFINITE) {
    nonfinite = 1;
    rc = QP3D_OK;
  }
  if (rc || (rc = qp3d_plan_record(p, record))) goto out;
  if (cfg.snapflag) {
    int hx, hy, hz;
    qp3d_plan_dims(p, &hx, &hy, &hz);
    float *snap = (float *)malloc(sizeof(float) * (size_t)hx * hy * hz);
    if (!snap) {
      qp3d_set_error("modeling3d: out of host memory for a snapshot");
      rc = QP3D_ERR_ALLOC;
      goto out;
    }
    for (int i = 0; i < qp3d_plan_snapshot_count(p) && rc == QP3D_OK; i++) {
      int k = 0;
      if (!(rc = qp3d_plan_snapshot(p, i, &k, snap))) rc = write_snapshot_files(snap_file, "", k, snap, hx, hy, hz);
    }
    free(snap);
    if (rc == QP3D_OK && cfg.snap_s) {
      float *sf = (float *)malloc(sizeof(float) * (size_t)nx * ny * nz);
      if (!sf) {
        qp3d_set_error("modeling3d: out of host memory for an S-field snapshot");
        rc = QP3D_ERR_ALLOC;
        goto out;
      }
      for (int i = 0; i < qp3d_plan_snapshot_count(p) && rc == QP3D_OK; i++) {
        int k = 0;
        if (!(rc = qp3d_plan_snapshot_s(p, i, &k, sf))) rc = write_snapshot_files(snap_file, "S", k, sf, nx, ny, nz);
      }
      free(sf);
    }
  }
out:
  qp3d_plan_destroy(p);
  if (rc == QP3D_OK && nonfinite) {
    qp3d_set_error("wavefield became non-finite (inf/NaN): dt violates the stability limit for this model");
    rc = QP3D_ERR_NONFINITE;
  }
  return rc;
}

void modeling3d(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt, int pml,
                int snapflag, int sx, int sy, int sz, float *vp, float *epsilon, float *delta, float *source,
                float *wavelet, float *record, float *dlr, float *ddlr, float *dtb, float *ddtb, float *dfb, float *ddfb,
                float *c, float *c2, const char *snap_file)
{
  int rc = qp3d_modeling3d(nx, ny, nz, nt, ntsnap, dx, dy, dz, dt, pml, snapflag, sx, sy, sz, vp, epsilon, delta,
                           source, wavelet, record, dlr, ddlr, dtb, ddtb, dfb, ddfb, c, c2, snap_file);
  if (rc) fprintf(stderr, "modeling3d: %s\n", qp3d_last_error());
}
