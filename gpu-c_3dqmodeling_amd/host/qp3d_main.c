/*
 * qp3d_main: the drop-in host driver. It reads the reference's parameter file
 * (./input/parameter.txt, or the path given as argv[1]) in the same order as Main.c:24-123,
 * runs the shot loop of Main.c:163-247 on the GPU through modeling3d(), and writes the same
 * output files with the same byte layout and open modes (read_write.c:29-77):
 *   direct-wave record (overwrite) when muteflag == 1,
 *   full-wave and muted records appended per shot,
 *   xz / yz receiver slices of the first shot (written twice; the final content is the
 *   full-wave slice, Main.c:219-245),
 *   snapshots <snap_file>k.dat, _xz, _yz, _xy every ntsnap steps when enabled.
 *
 * Differences from Main.c, each a defect fix:
 *   - the direct-wave model is a homogeneous 2500 m/s isotropic medium
 *     (3dPureqPModeling.c:212-219); Main.c passes never-filled zero arrays (Main.c:133,172);
 *   - missing/short files and model windows outside the model are reported and stop the run
 *     instead of reading out of bounds (Main.c:25-30,187);
 *   - GPU failures are reported with the library's status and message.
 * Optional extensions: QP3D_SRC_DEPTH=n places the source and receivers n cells below the
 * top of the interior (the reference forces depth 0, Main.c:200); QP3D_SNAP_S=1 makes
 * modeling3d() also write the S field of every snapshot step (<snap_file>_S<k>.dat, _Sxz,
 * _Syz, _Sxy; the reference's writer for them is commented out,
 * 3dPureqPModeling_single_shot.cu:381-409); QP3D_GMAX_MODE=2 makes modeling3d() take the
 * gradient maximum from the update kernels under a predicted threshold (qp3d_config.gmax_mode).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qp3d.h"

#define MAXS 4096

static int skip_line(FILE *fp)
{
  char buff[MAXS];
  return fgets(buff, sizeof buff, fp) != NULL;
}

static int write_raw(const char *name, const float *a, size_t n, const char *mode)
{
  FILE *fp = fopen(name, mode);
  if (!fp) {
    fprintf(stderr, "qp3d_main: cannot open %s\n", name);
    return -1;
  }
  size_t w = fwrite(a, sizeof(float), n, fp);
  fclose(fp);
  return w == n ? 0 : -1;
}

static float *zalloc(size_t n)
{
  float *p = (float *)calloc(n ? n : 1, sizeof(float));
  if (!p) {
    fprintf(stderr, "qp3d_main: out of host memory (%zu floats)\n", n);
    exit(3);
  }
  return p;
}

int main(int argc, char *argv[])
{
  const char *parfile = argc > 1 ? argv[1] : "./input/parameter.txt";
  int Nx0, Ny0, Nz0, pml, nt, hx, hy, nxshot, nyshot, dxshot, dyshot, multishot, muteflag, snapflag1, snapflag2,
      sx0, sy0, sz0, ntsnap, gx0, gy0, gz0;
  float dx, dy, dz, dt, fm, amp, alpha;
  char vp_file[MAXS], eps_file[MAXS], del_file[MAXS], recfull_file[MAXS], recfilexz[MAXS], recfileyz[MAXS],
      recdir_file[MAXS], recmut_file[MAXS], snap_file1[MAXS], snap_file2[MAXS];
  FILE *fp = fopen(parfile, "r");
  if (!fp) {
    fprintf(stderr, "%s does not exist\n", parfile);
    return 2;
  }
  /* Main.c:31-122: a header line before each group of values */
  int ok = 1;
  ok &= skip_line(fp);
  ok &= fscanf(fp, "%f %d %d", &dt, &nt, &ntsnap) == 3;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%d %d %d", &Nx0, &Ny0, &Nz0) == 3;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%f %f %f", &dx, &dy, &dz) == 3;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%f %f %f", &fm, &amp, &alpha) == 3;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%d %d %d", &sx0, &sy0, &sz0) == 3;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%d %d %d", &gx0, &gy0, &gz0) == 3;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%d %d %d %d", &nxshot, &nyshot, &dxshot, &dyshot) == 4;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%d %d", &hx, &hy) == 2;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%d", &pml) == 1;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%4095s %4095s %4095s", vp_file, eps_file, del_file) == 3;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%4095s %4095s %4095s", recfull_file, recdir_file, recmut_file) == 3;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%4095s %4095s", recfilexz, recfileyz) == 2;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%4095s %4095s", snap_file1, snap_file2) == 2;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%d", &multishot) == 1;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%d", &muteflag) == 1;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%d", &snapflag2) == 1;
  ok &= skip_line(fp) & skip_line(fp);
  ok &= fscanf(fp, "%d", &snapflag1) == 1;
  fclose(fp);
  if (!ok) {
    fprintf(stderr, "%s: malformed or short parameter file\n", parfile);
    return 2;
  }
  (void)gx0; (void)gy0; (void)gz0; (void)multishot; (void)sz0;
  printf("dt= %f  nt=%d ntsnap=%d\nNx0=%d ny0=%d nz0=%d\ndx=%f,dy=%f dz=%f\nfm=%f amp=%f alpha=%f\n", dt, nt, ntsnap,
         Nx0, Ny0, Nz0, dx, dy, dz, fm, amp, alpha);
  printf("hx=%d hy=%d pml=%d nshot=%d\n", hx, hy, pml, nxshot * nyshot);
  if (pml < 6 || hx < 1 || hy < 1 || Nz0 < 1 || nt < 1 || ntsnap < 1) {
    fprintf(stderr, "qp3d_main: need pml >= 6, positive window, nt and ntsnap\n");
    return 2;
  }
  int src_depth = 0;
  const char *e = getenv("QP3D_SRC_DEPTH");
  if (e) src_depth = atoi(e);

  const int nx = hx + 2 * pml, ny = hy + 2 * pml, nz = Nz0 + 2 * pml;
  const float t0 = (float)(1.0 / fm);
  const size_t N = (size_t)nx * ny * nz, H = (size_t)hx * hy * Nz0, R = (size_t)hx * hy * nt;
  float *vp1 = zalloc(H), *ep1 = zalloc(H), *de1 = zalloc(H);
  float *vp = zalloc(N), *ep = zalloc(N), *de = zalloc(N), *source = zalloc(N), *wavelet = zalloc(nt);
  float *rec_full = zalloc(R), *rec_dir = zalloc(R), *rec_mute = zalloc(R);
  float *dlr = zalloc(pml), *ddlr = zalloc(pml), *dtb = zalloc(pml), *ddtb = zalloc(pml), *dfb = zalloc(pml),
        *ddfb = zalloc(pml);
  float *recxz = zalloc((size_t)nt * hx), *recyz = zalloc((size_t)nt * hy);
  float c[36] = {0}, c2[36] = {0};
  qp3d_fdcoff1(c);
  qp3d_fdcoff2(c2);
  const int sx1 = (hx - 1) / 2 + pml, sy1 = (hy - 1) / 2 + pml, sz1 = pml + src_depth;
  int rc;

  if (muteflag == 1) {
    printf("/************* Modeling directwave ***************/\n");
    float *vpd = zalloc(N), *epd = zalloc(N), *ded = zalloc(N);
    for (size_t i = 0; i < N; i++) vpd[i] = 2500.0f;
    float vpmax = (1 + 2 * qp3d_absmax(epd, (long)N)) * qp3d_absmax(vpd, (long)N);
    qp3d_pmlcoff(pml, vpmax, dx, dy, dz, dlr, ddlr, dtb, ddtb, dfb, ddfb);
    qp3d_source(sx1, sy1, sz1, fm, amp, alpha, dt, dx, dy, dz, t0, nt, nx, ny, nz, source, wavelet);
    rc = qp3d_modeling3d(nx, ny, nz, nt, ntsnap, dx, dy, dz, dt, pml, snapflag1, sx1, sy1, sz1, vpd, epd, ded, source,
                         wavelet, rec_dir, dlr, ddlr, dtb, ddtb, dfb, ddfb, c, c2, snap_file1);
    free(vpd); free(epd); free(ded);
    if (rc) {
      fprintf(stderr, "modeling3d (direct wave) failed: %s\n", qp3d_last_error());
      return 4;
    }
    if (write_raw(recdir_file, rec_dir, R, "wb")) return 5;
  }
  printf("/************* Multi-shots Modeling  ***************/\n");
  for (int ixshot = 0; ixshot < nxshot; ixshot++)
    for (int iyshot = 0; iyshot < nyshot; iyshot++) {
      printf("shot=%d\n", ixshot * nyshot + iyshot);
      const int sx = sx0 + ixshot * dxshot, sy = sy0 + iyshot * dyshot;
      const int fny = sy - (hy - 1) / 2, fnx = sx - (hx - 1) / 2;
      if ((rc = qp3d_readmodel3d(vp_file, vp1, hy, hx, Nz0, fny, fnx, 0, Ny0, Nx0, Nz0)) ||
          (rc = qp3d_readmodel3d(eps_file, ep1, hy, hx, Nz0, fny, fnx, 0, Ny0, Nx0, Nz0)) ||
          (rc = qp3d_readmodel3d(del_file, de1, hy, hx, Nz0, fny, fnx, 0, Ny0, Nx0, Nz0))) {
        fprintf(stderr, "qp3d_main: %s\n", qp3d_last_error());
        return 2;
      }
      qp3d_extend_model(vp1, vp, nx, ny, nz, pml);
      qp3d_extend_model(ep1, ep, nx, ny, nz, pml);
      qp3d_extend_model(de1, de, nx, ny, nz, pml);
      float vpmax = (1 + 2 * qp3d_absmax(ep, (long)N)) * qp3d_absmax(vp, (long)N);
      qp3d_pmlcoff(pml, vpmax, dx, dy, dz, dlr, ddlr, dtb, ddtb, dfb, ddfb);
      qp3d_source(sx1, sy1, sz1, fm, amp, alpha, dt, dx, dy, dz, t0, nt, nx, ny, nz, source, wavelet);
      rc = qp3d_modeling3d(nx, ny, nz, nt, ntsnap, dx, dy, dz, dt, pml, snapflag2, sx1, sy1, sz1, vp, ep, de, source,
                           wavelet, rec_full, dlr, ddlr, dtb, ddtb, dfb, ddfb, c, c2, snap_file2);
      if (rc) {
        fprintf(stderr, "modeling3d failed: %s\n", qp3d_last_error());
        return 4;
      }
      for (size_t i = 0; i < R; i++) rec_mute[i] = rec_full[i] - rec_dir[i];
      if (write_raw(recfull_file, rec_full, R, "ab") || write_raw(recmut_file, rec_mute, R, "ab")) return 5;
      if (ixshot == 0 && iyshot == 0) {
        const float *srcs[2] = {rec_mute, rec_full};
        for (int pass = 0; pass < 2; pass++) {
          for (int k = 0; k < nt; k++) {
            for (int i = 0; i < hx; i++) recxz[(size_t)k * hx + i] = srcs[pass][((size_t)k * hx + i) * hy + (sy1 - pml)];
            for (int j = 0; j < hy; j++) recyz[(size_t)k * hy + j] = srcs[pass][((size_t)k * hx + (sx1 - pml)) * hy + j];
          }
          if (write_raw(recfilexz, recxz, (size_t)nt * hx, "wb") || write_raw(recfileyz, recyz, (size_t)nt * hy, "wb"))
            return 5;
        }
      }
    }
  free(vp1); free(ep1); free(de1); free(vp); free(ep); free(de); free(source); free(wavelet);
  free(rec_full); free(rec_dir); free(rec_mute); free(dlr); free(ddlr); free(dtb); free(ddtb); free(dfb);
  free(ddfb); free(recxz); free(recyz);
  return 0;
}
