/*
 * Host-side setup math of the qP3D solver: FD tables, PML profiles, Ricker wavelet and
 * spatial source, PML-domain model extension, model-window reads, and the synthetic
 * layered VTI model generator. Everything here runs once per shot on the host.
 *
 * The tables must be bit-identical to the reference's, so each function keeps the
 * reference's arithmetic types: where the reference (compiled as C++) promotes to double
 * through pow()/exp() or double literals this code does too, and where C++ overload
 * resolution picks the float function (exp of a float in the wavelet) this code calls the
 * float function explicitly.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qp3d.h"

#define QP3D_PI 3.1415926535898 /* Main.c:13, CPU_function.c:6 */

/* CPU_function.c:111-128 fdcoff1: first-derivative central-difference coefficients of orders
 * 2..10 in a 6x6 table; row 5 (entries 31..35) is the 10th-order set the solver reads. */
void qp3d_fdcoff1(float *c)
{
  static const struct { int i; float v; } t[] = {
      {7, 0.5f},     {13, 0.667f},   {14, -0.083f}, {19, 0.75f},    {20, -0.15f},
      {21, 0.0167f}, {25, 0.8f},     {26, -0.2f},   {27, 0.038f},   {28, -0.0036f},
      {31, 0.833f},  {32, -0.2381f}, {33, 0.059f},  {34, -0.0099f}, {35, 0.00079f}};
  for (size_t k = 0; k < sizeof t / sizeof t[0]; k++) c[t[k].i] = t[k].v;
}

/* CPU_function.c:131-148 fdcoff2: second-derivative coefficients, same table shape. */
void qp3d_fdcoff2(float *c)
{
  c[7] = 1.0f;
  c[13] = (float)(4.0 / 3.0);
  c[14] = (float)(-1.0 / 12.0);
  c[19] = 1.5f;
  c[20] = -0.15f;
  c[21] = (float)(1.0 / 90.0);
  c[25] = 1.6f;
  c[26] = -0.2f;
  c[27] = (float)(8.0 / 315.0);
  c[28] = (float)(-1.0 / 560);
  c[31] = 1.6667f;
  c[32] = -0.2381f;
  c[33] = 0.0397f;
  c[34] = -0.0050f;
  c[35] = 0.0003f;
}

/* CPU_function.c:150-169 pmlcoff: with L = pml*h and w = (i+1)*h,
 *   d(i)  = 4*vpmax/(2L) * (w/L)^3 * ln(1/R)
 *   d'(i) = 6*vpmax*w^2 / L^4 * ln(1/R),   R = 1e-5,
 * for h = dx (left/right), dz (top/bottom), dy (forward/backward). */
static void pml_profile(int pml, float vpmax, float h, float *d, float *dd)
{
  const float R = 0.00001f;
  const float wid = pml * h;
  const float lnr = logf(1 / R);
  for (int i = 0; i < pml; i++) {
    float w = (i + 1) * h;
    d[i] = (float)(4 * vpmax / (2 * wid) * pow((double)(w / wid), 3) * (double)lnr);
    dd[i] = (float)(6 * vpmax * w * w / pow((double)wid, 4) * (double)lnr);
  }
}

void qp3d_pmlcoff(int pml, float vpmax, float dx, float dy, float dz, float *dlr, float *ddlr, float *dtb,
                  float *ddtb, float *dfb, float *ddfb)
{
  pml_profile(pml, vpmax, dx, dlr, ddlr);
  pml_profile(pml, vpmax, dz, dtb, ddtb);
  pml_profile(pml, vpmax, dy, dfb, ddfb);
}

/* CPU_function.c:37-64 Source: spatial source amp*exp(-alpha^2 r^2) (r in metres, double) over the
 * padded grid, and the Ricker wavelet w[k] = (1-2a) e^-a with a = (pi fm (k dt - t0))^2 stored as
 * float before the float exponential. */
void qp3d_source(int sx, int sy, int sz, float fm, float amp, float alpha, float dt, float dx, float dy, float dz,
                 float t0, int nt, int nx, int ny, int nz, float *source, float *wavelet)
{
  if (source) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nz; i++)
      for (int j = 0; j < nx; j++)
        for (int k = 0; k < ny; k++) {
          double r2 = pow((double)((j - sx) * dx), 2) + pow((double)((k - sy) * dy), 2) +
                      pow((double)((i - sz) * dz), 2);
          source[(long)i * nx * ny + (long)j * ny + k] = (float)(amp * exp(-pow((double)alpha, 2) * r2));
        }
  }
  if (wavelet) {
    for (int k = 0; k < nt; k++) {
      float t = k * dt;
      float a = (float)pow(QP3D_PI * fm * (double)(t - t0), 2);
      wavelet[k] = (1 - 2 * a) * expf(-a);
    }
  }
}

/* CPU_function.c:67-108 extend_model: interior copy, then edge replication into the PML
 * along z, then x, then y (the later passes see the earlier ones). */
void qp3d_extend_model(const float *in, float *out, int nx, int ny, int nz, int pml)
{
  const long pln = (long)nx * ny;
  const int hx = nx - 2 * pml, hy = ny - 2 * pml;
  for (int i = pml; i < nz - pml; i++)
    for (int j = pml; j < nx - pml; j++)
      memcpy(out + i * pln + (long)j * ny + pml, in + ((long)(i - pml) * hx + (j - pml)) * hy,
             sizeof(float) * hy);
  for (int i = 0; i < nz; i++) {
    int src = i < pml ? pml : (i >= nz - pml ? nz - pml - 1 : -1);
    if (src < 0) continue;
    for (int j = pml; j < nx - pml; j++)
      memcpy(out + i * pln + (long)j * ny + pml, out + src * pln + (long)j * ny + pml, sizeof(float) * hy);
  }
  for (int i = 0; i < nz; i++)
    for (int j = 0; j < nx; j++) {
      int src = j < pml ? pml : (j >= nx - pml ? nx - pml - 1 : -1);
      if (src < 0) continue;
      memcpy(out + i * pln + (long)j * ny + pml, out + i * pln + (long)src * ny + pml, sizeof(float) * hy);
    }
  for (int i = 0; i < nz; i++)
    for (int j = 0; j < nx; j++) {
      float *row = out + i * pln + (long)j * ny;
      for (int k = 0; k < pml; k++) row[k] = row[pml];
      for (int k = ny - pml; k < ny; k++) row[k] = row[ny - pml - 1];
    }
}

/* CPU_function.c:10-22 absMaxval over n cells ('>=' compare; NaN never wins). */
float qp3d_absmax(const float *a, long n)
{
  float m = 0.0f;
  for (long i = 0; i < n; i++) {
    float v = fabsf(a[i]);
    if (v >= m) m = v;
  }
  return m;
}

/* CPU_function.c:234-261 readmodel3d: read the window [fny,fny+ny) x [fnx,fnx+nx) x [fnz,fnz+nz)
 * of a model file stored [vny][vnx][vnz] (z fastest, native float32) into a[z][x][y].
 * Returns 0, or a negative status if the file is missing, short, or the window is outside it. */
int qp3d_readmodel3d(const char *filename, float *a, int ny, int nx, int nz, int fny, int fnx, int fnz, int vny,
                     int vnx, int vnz)
{
  if (fny < 0 || fnx < 0 || fnz < 0 || fny + ny > vny || fnx + nx > vnx || fnz + nz > vnz) {
    qp3d_set_error("readmodel3d: window [%d,%d)x[%d,%d)x[%d,%d) outside model %dx%dx%d (y,x,z) of %s", fny,
                   fny + ny, fnx, fnx + nx, fnz, fnz + nz, vny, vnx, vnz, filename);
    return QP3D_ERR_ARG;
  }
  FILE *fp = fopen(filename, "rb");
  if (!fp) {
    qp3d_set_error("readmodel3d: cannot open %s", filename);
    return QP3D_ERR_IO;
  }
  float *line = (float *)malloc(sizeof(float) * (size_t)nz);
  int rc = 0;
  for (int i = 0; i < ny && rc == 0; i++)
    for (int j = 0; j < nx; j++) {
      long off = (((long)(fny + i) * vnx + fnx + j) * vnz + fnz) * (long)sizeof(float);
      if (fseek(fp, off, SEEK_SET) != 0 || fread(line, sizeof(float), (size_t)nz, fp) != (size_t)nz) {
        qp3d_set_error("readmodel3d: %s is shorter than %dx%dx%d floats", filename, vny, vnx, vnz);
        rc = QP3D_ERR_IO;
        break;
      }
      for (int k = 0; k < nz; k++) a[(long)k * nx * ny + (long)j * ny + i] = line[k];
    }
  free(line);
  fclose(fp);
  return rc;
}

/* ---- synthetic layered VTI model ----------------------------------------------------------
 * Four equal-thickness horizontal layers over nz (layer = z*4/nz) with
 *   vp 2000/2500/3000/3500 m/s, epsilon 0/0.10/0.15/0.20, delta 0/0.05/0.08/0.10,
 * each field multiplied by (1 + 0.03 g), g an independent smooth random field per field
 * (drawn in the order vp, epsilon, delta from one SplitMix64 stream, cells in file order,
 * one Box-Muller pair per two cells;
 * separable Gaussian filter with sigma = 5 cells truncated at 4 sigma, half-sample
 * reflection at the edges; scaled to max|g| = 1), then delta <- min(delta, epsilon).
 * Output arrays are in the model-file layout [ny][nx][nz] (z fastest). */
/* Draw d (0-based) of the SplitMix64 stream seeded with s: the generator's state after d+1
 * increments, mixed. Counter-based, so the stream can be filled in parallel. */
static uint64_t splitmix64_at(uint64_t s, uint64_t d)
{
  uint64_t z = s + (d + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static double unit_open(uint64_t s, uint64_t d) /* uniform in (0, 1] */
{
  return ((double)(splitmix64_at(s, d) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

static int reflect_index(int i, int n)
{
  if (n == 1) return 0;
  int p = 2 * n;
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - 1 - i;
}

/* Smooth one field g[ny][nx][nz] in place along each axis. */
static void gauss_smooth3(double *g, int ny, int nx, int nz, double sigma)
{
  int rad = (int)ceil(4.0 * sigma);
  double *w = (double *)malloc(sizeof(double) * (2 * rad + 1));
  double sum = 0;
  for (int k = -rad; k <= rad; k++) sum += (w[k + rad] = exp(-0.5 * k * k / (sigma * sigma)));
  for (int k = 0; k <= 2 * rad; k++) w[k] /= sum;
  int nmax = ny > nx ? ny : nx;
  if (nz > nmax) nmax = nz;
  const long sy = (long)nx * nz, sx = nz, sz = 1;
  for (int axis = 0; axis < 3; axis++) {
    int n = axis == 0 ? nz : (axis == 1 ? nx : ny);
    long stride = axis == 0 ? sz : (axis == 1 ? sx : sy);
    long nlines = (long)ny * nx * nz / n;
#pragma omp parallel
    {
      double *ext = (double *)malloc(sizeof(double) * (nmax + 2 * rad));
#pragma omp for schedule(static)
      for (long l = 0; l < nlines; l++) {
        long base;
        if (axis == 0) base = l * nz;
        else if (axis == 1) base = (l / nz) * sy + (l % nz);
        else base = l;
        /* line with rad reflected samples on each side */
        for (int i = -rad; i < n + rad; i++) ext[i + rad] = g[base + reflect_index(i, n) * stride];
        for (int i = 0; i < n; i++) {
          const double *e = ext + i;
          double acc = 0;
          for (int k = 0; k <= 2 * rad; k++) acc += w[k] * e[k];
          g[base + i * stride] = acc;
        }
      }
      free(ext);
    }
  }
  free(w);
}

int qp3d_genmodel(int nx, int ny, int nz, uint64_t seed, float sigma_cells, float amp_frac, float *vp, float *eps,
                  float *del)
{
  static const double lv[4] = {2000, 2500, 3000, 3500}, le[4] = {0, 0.10, 0.15, 0.20},
                      ld[4] = {0, 0.05, 0.08, 0.10};
  const long n = (long)nx * ny * nz;
  double *g = (double *)malloc(sizeof(double) * n);
  if (!g) {
    qp3d_set_error("genmodel: out of host memory for %ld cells", n);
    return QP3D_ERR_ALLOC;
  }
  float *outs[3] = {vp, eps, del};
  const double *lay[3] = {lv, le, ld};
  const long npair = (n + 1) / 2;
  for (int f = 0; f < 3; f++) {
    /* pair j of field f uses draws 2*(f*npair + j) and 2*(f*npair + j) + 1 */
#pragma omp parallel for schedule(static)
    for (long j = 0; j < npair; j++) {
      const uint64_t d = 2 * ((uint64_t)f * npair + j);
      double r = sqrt(-2.0 * log(unit_open(seed, d)));
      double th = 2.0 * QP3D_PI * unit_open(seed, d + 1);
      g[2 * j] = r * cos(th);
      if (2 * j + 1 < n) g[2 * j + 1] = r * sin(th);
    }
    if (sigma_cells > 0) gauss_smooth3(g, ny, nx, nz, sigma_cells);
    double gm = 0;
    for (long i = 0; i < n; i++) gm = fmax(gm, fabs(g[i]));
    double sc = gm > 0 ? 1.0 / gm : 0.0;
    for (long i = 0; i < n; i++) {
      int z = (int)(i % nz);
      int layer = (int)((long)z * 4 / nz);
      outs[f][i] = (float)(lay[f][layer] * (1.0 + amp_frac * g[i] * sc));
    }
  }
  for (long i = 0; i < n; i++)
    if (del[i] > eps[i]) del[i] = eps[i];
  free(g);
  return 0;
}
