// Harness around the reference's own CPU kernels (CPU_function.c, compiled from
// /root/reference by oracle/Makefile into oracle/_ref/libqp3d_ref.so).
// TEST INFRASTRUCTURE ONLY: it exists to pin the oracle against the reference.
//
// The reference's kernels are called unmodified. The harness supplies what the reference
// CPU path gets wrong (CPU_function.c:481-563 exchange reads an uninitialized index):
// a rotation loop with the GPU exchange's semantics (GPU_kernel.cu:733-812), zeroed
// state, and PML arrays with the GPU sizes (3dPureqPModeling_single_shot.cu:122-186).
#include <cstdlib>
#include <cstring>

#include "CPU_function.h"

namespace {
struct RefState {
  int nx, ny, nz, pml;
  float dx, dy, dz, dt;
  float *vp, *eps, *del, *src, *dlr, *ddlr, *dtb, *ddtb, *dfb, *ddfb, *c, *c2;
  float *u1, *u2, *u3, *ux, *uy, *uz, *S;
  float *aux[6][10];
  long aux_n[6];
  float G;
};
}  // namespace

extern "C" {

void *ref_create(int nx, int ny, int nz, int pml, float dx, float dy, float dz, float dt, float *vp, float *eps,
                 float *del, float *src, float *dlr, float *ddlr, float *dtb, float *ddtb, float *dfb, float *ddfb,
                 float *c, float *c2)
{
  RefState *s = (RefState *)calloc(1, sizeof(RefState));
  s->nx = nx; s->ny = ny; s->nz = nz; s->pml = pml;
  s->dx = dx; s->dy = dy; s->dz = dz; s->dt = dt;
  s->vp = vp; s->eps = eps; s->del = del; s->src = src;
  s->dlr = dlr; s->ddlr = ddlr; s->dtb = dtb; s->ddtb = ddtb; s->dfb = dfb; s->ddfb = ddfb;
  s->c = c; s->c2 = c2;
  long n = (long)nx * ny * nz;
  float **f[7] = {&s->u1, &s->u2, &s->u3, &s->ux, &s->uy, &s->uz, &s->S};
  for (int k = 0; k < 7; k++) *f[k] = (float *)calloc(n, sizeof(float));
  s->aux_n[0] = s->aux_n[1] = (long)pml * ny * nz;
  s->aux_n[2] = s->aux_n[3] = (long)pml * nx * nz;
  s->aux_n[4] = s->aux_n[5] = (long)pml * nx * ny;
  for (int fc = 0; fc < 6; fc++)
    for (int k = 0; k < 10; k++) s->aux[fc][k] = (float *)calloc(s->aux_n[fc], sizeof(float));
  return s;
}

void ref_destroy(void *h)
{
  RefState *s = (RefState *)h;
  if (!s) return;
  free(s->u1); free(s->u2); free(s->u3); free(s->ux); free(s->uy); free(s->uz); free(s->S);
  for (int fc = 0; fc < 6; fc++)
    for (int k = 0; k < 10; k++) free(s->aux[fc][k]);
  free(s);
}

static void rotate(RefState *s)
{
  const int nx = s->nx, ny = s->ny, nz = s->nz, pml = s->pml;
  for (int z = 0; z < nz; z++)
    for (int x = 0; x < nx; x++)
      for (int y = 0; y < ny; y++) {
        long i = (long)z * nx * ny + (long)x * ny + y;
        s->u1[i] = s->u2[i];
        s->u2[i] = s->u3[i];
        long j[6] = {(long)z * ny * pml + (long)x * ny + y,
                     (long)z * ny * pml + (long)(x - nx + pml) * ny + y,
                     (long)z * nx * pml + (long)x * pml + y,
                     (long)z * nx * pml + (long)x * pml + (y - ny + pml),
                     i,
                     (long)(z - nz + pml) * nx * ny + (long)x * ny + y};
        int in[6] = {x < pml, x >= nx - pml, y < pml, y >= ny - pml, z < pml, z >= nz - pml};
        for (int f = 0; f < 6; f++)
          if (in[f]) {
            float **a = s->aux[f];
            long q = j[f];
            a[0][q] = a[1][q]; a[1][q] = a[2][q];
            a[4][q] = a[5][q]; a[5][q] = a[6][q];
            a[7][q] = a[8][q]; a[8][q] = a[9][q];
          }
      }
}

// One step in the order of modeling3D.c:53-81, with the reference's kernels.
void ref_step(void *h, float w)
{
  RefState *s = (RefState *)h;
  const int nx = s->nx, ny = s->ny, nz = s->nz, pml = s->pml;
  grad(s->u2, s->ux, s->uy, s->uz, s->c, nx, ny, nz, s->dx, s->dy, s->dz);
  float gx = absMaxval(s->ux, nx, ny, nz);
  float gy = absMaxval(s->uy, nx, ny, nz);
  float gz = absMaxval(s->uz, nx, ny, nz);
  float G = (gx > gy) ? gx : gy;
  G = (G > gz) ? G : gz;
  s->G = G;
  scalar_operator(G, s->ux, s->uy, s->uz, s->eps, s->del, s->S, nx, ny, nz);
  float **L = s->aux[0], **R = s->aux[1], **F = s->aux[2], **K = s->aux[3], **T = s->aux[4], **B = s->aux[5];
  wavefield_update(s->c, s->c2, s->dlr, s->ddlr, s->dtb, s->ddtb, s->dfb, s->ddfb, s->eps, s->del, s->vp, s->dx,
                   s->dy, s->dz, s->dt, nx, ny, nz, pml, 0, s->ux, s->uy, s->uz, s->u1, s->u3, s->u2, s->S,
                   L[0], L[1], L[2], L[3], L[4], L[5], L[6], L[7], L[8], L[9],
                   R[0], R[1], R[2], R[3], R[4], R[5], R[6], R[7], R[8], R[9],
                   T[0], T[1], T[2], T[3], T[4], T[5], T[6], T[7], T[8], T[9],
                   B[0], B[1], B[2], B[3], B[4], B[5], B[6], B[7], B[8], B[9],
                   F[0], F[1], F[2], F[3], F[4], F[5], F[6], F[7], F[8], F[9],
                   K[0], K[1], K[2], K[3], K[4], K[5], K[6], K[7], K[8], K[9]);
  addsource(s->src, w, s->u3, nx, ny, nz);
  rotate(s);
}

float *ref_field(void *h, int which)
{
  RefState *s = (RefState *)h;
  float *f[7] = {s->u1, s->u2, s->u3, s->ux, s->uy, s->uz, s->S};
  return (which >= 0 && which < 7) ? f[which] : nullptr;
}
float *ref_aux(void *h, int face, int arr) { return ((RefState *)h)->aux[face][arr]; }
long ref_aux_len(void *h, int face) { return ((RefState *)h)->aux_n[face]; }
double ref_last_G(void *h) { return ((RefState *)h)->G; }

// Setup tables, straight through to the reference.
void ref_fdcoff1(float *c) { fdcoff1(c); }
void ref_fdcoff2(float *c) { fdcoff2(c); }
void ref_pmlcoff(int pml, float vpmax, float dx, float dy, float dz, float *dlr, float *ddlr, float *dtb,
                 float *ddtb, float *dfb, float *ddfb)
{
  pmlcoff(pml, vpmax, dx, dy, dz, dlr, ddlr, dtb, ddtb, dfb, ddfb);
}
void ref_source(int sx, int sy, int sz, float fm, float amp, float alpha, float dt, float dx, float dy, float dz,
                float t0, int nt, int nx, int ny, int nz, float *source, float *wavelet)
{
  Source(sx, sy, sz, fm, amp, alpha, dt, dx, dy, dz, t0, nt, nx, ny, nz, source, wavelet);
}
void ref_extend_model(float *in, float *out, int nx, int ny, int nz, int pml) { extend_model(in, out, nx, ny, nz, pml); }
float ref_absmax(float *a, int nx, int ny, int nz) { return absMaxval(a, nx, ny, nz); }
void ref_readmodel3d(const char *fn, float *a, int ny, int nx, int nz, int fny, int fnx, int fnz, int vny, int vnx,
                     int vnz)
{
  readmodel3d(fn, a, ny, nx, nz, fny, fnx, fnz, vny, vnx, vnz);
}

}  // extern "C"
