/*
 * Templated body of the CPU oracle for the 3D pure-qP VTI time loop.
 * TEST INFRASTRUCTURE ONLY: it is the checker for the HIP path, never part of it.
 *
 * Included twice by qp3d_oracle.c: once with REAL=float (the reference's own
 * FP32 arithmetic, including every place where the reference promotes to
 * double through pow()/double literals) and once with REAL=double (an FP64
 * restatement of the same physics).
 *
 * Layout follows the reference: fields are [nz][nx][ny], y unit-stride,
 * linear index z*nx*ny + x*ny + y (GPU_kernel.cu:551, array_new.c:2).
 * PML state follows the reference's per-face arrays with the GPU sizes
 * (3dPureqPModeling_single_shot.cu:122-186):
 *   L/R  [nz][pml][ny]   index z*ny*pml + x'*ny + y
 *   F/Bk [nz][nx][pml]   index z*nx*pml + x*pml + y'
 *   T/Bo [pml][nx][ny]   index z'*nx*ny + x*ny + y
 * and ten arrays per face: w11 w12 w13 w21 w31 w32 w33 p1 p2 p3.
 *
 * Required macros: REAL, SFX(name), SQRTR (sqrt of REAL that matches the
 * C++ overload the reference resolves to for a float argument).
 */

#define ST SFX(oracle_state)

typedef struct ST {
  int nx, ny, nz, pml;
  float dx, dy, dz, dt;
  const float *vp, *eps, *del, *src;
  const float *dlr, *ddlr, *dtb, *ddtb, *dfb, *ddfb, *c, *c2;
  REAL *u1, *u2, *u3, *ux, *uy, *uz, *S;
  REAL *aux[6][10];
  long aux_n[6];
  REAL G;
  int nthreads;
  /* frozen decisions (test hook): dmode 1 records, per step, one bit per cell for
     |grad u| > 0.01*G into dbits; dmode 2 takes each cell's decision from dbits instead of
     comparing (the threshold G and every cell's side of it held at a base run's) */
  unsigned char *dbits;
  int dmode;
  long kstep;
} ST;

/* CPU_function.c:10-22 absMaxval: max of fabs over all cells; '>=' compare, so NaN never wins. */
REAL SFX(oracle_absmax)(const REAL *a, long n, int nthreads)
{
  REAL m = 0;
#pragma omp parallel for reduction(max : m) num_threads(nthreads) schedule(static)
  for (long i = 0; i < n; i++) {
    REAL v = a[i] < 0 ? -a[i] : a[i];
    if (v >= m) m = v;
  }
  return m;
}

/* CPU_function.c:301-324 grad: 10th-order first derivatives on [5,n-5)^3, each term divided separately. */
void SFX(oracle_grad)(const REAL *u2, REAL *ux, REAL *uy, REAL *uz, const float *c, int nx, int ny, int nz,
                      float dx, float dy, float dz, int nthreads)
{
  const long pln = (long)nx * ny;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int z = 5; z < nz - 5; z++)
    for (int x = 5; x < nx - 5; x++)
      for (int y = 5; y < ny - 5; y++) {
        long i = z * pln + (long)x * ny + y;
        REAL Ux = 0, Uy = 0, Uz = 0;
        for (int m = 1; m <= 5; m++) {
          REAL cm = c[30 + m];
          Ux = Ux + cm * (u2[i + (long)m * ny] - u2[i - (long)m * ny]) / (REAL)dx;
          Uy = Uy + cm * (u2[i + m] - u2[i - m]) / (REAL)dy;
          Uz = Uz + cm * (u2[i + m * pln] - u2[i - m * pln]) / (REAL)dz;
        }
        ux[i] = Ux;
        uy[i] = Uy;
        uz[i] = Uz;
      }
}

/* CPU_function.c:339-370 / GPU_kernel.cu:87-115 scalar_operator. S=1 everywhere; on [5,n-5)^3
 * the direction n = grad/(|grad|+2.22e-17) if |grad| > 0.01*G (double compare), else 0, and
 * S = 0.5*(1+sqrt(1 - 8(eps-del)(nx^2+ny^2)nz^2 / (1+2eps(nx^2+ny^2))^2)).
 * The numerator is REAL arithmetic, the denominator and the outer sqrt are double (pow). */
void SFX(oracle_scalar_operator)(REAL G, const REAL *ux, const REAL *uy, const REAL *uz, const float *eps,
                                 const float *del, REAL *S, int nx, int ny, int nz, int nthreads, unsigned char *dbits,
                                 int dmode)
{
  const long pln = (long)nx * ny;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int z = 0; z < nz; z++)
    for (int x = 0; x < nx; x++)
      for (int y = 0; y < ny; y++) {
        long i = z * pln + (long)x * ny + y;
        S[i] = 1;
        if (y >= 5 && y < ny - 5 && x >= 5 && x < nx - 5 && z >= 5 && z < nz - 5) {
          REAL mode = SQRTR(ux[i] * ux[i] + uy[i] * uy[i] + uz[i] * uz[i]);
          REAL nxv, nyv, nzv;
          int above = (double)mode > 1.0e-02 * (double)G;
          if (dmode == 1 && above) {
#pragma omp atomic
            dbits[i >> 3] |= (unsigned char)(1u << (i & 7));
          }
          if (dmode == 2) above = (dbits[i >> 3] >> (i & 7)) & 1;
          if (above) {
            nxv = (REAL)((double)ux[i] / ((double)mode + 2.22e-17));
            nyv = (REAL)((double)uy[i] / ((double)mode + 2.22e-17));
            nzv = (REAL)((double)uz[i] / ((double)mode + 2.22e-17));
          } else {
            nxv = 0;
            nyv = 0;
            nzv = 0;
          }
          REAL e = (REAL)eps[i], d = (REAL)del[i];
          REAL q = nxv * nxv + nyv * nyv;
          REAL num = 8 * (e - d) * q * nzv * nzv;
          double den = pow((double)(1 + 2 * e * q), 2);
          S[i] = (REAL)(0.5 * (1 + sqrt(1 - (double)num / den)));
        }
      }
}

/* One PML face cell, CPU_function.c:405-476 / GPU_kernel.cu:566-637.
 * cn_is_A: the normal coefficient is (1+2eps)*S*(vp*dt)^2 (x and y faces) or S*(vp*dt)^2 (z faces).
 * sgn: +1 on the low faces (L, F, T), -1 on the high faces (R, Bk, Bo). */
static inline REAL SFX(pml_cell)(REAL e, REAL S, REAL vp, REAL dt, REAL Unn, REAL T, REAL un, int cn_is_A, int sgn,
                                 REAL d, REAL dd, REAL *w11, REAL *w12, REAL *w13, REAL *w21, REAL *w31, REAL *w32,
                                 REAL *w33, REAL *p1, REAL *p2, REAL *p3, long j)
{
  double pv = pow((double)(vp * dt), 2);
  double Cn = cn_is_A ? (double)((1 + 2 * e) * S) * pv : (double)S * pv;
  double k1 = pow((double)(d * dt + 1), 2) - 3;
  REAL k2 = 2 * d * dt - 1;
  w13[j] = (REAL)(Cn * (double)Unn - k1 * (double)w12[j] + (double)(k2 * w11[j]));
  double pterm = Cn * (double)dd * (double)un;
  double pbase = (3 - pow((double)(1 + dt * d), 2)) * (double)p2[j] + (double)(k2 * p1[j]);
  p3[j] = (REAL)(sgn > 0 ? pbase + pterm : pbase - pterm);
  w21[j] = dt * p3[j] + w21[j] * (1 - dt * d);
  double B = (double)S * pv;
  double tang = cn_is_A ? B * (double)T : (double)((1 + 2 * e) * S) * pv * (double)T;
  w33[j] = (REAL)(tang + (double)(2 * w32[j]) - (double)w31[j]);
  return w13[j] + w21[j] + w33[j];
}

/* CPU_function.c:371-480 / GPU_kernel.cu:534-639 wavefield_update on [5,n-5)^3.
 * First matching branch wins: interior, L, R, F, Bk, T, Bo. */
void SFX(oracle_wavefield_update)(ST *s)
{
  const int nx = s->nx, ny = s->ny, nz = s->nz, pml = s->pml;
  const long pln = (long)nx * ny;
  const REAL dt = s->dt;
  const REAL dx2 = (REAL)s->dx * (REAL)s->dx, dy2 = (REAL)s->dy * (REAL)s->dy, dz2 = (REAL)s->dz * (REAL)s->dz;
  const float *c2 = s->c2;
  REAL *u1 = s->u1, *u2 = s->u2, *u3 = s->u3, *S = s->S;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int z = 5; z < nz - 5; z++)
    for (int x = 5; x < nx - 5; x++)
      for (int y = 5; y < ny - 5; y++) {
        long i = z * pln + (long)x * ny + y;
        REAL Uxx = 0, Uyy = 0, Uzz = 0;
        for (int m = 1; m <= 5; m++) {
          REAL cm = c2[30 + m];
          Uxx = Uxx + cm * (u2[i + (long)m * ny] - 2 * u2[i] + u2[i - (long)m * ny]) / dx2;
          Uyy = Uyy + cm * (u2[i + m] - 2 * u2[i] + u2[i - m]) / dy2;
          Uzz = Uzz + cm * (u2[i + m * pln] - 2 * u2[i] + u2[i - m * pln]) / dz2;
        }
        REAL e = (REAL)s->eps[i], vp = (REAL)s->vp[i], Si = S[i];
        if (x >= pml && x < nx - pml && y >= pml && y < ny - pml && z >= pml && z < nz - pml) {
          u3[i] = 2 * u2[i] - u1[i] + dt * dt * ((1 + 2 * e) * (Uxx + Uyy) + Uzz) * Si * vp * vp;
        } else if (x < pml) {
          REAL **a = s->aux[0];
          long j = (long)z * ny * pml + (long)x * ny + y;
          u3[i] = SFX(pml_cell)(e, Si, vp, dt, Uxx, Uzz + (1 + 2 * e) * Uyy, s->ux[i], 1, +1, s->dlr[pml - x],
                                s->ddlr[pml - x], a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], j);
        } else if (x >= nx - pml) {
          REAL **a = s->aux[1];
          long j = (long)z * ny * pml + (long)(x - nx + pml) * ny + y;
          u3[i] = SFX(pml_cell)(e, Si, vp, dt, Uxx, Uzz + (1 + 2 * e) * Uyy, s->ux[i], 1, -1, s->dlr[x - nx + pml],
                                s->ddlr[x - nx + pml], a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], j);
        } else if (y < pml) {
          REAL **a = s->aux[2];
          long j = (long)z * nx * pml + (long)x * pml + y;
          u3[i] = SFX(pml_cell)(e, Si, vp, dt, Uyy, Uzz + (1 + 2 * e) * Uxx, s->uy[i], 1, +1, s->dfb[pml - y],
                                s->ddfb[pml - y], a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], j);
        } else if (y >= ny - pml) {
          REAL **a = s->aux[3];
          long j = (long)z * nx * pml + (long)x * pml + (y - ny + pml);
          u3[i] = SFX(pml_cell)(e, Si, vp, dt, Uyy, Uzz + (1 + 2 * e) * Uxx, s->uy[i], 1, -1, s->dfb[y - ny + pml],
                                s->ddfb[y - ny + pml], a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], j);
        } else if (z < pml) {
          REAL **a = s->aux[4];
          long j = (long)z * nx * ny + (long)x * ny + y;
          u3[i] = SFX(pml_cell)(e, Si, vp, dt, Uzz, Uxx + Uyy, s->uz[i], 0, +1, s->dtb[pml - z], s->ddtb[pml - z],
                                a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], j);
        } else {
          REAL **a = s->aux[5];
          long j = (long)(z - nz + pml) * nx * ny + (long)x * ny + y;
          u3[i] = SFX(pml_cell)(e, Si, vp, dt, Uzz, Uxx + Uyy, s->uz[i], 0, -1, s->dtb[z - nz + pml],
                                s->ddtb[z - nz + pml], a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], j);
        }
      }
}

/* CPU_function.c:326-338 addsource: u3 += source*wavelet over every cell. */
void SFX(oracle_addsource)(const float *src, float w, REAL *u3, long n, int nthreads)
{
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (long i = 0; i < n; i++) u3[i] += (REAL)src[i] * (REAL)w;
}

/* GPU_kernel.cu:733-812 exchange (the CPU twin at CPU_function.c:481-563 reads an uninitialized
 * index): u1<-u2, u2<-u3 everywhere; per face region, as independent tests,
 * w11<-w12<-w13, w31<-w32<-w33, p1<-p2<-p3; w21 is not rotated. */
void SFX(oracle_exchange)(ST *s)
{
  const int nx = s->nx, ny = s->ny, nz = s->nz, pml = s->pml;
  const long pln = (long)nx * ny;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int z = 0; z < nz; z++)
    for (int x = 0; x < nx; x++)
      for (int y = 0; y < ny; y++) {
        long i = z * pln + (long)x * ny + y;
        s->u1[i] = s->u2[i];
        s->u2[i] = s->u3[i];
        long j[6];
        int in[6];
        in[0] = x < pml;
        j[0] = (long)z * ny * pml + (long)x * ny + y;
        in[1] = x >= nx - pml;
        j[1] = (long)z * ny * pml + (long)(x - nx + pml) * ny + y;
        in[2] = y < pml;
        j[2] = (long)z * nx * pml + (long)x * pml + y;
        in[3] = y >= ny - pml;
        j[3] = (long)z * nx * pml + (long)x * pml + (y - ny + pml);
        in[4] = z < pml;
        j[4] = i;
        in[5] = z >= nz - pml;
        j[5] = (long)(z - nz + pml) * nx * ny + (long)x * ny + y;
        for (int f = 0; f < 6; f++)
          if (in[f]) {
            REAL **a = s->aux[f];
            long q = j[f];
            a[0][q] = a[1][q];
            a[1][q] = a[2][q];
            a[4][q] = a[5][q];
            a[5][q] = a[6][q];
            a[7][q] = a[8][q];
            a[8][q] = a[9][q];
          }
      }
}

ST *SFX(oracle_create)(int nx, int ny, int nz, int pml, float dx, float dy, float dz, float dt, const float *vp,
                       const float *eps, const float *del, const float *src, const float *dlr, const float *ddlr,
                       const float *dtb, const float *ddtb, const float *dfb, const float *ddfb, const float *c,
                       const float *c2, int nthreads)
{
  ST *s = (ST *)calloc(1, sizeof(ST));
  if (!s) return NULL;
  s->nx = nx; s->ny = ny; s->nz = nz; s->pml = pml;
  s->dx = dx; s->dy = dy; s->dz = dz; s->dt = dt;
  s->vp = vp; s->eps = eps; s->del = del; s->src = src;
  s->dlr = dlr; s->ddlr = ddlr; s->dtb = dtb; s->ddtb = ddtb; s->dfb = dfb; s->ddfb = ddfb;
  s->c = c; s->c2 = c2;
  s->nthreads = nthreads > 0 ? nthreads : 1;
  long n = (long)nx * ny * nz;
  REAL **f[7] = {&s->u1, &s->u2, &s->u3, &s->ux, &s->uy, &s->uz, &s->S};
  for (int k = 0; k < 7; k++) {
    *f[k] = (REAL *)calloc(n, sizeof(REAL));
    if (!*f[k]) return NULL;
  }
  s->aux_n[0] = s->aux_n[1] = (long)pml * ny * nz;
  s->aux_n[2] = s->aux_n[3] = (long)pml * nx * nz;
  s->aux_n[4] = s->aux_n[5] = (long)pml * nx * ny;
  for (int fc = 0; fc < 6; fc++)
    for (int k = 0; k < 10; k++) {
      s->aux[fc][k] = (REAL *)calloc(s->aux_n[fc], sizeof(REAL));
      if (!s->aux[fc][k]) return NULL;
    }
  return s;
}

void SFX(oracle_destroy)(ST *s)
{
  if (!s) return;
  free(s->u1); free(s->u2); free(s->u3); free(s->ux); free(s->uy); free(s->uz); free(s->S);
  for (int fc = 0; fc < 6; fc++)
    for (int k = 0; k < 10; k++) free(s->aux[fc][k]);
  free(s);
}

/* One time step in the order of modeling3D.c:53-81: grad, max, S, update, source, exchange.
 * Gthr < 0: the threshold of S is 0.01*G of this step (the reference). Gthr >= 0 plants a
 * mistake for the tests: S takes its threshold from Gthr (s->G is still this step's G). */
static void SFX(oracle_step_impl)(ST *s, float w, double Gthr);
void SFX(oracle_step)(ST *s, float w) { SFX(oracle_step_impl)(s, w, -1.0); }
void SFX(oracle_step_planted_G)(ST *s, float w, double Gthr) { SFX(oracle_step_impl)(s, w, Gthr); }
static void SFX(oracle_step_impl)(ST *s, float w, double Gthr)
{
  long n = (long)s->nx * s->ny * s->nz;
  SFX(oracle_grad)(s->u2, s->ux, s->uy, s->uz, s->c, s->nx, s->ny, s->nz, s->dx, s->dy, s->dz, s->nthreads);
  REAL gx = SFX(oracle_absmax)(s->ux, n, s->nthreads);
  REAL gy = SFX(oracle_absmax)(s->uy, n, s->nthreads);
  REAL gz = SFX(oracle_absmax)(s->uz, n, s->nthreads);
  REAL G = gx > gy ? gx : gy;
  G = G > gz ? G : gz;
  s->G = G;
  if (Gthr >= 0) G = (REAL)Gthr;
  unsigned char *bits = s->dmode ? s->dbits + s->kstep * ((n + 7) / 8) : NULL;
  if (s->dmode == 1) memset(bits, 0, (size_t)((n + 7) / 8));
  SFX(oracle_scalar_operator)(G, s->ux, s->uy, s->uz, s->eps, s->del, s->S, s->nx, s->ny, s->nz, s->nthreads, bits,
                              s->dmode);
  SFX(oracle_wavefield_update)(s);
  SFX(oracle_addsource)(s->src, w, s->u3, n, s->nthreads);
  SFX(oracle_exchange)(s);
  s->kstep++;
}

/* Frozen decisions (test hook; see ST): mode 1 records the decisions of the following steps
 * into bits (ceil(N/8) bytes per step, for as many steps as will be run), mode 2 replays them,
 * mode 0 turns the hook off. The step count restarts at 0. */
void SFX(oracle_set_decisions)(ST *s, unsigned char *bits, int mode)
{
  s->dbits = bits;
  s->dmode = bits ? mode : 0;
  s->kstep = 0;
}

/* nt steps with the given wavelet samples; record[k] = u(sz) interior plane after step k. */
void SFX(oracle_run_record)(ST *s, const float *wavelet, int nt, int sz, REAL *record)
{
  const int hx = s->nx - 2 * s->pml, hy = s->ny - 2 * s->pml;
  const long pln = (long)s->nx * s->ny;
  for (int k = 0; k < nt; k++) {
    SFX(oracle_step)(s, wavelet[k]);
    for (int x = 0; x < hx; x++)
      for (int y = 0; y < hy; y++)
        record[(long)k * hx * hy + (long)x * hy + y] = s->u2[sz * pln + (long)(x + s->pml) * s->ny + (y + s->pml)];
  }
}

/* Field access for tests: 0..6 = u1 u2 u3 ux uy uz S. */
REAL *SFX(oracle_field)(ST *s, int which)
{
  switch (which) {
  case 0: return s->u1;
  case 1: return s->u2;
  case 2: return s->u3;
  case 3: return s->ux;
  case 4: return s->uy;
  case 5: return s->uz;
  case 6: return s->S;
  }
  return NULL;
}

REAL *SFX(oracle_aux)(ST *s, int face, int arr) { return s->aux[face][arr]; }
long SFX(oracle_aux_len)(ST *s, int face) { return s->aux_n[face]; }
double SFX(oracle_last_G)(ST *s) { return (double)s->G; }

/* modeling3D.c:17-153: nt steps, record after each step (:83-87), snapshots when
 * snapflag==1 and k%ntsnap==0 (:89-124). record is [nt][nx-2pml][ny-2pml]. Snapshot files are
 * written only by the float instance, byte-compatible with read_write.c:29-52; with
 * QP3D_SNAP_S=1 in the environment the S field's files too. */
int SFX(oracle_modeling3d)(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt, int pml,
                           int snapflag, int sz, const float *vp, const float *eps, const float *del,
                           const float *source, const float *wavelet, REAL *record, const float *dlr,
                           const float *ddlr, const float *dtb, const float *ddtb, const float *dfb, const float *ddfb,
                           const float *c, const float *c2, const char *snap_file, int nthreads)
{
  ST *s = SFX(oracle_create)(nx, ny, nz, pml, dx, dy, dz, dt, vp, eps, del, source, dlr, ddlr, dtb, ddtb, dfb, ddfb,
                             c, c2, nthreads);
  if (!s) return -1;
  const int hx = nx - 2 * pml, hy = ny - 2 * pml, hz = nz - 2 * pml;
  const long pln = (long)nx * ny;
  REAL *snap = NULL;
  if (snapflag == 1 && sizeof(REAL) == sizeof(float)) snap = (REAL *)calloc((long)hx * hy * hz, sizeof(REAL));
  /* QP3D_SNAP_S=1: also the S-field snapshot files, as the library's drop-in call writes them */
  const char *ss = getenv("QP3D_SNAP_S");
  const int snap_s = ss && atoi(ss) == 1;
  for (int k = 0; k < nt; k++) {
    SFX(oracle_step)(s, wavelet[k]);
    for (int x = pml; x < nx - pml; x++)
      for (int y = pml; y < ny - pml; y++)
        record[(long)k * hx * hy + (long)(x - pml) * hy + (y - pml)] = s->u2[sz * pln + (long)x * ny + y];
    if (snap && k % ntsnap == 0) {
      char name[1024];
      for (int z = 0; z < hz; z++)
        for (int x = 0; x < hx; x++)
          for (int y = 0; y < hy; y++)
            snap[((long)z * hx + x) * hy + y] = s->u2[(z + pml) * pln + (long)(x + pml) * ny + (y + pml)];
      snprintf(name, sizeof name, "%s%d.dat", snap_file, k);
      FILE *fp = fopen(name, "wb");
      if (fp) { fwrite(snap, sizeof(REAL), (size_t)hx * hy * hz, fp); fclose(fp); }
      REAL *xz = (REAL *)malloc(sizeof(REAL) * (size_t)hz * hx);
      REAL *yz = (REAL *)malloc(sizeof(REAL) * (size_t)hz * hy);
      REAL *xy = (REAL *)malloc(sizeof(REAL) * (size_t)hx * hy);
      for (int z = 0; z < hz; z++)
        for (int x = 0; x < hx; x++) xz[(long)z * hx + x] = snap[((long)z * hx + x) * hy + (hy - 1) / 2];
      for (int z = 0; z < hz; z++)
        for (int y = 0; y < hy; y++) yz[(long)z * hy + y] = snap[((long)z * hx + (hx - 1) / 2) * hy + y];
      for (int x = 0; x < hx; x++)
        for (int y = 0; y < hy; y++) xy[(long)x * hy + y] = snap[((long)((hz - 1) / 2) * hx + x) * hy + y];
      snprintf(name, sizeof name, "%s_xz%d.dat", snap_file, k);
      if ((fp = fopen(name, "wb"))) { fwrite(xz, sizeof(REAL), (size_t)hz * hx, fp); fclose(fp); }
      snprintf(name, sizeof name, "%s_yz%d.dat", snap_file, k);
      if ((fp = fopen(name, "wb"))) { fwrite(yz, sizeof(REAL), (size_t)hz * hy, fp); fclose(fp); }
      snprintf(name, sizeof name, "%s_xy%d.dat", snap_file, k);
      if ((fp = fopen(name, "wb"))) { fwrite(xy, sizeof(REAL), (size_t)hx * hy, fp); fclose(fp); }
      free(xz); free(yz); free(xy);
      if (snap_s) {
        /* the S field of this step over the padded grid and its centre slices, the writer the
           reference has commented out (3dPureqPModeling_single_shot.cu:381-409) */
        snprintf(name, sizeof name, "%s_S%d.dat", snap_file, k);
        if ((fp = fopen(name, "wb"))) { fwrite(s->S, sizeof(REAL), (size_t)nz * pln, fp); fclose(fp); }
        REAL *sxz = (REAL *)malloc(sizeof(REAL) * (size_t)nz * nx);
        REAL *syz = (REAL *)malloc(sizeof(REAL) * (size_t)nz * ny);
        REAL *sxy = (REAL *)malloc(sizeof(REAL) * (size_t)nx * ny);
        for (int z = 0; z < nz; z++)
          for (int x = 0; x < nx; x++) sxz[(long)z * nx + x] = s->S[z * pln + (long)x * ny + (ny - 1) / 2];
        for (int z = 0; z < nz; z++)
          for (int y = 0; y < ny; y++) syz[(long)z * ny + y] = s->S[z * pln + (long)((nx - 1) / 2) * ny + y];
        for (int x = 0; x < nx; x++)
          for (int y = 0; y < ny; y++) sxy[(long)x * ny + y] = s->S[((nz - 1) / 2) * pln + (long)x * ny + y];
        snprintf(name, sizeof name, "%s_Sxz%d.dat", snap_file, k);
        if ((fp = fopen(name, "wb"))) { fwrite(sxz, sizeof(REAL), (size_t)nz * nx, fp); fclose(fp); }
        snprintf(name, sizeof name, "%s_Syz%d.dat", snap_file, k);
        if ((fp = fopen(name, "wb"))) { fwrite(syz, sizeof(REAL), (size_t)nz * ny, fp); fclose(fp); }
        snprintf(name, sizeof name, "%s_Sxy%d.dat", snap_file, k);
        if ((fp = fopen(name, "wb"))) { fwrite(sxy, sizeof(REAL), (size_t)nx * ny, fp); fclose(fp); }
        free(sxz); free(syz); free(sxy);
      }
    }
  }
  free(snap);
  SFX(oracle_destroy)(s);
  return 0;
}

#undef ST
