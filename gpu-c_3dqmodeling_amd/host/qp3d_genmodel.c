/*
 * qp3d_genmodel NX NY NZ SEED OUTDIR [SIGMA AMP]
 * Writes OUTDIR/vp.dat, OUTDIR/epsilon.dat, OUTDIR/delta.dat in the reference's model-file
 * layout (native float32, [NY][NX][NZ] with z fastest, as readmodel3d reads them,
 * CPU_function.c:244-251): four horizontal layers plus a smooth random perturbation
 * (qp3d_genmodel in the library).
 */
#include <stdio.h>
#include <stdlib.h>

#include "qp3d.h"

int main(int argc, char **argv)
{
  if (argc < 6) {
    fprintf(stderr, "usage: %s NX NY NZ SEED OUTDIR [SIGMA AMP]\n", argv[0]);
    return 2;
  }
  int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]);
  unsigned long long seed = strtoull(argv[4], NULL, 10);
  float sigma = argc > 6 ? (float)atof(argv[6]) : 5.0f, amp = argc > 7 ? (float)atof(argv[7]) : 0.03f;
  if (nx < 1 || ny < 1 || nz < 1) {
    fprintf(stderr, "bad dimensions\n");
    return 2;
  }
  size_t n = (size_t)nx * ny * nz;
  float *f[3];
  for (int i = 0; i < 3; i++)
    if (!(f[i] = (float *)malloc(sizeof(float) * n))) {
      fprintf(stderr, "out of memory\n");
      return 3;
    }
  if (qp3d_genmodel(nx, ny, nz, seed, sigma, amp, f[0], f[1], f[2])) {
    fprintf(stderr, "%s\n", qp3d_last_error());
    return 3;
  }
  const char *names[3] = {"vp.dat", "epsilon.dat", "delta.dat"};
  for (int i = 0; i < 3; i++) {
    char path[4096];
    snprintf(path, sizeof path, "%s/%s", argv[5], names[i]);
    FILE *fp = fopen(path, "wb");
    if (!fp || fwrite(f[i], sizeof(float), n, fp) != n) {
      fprintf(stderr, "cannot write %s\n", path);
      return 5;
    }
    fclose(fp);
    free(f[i]);
  }
  return 0;
}
