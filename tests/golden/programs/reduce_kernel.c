/*
 * Kernel reductions: PSReduce(&v, op, kernel, dom, args...) with reduction
 * kernels that return float or double (the interface of the reference's
 * docs/programming.md, "Reduction"), on 3-D float and double grids, a 2-D
 * and a 1-D float grid, over interior, whole, small and empty domains.
 *
 *   reduce_kernel NX NY NZ IN BF16
 *
 * reads NX*NY*NZ floats (x fastest) from IN into the float grid g (stored
 * in BF16 when BF16 is 1), fills the double grid h with 0.5 * f + 0.25,
 * the 2-D grid a with the first plane of f and the 1-D grid b with its
 * first row, and prints one "name value" line per reduction.  Built by
 * tests/golden/build_corpus.py.
 */
#include <stdio.h>
#include <stdlib.h>

#include "physis/physis.h"

float sqdiff(const int x, const int y, const int z, PSGrid3DFloat g, float w) {
  float d = PSGridGet(g, x + 1, y, z) - PSGridGet(g, x, y, z);
  return w * d * d;
}

float curv(const int x, const int y, const int z, PSGrid3DFloat g) {
  float c = PSGridGet(g, x, y, z);
  float v = c - 0.5f * (PSGridGetPeriodic(g, x, y - 1, z) + PSGridGetPeriodic(g, x, y + 1, z));
  if (v < 0.0f) return -v;
  return v;
}

double dsq(const int x, const int y, const int z, PSGrid3DDouble h, PSGrid3DFloat g) {
  double a = PSGridGet(h, x, y, z);
  return a * a + PSGridGet(g, x, y, z);
}

double dprod(const int x, const int y, const int z, PSGrid3DDouble h) {
  return 1.0 + 1e-3 * PSGridGet(h, x, y, z);
}

float row2(const int x, const int y, PSGrid2DFloat a, float s) {
  return s * PSGridGet(a, x, y);
}

float line1(const int x, PSGrid1DFloat b) {
  return PSGridGet(b, x) - PSGridGetPeriodic(b, x + 1);
}

int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 0, 0, 0);
  if (argc != 6) {
    fprintf(stderr, "usage: %s NX NY NZ IN BF16\n", argv[0]);
    return 2;
  }
  int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]);
  int attr = atoi(argv[5]) ? PS_GRID_STORAGE_BF16 : 0;
  size_t n = (size_t)nx * ny * nz;
  float *f = (float *)malloc(sizeof(float) * n);
  double *d = (double *)malloc(sizeof(double) * n);
  FILE *in = fopen(argv[4], "rb");
  if (!f || !d || !in || fread(f, sizeof(float), n, in) != n) {
    fprintf(stderr, "cannot read %s\n", argv[4]);
    return 1;
  }
  fclose(in);
  for (size_t i = 0; i < n; ++i) d[i] = 0.5 * f[i] + 0.25;
  PSGrid3DFloat g = PSGrid3DFloatNew(nx, ny, nz, attr);
  PSGrid3DDouble h = PSGrid3DDoubleNew(nx, ny, nz);
  PSGrid2DFloat a = PSGrid2DFloatNew(nx, ny);
  PSGrid1DFloat b = PSGrid1DFloatNew(nx);
  PSGridCopyin(g, f);
  PSGridCopyin(h, d);
  PSGridCopyin(a, f);
  PSGridCopyin(b, f);

  float fs, fmax, fmin, r2, es, emax, l1;
  double ds, dp;
  PSReduce(&fs, PS_SUM, sqdiff, PSDomain3DNew(0, nx - 1, 0, ny, 0, nz), g, 0.75f);
  PSReduce(&fmax, PS_MAX, curv, PSDomain3DNew(0, nx, 0, ny, 1, nz - 1), g);
  PSReduce(&fmin, PS_MIN, curv, PSDomain3DNew(0, nx, 0, ny, 1, nz - 1), g);
  PSReduce(&ds, PS_SUM, dsq, PSDomain3DNew(0, nx, 0, ny, 0, nz), h, g);
  PSReduce(&dp, PS_PROD, dprod, PSDomain3DNew(0, nx, 0, 2, 0, 2), h);
  PSReduce(&r2, PS_SUM, row2, PSDomain2DNew(0, nx, 0, ny), a, 0.25f);
  PSReduce(&l1, PS_MIN, line1, PSDomain1DNew(0, nx), b);
  PSReduce(&es, PS_SUM, sqdiff, PSDomain3DNew(0, 0, 0, ny, 0, nz), g, 1.0f);
  PSReduce(&emax, PS_MAX, curv, PSDomain3DNew(0, nx, 2, 2, 0, nz), g);
  printf("sqdiff_sum %.9g\n", (double)fs);
  printf("curv_max %.9g\n", (double)fmax);
  printf("curv_min %.9g\n", (double)fmin);
  printf("dsq_sum %.17g\n", ds);
  printf("dprod_prod %.17g\n", dp);
  printf("row2_sum %.9g\n", (double)r2);
  printf("line1_min %.9g\n", (double)l1);
  printf("empty_sum %.9g\n", (double)es);
  printf("empty_max %.9g\n", (double)emax);
  PSGridFree(g);
  PSGridFree(h);
  PSGridFree(a);
  PSGridFree(b);
  free(f);
  free(d);
  PSFinalize();
  return 0;
}
