/*
 * Star-form 7-point diffusion for rows of any length, clamped (the
 * clamped-face form of the reference's diffusion benchmark) or periodic.
 *
 *   star_ragged NX NY NZ ITER PERIODIC IN OUT
 *
 * reads NX*NY*NZ floats (x fastest) from IN into g1, runs
 * PSStencilRun(map(g1 -> g2), map(g2 -> g1), ITER) and writes g1 then g2
 * to OUT.  Coefficients (ce, cw, cn, cs, ct, cb, cc) are asymmetric so that
 * a mixed-up direction shows.  Built by tests/golden/build_corpus.py.
 */
#include <stdio.h>
#include <stdlib.h>

#include "physis/physis.h"

#define CE 0.11f
#define CW 0.07f
#define CN 0.05f
#define CS 0.13f
#define CT 0.02f
#define CB 0.17f
#define CC 0.45f

static void clamped(const int x, const int y, const int z, PSGrid3DFloat g1, PSGrid3DFloat g2,
                    float ce, float cw, float cn, float cs, float ct, float cb, float cc) {
  int nx, ny, nz;
  nx = PSGridDim(g1, 0);
  ny = PSGridDim(g1, 1);
  nz = PSGridDim(g1, 2);
  float c, w, e, n, s, b, t;
  c = PSGridGet(g1, x, y, z);
  if (x == 0) w = PSGridGet(g1, x, y, z); else w = PSGridGet(g1, x - 1, y, z);
  if (x == nx - 1) e = PSGridGet(g1, x, y, z); else e = PSGridGet(g1, x + 1, y, z);
  if (y == 0) n = PSGridGet(g1, x, y, z); else n = PSGridGet(g1, x, y - 1, z);
  if (y == ny - 1) s = PSGridGet(g1, x, y, z); else s = PSGridGet(g1, x, y + 1, z);
  if (z == 0) b = PSGridGet(g1, x, y, z); else b = PSGridGet(g1, x, y, z - 1);
  if (z == nz - 1) t = PSGridGet(g1, x, y, z); else t = PSGridGet(g1, x, y, z + 1);
  PSGridEmit(g2, cc * c + cw * w + ce * e + cs * s + cn * n + cb * b + ct * t);
  return;
}

static void periodic(const int x, const int y, const int z, PSGrid3DFloat g1, PSGrid3DFloat g2,
                     float ce, float cw, float cn, float cs, float ct, float cb, float cc) {
  float c, w, e, n, s, b, t;
  c = PSGridGet(g1, x, y, z);
  w = PSGridGetPeriodic(g1, x - 1, y, z);
  e = PSGridGetPeriodic(g1, x + 1, y, z);
  n = PSGridGetPeriodic(g1, x, y - 1, z);
  s = PSGridGetPeriodic(g1, x, y + 1, z);
  b = PSGridGetPeriodic(g1, x, y, z - 1);
  t = PSGridGetPeriodic(g1, x, y, z + 1);
  PSGridEmit(g2, cc * c + cw * w + ce * e + cs * s + cn * n + cb * b + ct * t);
  return;
}

int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 0, 0, 0);
  if (argc != 8) {
    fprintf(stderr, "usage: %s NX NY NZ ITER PERIODIC IN OUT\n", argv[0]);
    return 2;
  }
  int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]);
  int iter = atoi(argv[4]), per = atoi(argv[5]);
  size_t n = (size_t)nx * ny * nz;
  float *buf = (float *)malloc(sizeof(float) * n);
  FILE *in = fopen(argv[6], "rb");
  if (!buf || !in || fread(buf, sizeof(float), n, in) != n) {
    fprintf(stderr, "cannot read %s\n", argv[6]);
    return 1;
  }
  fclose(in);
  PSGrid3DFloat g1 = PSGrid3DFloatNew(nx, ny, nz);
  PSGrid3DFloat g2 = PSGrid3DFloatNew(nx, ny, nz);
  PSDomain3D d = PSDomain3DNew(0, nx, 0, ny, 0, nz);
  PSGridCopyin(g1, buf);
  if (per) {
    PSStencilRun(PSStencilMap(periodic, d, g1, g2, CE, CW, CN, CS, CT, CB, CC),
                 PSStencilMap(periodic, d, g2, g1, CE, CW, CN, CS, CT, CB, CC), iter);
  } else {
    PSStencilRun(PSStencilMap(clamped, d, g1, g2, CE, CW, CN, CS, CT, CB, CC),
                 PSStencilMap(clamped, d, g2, g1, CE, CW, CN, CS, CT, CB, CC), iter);
  }
  FILE *out = fopen(argv[7], "wb");
  if (!out) return 1;
  PSGridCopyout(g1, buf);
  fwrite(buf, sizeof(float), n, out);
  PSGridCopyout(g2, buf);
  fwrite(buf, sizeof(float), n, out);
  fclose(out);
  free(buf);
  PSGridFree(g1);
  PSGridFree(g2);
  PSFinalize();
  return 0;
}
