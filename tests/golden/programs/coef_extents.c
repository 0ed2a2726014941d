/*
 * A map reading several grids at (or next to) each point besides the staged
 * one, for the translator's two instantiations of a 3-D plan: with MODE 0
 * every 3-D grid has the extents of p and the launch takes the uniform-
 * extents plan (one 32-bit element index shared by all grids); with MODE 1
 * grid c has one more row (so p's extents would address it wrongly) and the
 * launch keeps per-grid 64-bit offsets.
 *
 *   coef_extents NX NY NZ MODE ITER IN OUT
 *
 * reads p, a, b (NX*NY*NZ floats each, x fastest), then c (NX*(NY+MODE)*NZ
 * floats) from IN; runs PSStencilRun(map(relax, interior, p -> q),
 * map(relax, interior, q -> p), ITER) and writes p then q to OUT.  Built by
 * tests/golden/build_corpus.py.
 */
#include <stdio.h>
#include <stdlib.h>

#include "physis/physis.h"

static void relax(const int x, const int y, const int z, PSGrid3DFloat p, PSGrid3DFloat q, PSGrid3DFloat a,
                  PSGrid3DFloat b, PSGrid3DFloat c, float w) {
  float v = PSGridGet(a, x, y, z) * PSGridGet(p, x + 1, y, z) + PSGridGet(b, x, y, z) * PSGridGet(p, x, y - 1, z)
            + PSGridGet(c, x, y, z + 1) * PSGridGet(p, x, y, z) + w * PSGridGetPeriodic(a, x - 1, y, z);
  PSGridEmit(q, v);
  return;
}

int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 0, 0, 0);
  if (argc != 8) {
    fprintf(stderr, "usage: %s NX NY NZ MODE ITER IN OUT\n", argv[0]);
    return 2;
  }
  int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]), mode = atoi(argv[4]), iter = atoi(argv[5]);
  size_t n = (size_t)nx * ny * nz, nc = (size_t)nx * (ny + mode) * nz;
  float *buf = (float *)malloc(sizeof(float) * (3 * n + nc));
  FILE *in = fopen(argv[6], "rb");
  if (!buf || !in || fread(buf, sizeof(float), 3 * n + nc, in) != 3 * n + nc) {
    fprintf(stderr, "cannot read %s\n", argv[6]);
    return 1;
  }
  fclose(in);
  PSGrid3DFloat p = PSGrid3DFloatNew(nx, ny, nz);
  PSGrid3DFloat q = PSGrid3DFloatNew(nx, ny, nz);
  PSGrid3DFloat a = PSGrid3DFloatNew(nx, ny, nz);
  PSGrid3DFloat b = PSGrid3DFloatNew(nx, ny, nz);
  PSGrid3DFloat c = PSGrid3DFloatNew(nx, ny + mode, nz);
  PSGridCopyin(p, buf);
  PSGridCopyin(q, buf);
  PSGridCopyin(a, buf + n);
  PSGridCopyin(b, buf + 2 * n);
  PSGridCopyin(c, buf + 3 * n);
  PSDomain3D dom = PSDomain3DNew(1, nx - 1, 1, ny - 1, 1, nz - 1);
  PSStencilRun(PSStencilMap(relax, dom, p, q, a, b, c, 0.25f), PSStencilMap(relax, dom, q, p, a, b, c, 0.25f),
               iter);
  FILE *out = fopen(argv[7], "wb");
  if (!out) return 1;
  PSGridCopyout(p, buf);
  PSGridCopyout(q, buf + n);
  fwrite(buf, sizeof(float), 2 * n, out);
  fclose(out);
  free(buf);
  PSFinalize();
  return 0;
}
