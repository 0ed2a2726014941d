/*
 * A relaxation in the point form physisc-hip fuses two sweeps per HBM pass
 * for (psk::PointTB): the staged grid read within radius 1, diagonals
 * included, three coefficient grids read at the point itself, one scalar,
 * a ping-pong PSStencilRun over the interior.
 *
 *   point_relax NX NY NZ ITER IN OUT
 *
 * reads p, q, a, b, c (NX*NY*NZ floats each, x fastest) from IN, runs
 * PSStencilRun(map(relax, interior, p -> q), map(relax, interior, q -> p),
 * ITER) and writes p then q to OUT.  Built by tests/golden/build_corpus.py.
 */
#include <stdio.h>
#include <stdlib.h>

#include "physis/physis.h"

static void relax(const int x, const int y, const int z, PSGrid3DFloat p, PSGrid3DFloat q, PSGrid3DFloat a,
                  PSGrid3DFloat b, PSGrid3DFloat c, float w) {
  float s = PSGridGet(a, x, y, z) * (PSGridGet(p, x + 1, y, z) + PSGridGet(p, x - 1, y, z))
            + PSGridGet(b, x, y, z) * (PSGridGet(p, x, y + 1, z - 1) - PSGridGet(p, x, y - 1, z + 1))
            + PSGridGet(c, x, y, z) * PSGridGet(p, x + 1, y + 1, z + 1);
  float v = PSGridGet(p, x, y, z) + w * (s - PSGridGet(p, x, y, z));
  PSGridEmit(q, v);
  return;
}

int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 0, 0, 0);
  if (argc != 7) {
    fprintf(stderr, "usage: %s NX NY NZ ITER IN OUT\n", argv[0]);
    return 2;
  }
  int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]), iter = atoi(argv[4]);
  size_t n = (size_t)nx * ny * nz;
  float *buf = (float *)malloc(sizeof(float) * 5 * n);
  FILE *in = fopen(argv[5], "rb");
  if (!buf || !in || fread(buf, sizeof(float), 5 * n, in) != 5 * n) {
    fprintf(stderr, "cannot read %s\n", argv[5]);
    return 1;
  }
  fclose(in);
  PSGrid3DFloat p = PSGrid3DFloatNew(nx, ny, nz);
  PSGrid3DFloat q = PSGrid3DFloatNew(nx, ny, nz);
  PSGrid3DFloat a = PSGrid3DFloatNew(nx, ny, nz);
  PSGrid3DFloat b = PSGrid3DFloatNew(nx, ny, nz);
  PSGrid3DFloat c = PSGrid3DFloatNew(nx, ny, nz);
  PSGridCopyin(p, buf);
  PSGridCopyin(q, buf + n);
  PSGridCopyin(a, buf + 2 * n);
  PSGridCopyin(b, buf + 3 * n);
  PSGridCopyin(c, buf + 4 * n);
  PSDomain3D dom = PSDomain3DNew(1, nx - 1, 1, ny - 1, 1, nz - 1);
  PSStencilRun(PSStencilMap(relax, dom, p, q, a, b, c, 0.7f), PSStencilMap(relax, dom, q, p, a, b, c, 0.7f),
               iter);
  FILE *out = fopen(argv[6], "wb");
  if (!out) return 1;
  PSGridCopyout(p, buf);
  PSGridCopyout(q, buf + n);
  fwrite(buf, sizeof(float), 2 * n, out);
  fclose(out);
  free(buf);
  PSFinalize();
  return 0;
}
