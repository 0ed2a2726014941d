/*
 * Leapfrog wave equation on two grids: each map reads the grid it emits
 * (the previous time level at the point), so its two ping-pong maps must not
 * be fused into one pass.
 *
 *   wave_pingpong NX NY NZ ITER IN OUT
 *
 * reads 2*NX*NY*NZ floats (x fastest) from IN: u, then v.  Runs
 * PSStencilRun(map(u -> v), map(v -> u), ITER), where map(a -> b) sets
 * b = 2a - b + K * lap(a) with a periodic 7-point Laplacian, and writes u
 * then v to OUT.  Built by tests/golden/build_corpus.py.
 */
#include <stdio.h>
#include <stdlib.h>

#include "physis/physis.h"

#define K 0.1f

static void wave(const int x, const int y, const int z, PSGrid3DFloat a, PSGrid3DFloat b, float k) {
  float c, w, e, n, s, d, t, lap;
  c = PSGridGet(a, x, y, z);
  w = PSGridGetPeriodic(a, x - 1, y, z);
  e = PSGridGetPeriodic(a, x + 1, y, z);
  n = PSGridGetPeriodic(a, x, y - 1, z);
  s = PSGridGetPeriodic(a, x, y + 1, z);
  d = PSGridGetPeriodic(a, x, y, z - 1);
  t = PSGridGetPeriodic(a, x, y, z + 1);
  lap = w + e + n + s + d + t - 6.0f * c;
  PSGridEmit(b, 2.0f * c - PSGridGet(b, x, y, z) + k * lap);
  return;
}

int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 0, 0, 0);
  if (argc != 7) {
    fprintf(stderr, "usage: %s NX NY NZ ITER IN OUT\n", argv[0]);
    return 2;
  }
  int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]);
  int iter = atoi(argv[4]);
  size_t n = (size_t)nx * ny * nz;
  float *buf = (float *)malloc(sizeof(float) * 2 * n);
  FILE *in = fopen(argv[5], "rb");
  if (!buf || !in || fread(buf, sizeof(float), 2 * n, in) != 2 * n) {
    fprintf(stderr, "cannot read %s\n", argv[5]);
    return 1;
  }
  fclose(in);
  PSGrid3DFloat u = PSGrid3DFloatNew(nx, ny, nz);
  PSGrid3DFloat v = PSGrid3DFloatNew(nx, ny, nz);
  PSDomain3D dom = PSDomain3DNew(0, nx, 0, ny, 0, nz);
  PSGridCopyin(u, buf);
  PSGridCopyin(v, buf + n);
  PSStencilRun(PSStencilMap(wave, dom, u, v, K), PSStencilMap(wave, dom, v, u, K), iter);
  FILE *out = fopen(argv[6], "wb");
  if (!out) return 1;
  PSGridCopyout(u, buf);
  PSGridCopyout(v, buf + n);
  fwrite(buf, sizeof(float), 2 * n, out);
  fclose(out);
  free(buf);
  PSGridFree(u);
  PSGridFree(v);
  PSFinalize();
  return 0;
}
