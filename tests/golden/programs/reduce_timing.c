/*
 * Times a kernel reduction: PSReduce(&v, PS_SUM, sqdiff, dom, g) with
 * sqdiff = (g(x+1) - g(x))^2 over [0, N-1) x [0, N) x [0, N) of an N^3
 * float grid (one read of the grid per reduction).
 *
 *   reduce_timing N REPS
 *
 * fills g with a seeded pattern, runs the reduction twice untimed, then
 * REPS times, and prints the mean milliseconds per reduction (each call
 * returns its result to the host) and the last value.  Built by
 * tests/golden/build_corpus.py.
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "physis/physis.h"

float sqdiff(const int x, const int y, const int z, PSGrid3DFloat g) {
  float d = PSGridGet(g, x + 1, y, z) - PSGridGet(g, x, y, z);
  return d * d;
}

static double now_ms(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 0, 0, 0);
  if (argc != 3) {
    fprintf(stderr, "usage: %s N REPS\n", argv[0]);
    return 2;
  }
  int n = atoi(argv[1]), reps = atoi(argv[2]);
  size_t count = (size_t)n * n * n;
  float *f = (float *)malloc(sizeof(float) * count);
  if (!f) return 1;
  unsigned s = 12345u;
  for (size_t i = 0; i < count; ++i) {
    s = s * 1664525u + 1013904223u;
    f[i] = (float)(s >> 8) * (1.0f / 16777216.0f);
  }
  PSGrid3DFloat g = PSGrid3DFloatNew(n, n, n);
  PSGridCopyin(g, f);
  PSDomain3D dom = PSDomain3DNew(0, n - 1, 0, n, 0, n);
  float v = 0.0f;
  for (int i = 0; i < 2; ++i) PSReduce(&v, PS_SUM, sqdiff, dom, g);
  double t0 = now_ms();
  for (int i = 0; i < reps; ++i) PSReduce(&v, PS_SUM, sqdiff, dom, g);
  double ms = (now_ms() - t0) / reps;
  printf("reduce_ms %.4f\n", ms);
  printf("value %.9g\n", (double)v);
  PSGridFree(g);
  free(f);
  PSFinalize();
  return 0;
}
