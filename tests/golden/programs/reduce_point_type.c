/*
 * A kernel reduction over a grid of a user-defined point type (SoA planes
 * on the device): PSReduce(&v, op, kernel, dom, g) with the kernel reading
 * members of g at the point and at x + 1.
 *
 *   reduce_point_type N
 *
 * fills an N^3 grid of struct Point with x = 0.01 * i, y = 1 - 0.001 * i,
 * z = 0.5 (i the linear index, x fastest) and prints the SUM of
 * x(p)^2 + y(p + e_x) * z(p) over [0, N-1) x [0, N) x [0, N) and the MAX of
 * x(p) - y(p) over the whole grid.  Built by tests/golden/build_corpus.py.
 */
#include <stdio.h>
#include <stdlib.h>

#include "physis/physis.h"

struct Point {
  float x;
  float y;
  float z;
};

DeclareGrid3D(Point, struct Point);

float norm2(const int x, const int y, const int z, PSGrid3DPoint g) {
  float a = PSGridGet(g, x, y, z).x, b = PSGridGet(g, x + 1, y, z).y;
  return a * a + b * PSGridGet(g, x, y, z).z;
}

float spread(const int x, const int y, const int z, PSGrid3DPoint g) {
  return PSGridGet(g, x, y, z).x - PSGridGet(g, x, y, z).y;
}

int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 0, 0, 0);
  if (argc != 2) {
    fprintf(stderr, "usage: %s N\n", argv[0]);
    return 2;
  }
  int n = atoi(argv[1]);
  size_t count = (size_t)n * n * n;
  struct Point *buf = (struct Point *)malloc(sizeof(struct Point) * count);
  if (!buf) return 1;
  for (size_t i = 0; i < count; ++i) {
    buf[i].x = 0.01f * (float)i;
    buf[i].y = 1.0f - 0.001f * (float)i;
    buf[i].z = 0.5f;
  }
  PSGrid3DPoint g = PSGrid3DPointNew(n, n, n);
  PSGridCopyin(g, buf);
  float s, m;
  PSReduce(&s, PS_SUM, norm2, PSDomain3DNew(0, n - 1, 0, n, 0, n), g);
  PSReduce(&m, PS_MAX, spread, PSDomain3DNew(0, n, 0, n, 0, n), g);
  printf("norm2_sum %.9g\n", (double)s);
  printf("spread_max %.9g\n", (double)m);
  PSGridFree(g);
  free(buf);
  PSFinalize();
  return 0;
}
