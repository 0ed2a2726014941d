/*
 * Driver for the Himeno benchmark's original jacobi() (reference
 * examples/himeno/himenobmtxpa_original.c:299-346), which oracle/Makefile
 * compiles from the reference sources with its main renamed and links with
 * this file into oracle/_ref/himeno_original_jacobi.exe.  It lets the tests
 * run the original solver on inputs of their own:
 *
 *   himeno_original_jacobi NX NY NZ SWEEPS IN OUT
 *
 * IN holds 13 arrays of NZ*NY*NX floats, each with x fastest: p, a0, a1, a2,
 * a3, b0, b1, b2, c0, c1, c2, bnd, wrk1.  The original indexes its matrices
 * as (i, j, k) with k fastest, so its i is z, j is y and k is x here.  OUT
 * receives p after SWEEPS sweeps, then the float gosa of the last sweep.
 */
#include <stdio.h>
#include <stdlib.h>

/* The original's matrix type (himenobmtxpa_original.c:45-51). */
struct Mat {
  float *m;
  int mnums;
  int mrows;
  int mcols;
  int mdeps;
};
typedef struct Mat Matrix;

float jacobi(int nn, Matrix *a, Matrix *b, Matrix *c, Matrix *p, Matrix *bnd, Matrix *wrk1, Matrix *wrk2);
int newMat(Matrix *m, int mnums, int mrows, int mcols, int mdeps);

static int load(FILE *f, Matrix *m, int comp, size_t n) {
  return fread(m->m + comp * n, sizeof(float), n, f) == n;
}

int main(int argc, char **argv) {
  if (argc != 7) {
    fprintf(stderr, "usage: %s NX NY NZ SWEEPS IN OUT\n", argv[0]);
    return 2;
  }
  const int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]), sweeps = atoi(argv[4]);
  if (nx < 3 || ny < 3 || nz < 3 || sweeps < 1) return 2;
  const size_t n = (size_t)nx * ny * nz;
  Matrix a, b, c, p, bnd, wrk1, wrk2;
  if (!newMat(&p, 1, nz, ny, nx) || !newMat(&bnd, 1, nz, ny, nx) || !newMat(&wrk1, 1, nz, ny, nx) ||
      !newMat(&wrk2, 1, nz, ny, nx) || !newMat(&a, 4, nz, ny, nx) || !newMat(&b, 3, nz, ny, nx) ||
      !newMat(&c, 3, nz, ny, nx))
    return 1;
  FILE *in = fopen(argv[5], "rb");
  if (!in) return 1;
  int ok = load(in, &p, 0, n);
  for (int i = 0; i < 4; ++i) ok = ok && load(in, &a, i, n);
  for (int i = 0; i < 3; ++i) ok = ok && load(in, &b, i, n);
  for (int i = 0; i < 3; ++i) ok = ok && load(in, &c, i, n);
  ok = ok && load(in, &bnd, 0, n) && load(in, &wrk1, 0, n);
  fclose(in);
  if (!ok) {
    fprintf(stderr, "short input %s\n", argv[5]);
    return 1;
  }
  for (size_t i = 0; i < n; ++i) wrk2.m[i] = 0.0f;
  float gosa = jacobi(sweeps, &a, &b, &c, &p, &bnd, &wrk1, &wrk2);
  FILE *out = fopen(argv[6], "wb");
  if (!out) return 1;
  fwrite(p.m, sizeof(float), n, out);
  fwrite(&gosa, sizeof(float), 1, out);
  fclose(out);
  return 0;
}
