// Times the drop-in path on the reference's Himeno program
// (examples/himeno/himenobmtxpa_physis.c) translated by physisc-hip with
// TRACE_KERNEL and its main renamed, driven through the program's own
// functions: mat_set / mat_set_init (himenobmtxpa_physis.c:298-329) and
// jacobi (himenobmtxpa_physis.c:364-392).
//
//   himeno_dropin_timing NX NY NZ SWEEPS [bf16] --physis-trace
//
// makes the program's fourteen NX x NY x NZ grids (FP32, or BF16 storage
// with "bf16"), initialises them as the program's main does
// (himenobmtxpa_physis.c:120-151), runs jacobi with 4 sweeps (the first run
// loads the kernels and allocates the fused-sweep buffer), then with SWEEPS
// sweeps.  Each traced PSStencilRun waits for its kernels and prints
// "Physis: Stencil finished (time: T)" with T in milliseconds.  Prints the
// FP64 sum of p0 afterwards (in memory order, x fastest).  Built by
// tests/golden/build_corpus.py.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "physis/physis_hip_ext.h"

// The program's functions (C++ linkage: the translated file is C++).
void mat_set(PSGrid3DFloat mat, float val, float *buf);
void mat_set_init(PSGrid3DFloat mat, float *buf);
float jacobi(int nn, PSGrid3DFloat a0, PSGrid3DFloat a1, PSGrid3DFloat a2, PSGrid3DFloat a3, PSGrid3DFloat b0,
             PSGrid3DFloat b1, PSGrid3DFloat b2, PSGrid3DFloat c0, PSGrid3DFloat c1, PSGrid3DFloat c2,
             PSGrid3DFloat p0, PSGrid3DFloat p1, PSGrid3DFloat bnd, PSGrid3DFloat wrk1);

int main(int argc, char **argv) {
  std::vector<int> nums;
  int storage = PS_HIP_STORAGE_NATIVE;
  for (int i = 1; i < argc; ++i) {
    if (std::strcmp(argv[i], "bf16") == 0)
      storage = PS_HIP_STORAGE_BF16;
    else if (std::strncmp(argv[i], "--", 2) != 0)
      nums.push_back(std::atoi(argv[i]));
  }
  if (nums.size() != 4 || nums[0] < 3 || nums[1] < 3 || nums[2] < 3 || nums[3] < 2 || nums[3] % 2) {
    std::fprintf(stderr, "usage: %s NX NY NZ SWEEPS(even) [bf16] [--physis-trace]\n", argv[0]);
    return 2;
  }
  const int dim[3] = {nums[0], nums[1], nums[2]};
  const int sweeps = nums[3];
  PSInit(&argc, &argv, 3, dim[0], dim[1], dim[2]);
  enum { P0, P1, BND, WRK1, A0, A1, A2, A3, B0, B1, B2, C0, C1, C2, NGRIDS };
  __PSGrid *g[NGRIDS];
  for (int i = 0; i < NGRIDS; ++i) {
    if (PSHipGridNew(&g[i], sizeof(float), 3, dim, storage) != 0) {
      std::fprintf(stderr, "grid allocation failed\n");
      return 1;
    }
  }
  std::vector<float> buf((size_t)dim[0] * dim[1] * dim[2]);
  mat_set_init(g[P0], buf.data());
  mat_set_init(g[P1], buf.data());
  mat_set(g[BND], 1.0f, buf.data());
  mat_set(g[A0], 1.0f, buf.data());
  mat_set(g[A1], 1.0f, buf.data());
  mat_set(g[A2], 1.0f, buf.data());
  mat_set(g[A3], 1.0f / 6.0f, buf.data());
  mat_set(g[B0], 0.0f, buf.data());
  mat_set(g[B1], 0.0f, buf.data());
  mat_set(g[B2], 0.0f, buf.data());
  mat_set(g[C0], 1.0f, buf.data());
  mat_set(g[C1], 1.0f, buf.data());
  mat_set(g[C2], 1.0f, buf.data());
  for (int nn : {4, sweeps}) {
    if (nn == sweeps)
      std::fprintf(stderr, "timed run: %dx%dx%d, %d sweeps, %s\n", dim[0], dim[1], dim[2], sweeps,
                   storage == PS_HIP_STORAGE_BF16 ? "bf16" : "fp32");
    jacobi(nn, g[A0], g[A1], g[A2], g[A3], g[B0], g[B1], g[B2], g[C0], g[C1], g[C2], g[P0], g[P1], g[BND], g[WRK1]);
  }
  PSGridCopyout(g[P0], buf.data());
  double sum = 0.0;
  for (float v : buf) sum += v;
  std::fprintf(stderr, "p0 sum: %.17g\n", sum);
  for (int i = 0; i < NGRIDS; ++i) PSGridFree(g[i]);
  PSFinalize();
  return 0;
}
