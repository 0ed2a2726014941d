// Times the drop-in path on the reference's own DSL program: the diffusion
// benchmark's Physis variant (reference
// examples/diffusion-benchmark/diffusion3d_physis.c) translated by
// physisc-hip with TRACE_KERNEL, driven through its exported entry points.
//
//   diffusion_dropin_timing N COUNT --physis-trace
//
// initialises an N^3 field (seeded uniform [0,1)), runs the program once with
// 4 steps (the first run allocates the fused-step buffer and loads the
// kernels), then once with COUNT steps.  Each PSStencilRun of the traced
// build waits for its kernels and prints "Physis: Stencil finished (time: T)"
// with T in milliseconds; the host copies the program makes around it are
// outside that time.  Built by tests/golden/build_corpus.py.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

extern "C" {
void initialize_physis(int argc, char **argv, int nx, int ny, int nz);
void initialize_benchmark_physis(int nx, int ny, int nz);
void finalize_benchmark_physis();
void run_kernel_physis(int count, float *f1_host, int nx, int ny, int nz, float ce, float cw, float cn, float cs,
                       float ct, float cb, float cc);
}

int main(int argc, char **argv) {
  std::vector<int> nums;
  for (int i = 1; i < argc; ++i)
    if (std::strncmp(argv[i], "--", 2) != 0) nums.push_back(std::atoi(argv[i]));
  if (nums.size() != 2 || nums[0] < 1 || nums[1] < 0) {
    std::fprintf(stderr, "usage: %s N COUNT [--physis-trace]\n", argv[0]);
    return 2;
  }
  const int n = nums[0], count = nums[1];
  initialize_physis(argc, argv, n, n, n);
  std::vector<float> f((size_t)n * n * n);
  uint64_t s = 12345;
  for (float &v : f) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    v = (float)(s >> 40) * (1.0f / 16777216.0f);
  }
  initialize_benchmark_physis(n, n, n);
  run_kernel_physis(4, f.data(), n, n, n, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.4f);
  std::fprintf(stderr, "timed run: %d^3, %d steps\n", n, count);
  run_kernel_physis(count, f.data(), n, n, n, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.1f, 0.4f);
  finalize_benchmark_physis();
  return 0;
}
