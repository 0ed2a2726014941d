// buffer_hip.hip -- device-side conversion between FP32 host data and the
// 16-bit HBM storage of BF16/FP16 grids (used by copyin/copyout/set).
#include "runtime_hip.h"
#include "physis/hip/storage.h"

namespace physis_hip {
namespace {


template <class S>
__global__ void __launch_bounds__(256) to_storage(uint16_t *dst, const float *src, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = S::narrow(src[i]);
}

template <class S>
__global__ void __launch_bounds__(256) from_storage(float *dst, const uint16_t *src, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = S::widen(src[i]);
}

__device__ __forceinline__ float UnitF32(uint64_t seed, int64_t i) {
  uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return (float)(z >> 40) * (1.0f / 16777216.0f);
}

// Seeded uniform [lo, hi) field: element i = lo + (hi - lo) * u_i with
// u_i = (splitmix64(seed + (i+1) * 0x9E3779B97F4A7C15) >> 40) * 2^-24.
template <class S>
__global__ void __launch_bounds__(256)
fill_uniform(typename S::stored_t *dst, int64_t n, uint64_t seed, float lo, float hi) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const float span = hi - lo;
  for (; i < n; i += stride) S::store_cached(dst + i, (typename S::value_t)(lo + span * UnitF32(seed, i)));
}

unsigned BlocksFor(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (unsigned)(b < 8192 ? (b < 1 ? 1 : b) : 8192);
}

}  // namespace

void ConvertToStorage(void *dst, const float *src, int64_t n, int storage, hipStream_t s) {
  if (n <= 0) return;
  uint16_t *d = static_cast<uint16_t *>(dst);
  if (storage == PS_HIP_STORAGE_BF16)
    hipLaunchKernelGGL(to_storage<psk::BF16>, dim3(BlocksFor(n)), dim3(256), 0, s, d, src, n);
  else if (storage == PS_HIP_STORAGE_FP16)
    hipLaunchKernelGGL(to_storage<psk::FP16>, dim3(BlocksFor(n)), dim3(256), 0, s, d, src, n);
  else
    Fatal("ConvertToStorage", "not a 16-bit storage");
  CheckHip(hipGetLastError(), "to_storage launch");
}

void ConvertFromStorage(float *dst, const void *src, int64_t n, int storage, hipStream_t s) {
  if (n <= 0) return;
  const uint16_t *p = static_cast<const uint16_t *>(src);
  if (storage == PS_HIP_STORAGE_BF16)
    hipLaunchKernelGGL(from_storage<psk::BF16>, dim3(BlocksFor(n)), dim3(256), 0, s, dst, p, n);
  else if (storage == PS_HIP_STORAGE_FP16)
    hipLaunchKernelGGL(from_storage<psk::FP16>, dim3(BlocksFor(n)), dim3(256), 0, s, dst, p, n);
  else
    Fatal("ConvertFromStorage", "not a 16-bit storage");
  CheckHip(hipGetLastError(), "from_storage launch");
}

}  // namespace physis_hip

using namespace physis_hip;

extern "C" int PSHipGridFillUniform(__PSGrid *g, uint64_t seed, float lo, float hi) {
  if (!g || !g->p0) return PS_HIP_ERR_INVALID;
  const GridRec *r = Rec(g);
  hipStream_t s = Runtime::Get().stream();
  const int64_t n = g->num_elms;
  const unsigned nb = BlocksFor(n);
  if (r->storage == PS_HIP_STORAGE_BF16)
    hipLaunchKernelGGL(fill_uniform<psk::BF16>, dim3(nb), dim3(256), 0, s, (uint16_t *)g->p0, n, seed, lo, hi);
  else if (r->storage == PS_HIP_STORAGE_FP16)
    hipLaunchKernelGGL(fill_uniform<psk::FP16>, dim3(nb), dim3(256), 0, s, (uint16_t *)g->p0, n, seed, lo, hi);
  else if (g->elm_size == 4)
    hipLaunchKernelGGL(fill_uniform<psk::Native<float>>, dim3(nb), dim3(256), 0, s, (float *)g->p0, n, seed, lo, hi);
  else if (g->elm_size == 8)
    hipLaunchKernelGGL(fill_uniform<psk::Native<double>>, dim3(nb), dim3(256), 0, s, (double *)g->p0, n, seed, lo, hi);
  else
    return PS_HIP_ERR_UNSUPPORTED;
  return hipGetLastError() == hipSuccess ? PS_HIP_OK : PS_HIP_ERR_DEVICE;
}
