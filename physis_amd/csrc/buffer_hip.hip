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

// Plain streaming copy (16 bytes per lane per iteration): the bandwidth
// yardstick that stencil kernels and HBM counter readings are scaled by.
typedef float v4f __attribute__((ext_vector_type(4)));

// Each thread keeps four 16-byte loads in flight (issued before any store)
// and both directions stream past the caches: the fastest plain copy
// measured on MI355X (about 6.3 TB/s read + write, against 5.4 TB/s for one
// load in flight with cached loads).
constexpr int kCopyUnroll = 4;

__global__ void __launch_bounds__(256)
stream_copy(v4f *__restrict__ dst, const v4f *__restrict__ src, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * 256 * kCopyUnroll;
  for (int64_t base = (int64_t)blockIdx.x * 256 * kCopyUnroll + threadIdx.x; base < n; base += stride) {
    v4f v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const int64_t i = base + u * 256;
      if (i < n) v[u] = __builtin_nontemporal_load(src + i);
    }
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const int64_t i = base + u * 256;
      if (i < n) __builtin_nontemporal_store(v[u], dst + i);
    }
  }
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

extern "C" int PSHipStreamCopy(__PSGrid *dst, const __PSGrid *src) {
  if (!dst || !src || !dst->p0 || !src->p0) return PS_HIP_ERR_INVALID;
  const GridRec *rs = Rec(src);
  if (rs->bytes != Rec(dst)->bytes || rs->bytes % 16 != 0) return PS_HIP_ERR_INVALID;
  const int64_t n = (int64_t)(rs->bytes / 16);
  int64_t b = (n + 256 * kCopyUnroll - 1) / (256 * kCopyUnroll);
  if (b > 256 * 64) b = 256 * 64;
  hipLaunchKernelGGL(stream_copy, dim3((unsigned)b), dim3(256), 0, Runtime::Get().stream(),
                     static_cast<v4f *>(dst->p0), static_cast<const v4f *>(src->p0), n);
  return hipGetLastError() == hipSuccess ? PS_HIP_OK : PS_HIP_ERR_DEVICE;
}

namespace physis_hip {
namespace {

// p(x, y, z) = (float)(z*z) / ((nz-1)*(nz-1)) when kProfile, else `value`;
// the profile is the Himeno initial pressure (reference
// examples/himeno/himenobmtxpa_physis.c:313-329, Physis z = the original's i).
template <class S>
__global__ void __launch_bounds__(256)
fill_pattern(typename S::stored_t *dst, int64_t n, int64_t plane, int nz, int profile, float value) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    float v = value;
    if (profile) {
      const int k = (int)(i / plane);
      v = (float)(k * k) / ((nz - 1) * (nz - 1));
    }
    S::store_cached(dst + i, v);
  }
}

}  // namespace

// Seeded uniform fill on stream s (see PSHipGridFillUniform).
int FillUniformOn(__PSGrid *g, uint64_t seed, float lo, float hi, hipStream_t s) {
  if (!g || !g->p0) return PS_HIP_ERR_INVALID;
  const GridRec *r = Rec(g);
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

// Constant fill (profile = 0) or the Himeno z profile (profile = 1) of a
// float grid on stream s.
int FillPatternOn(__PSGrid *g, int profile, float value, hipStream_t s) {
  if (!g || !g->p0 || g->elm_size != 4) return PS_HIP_ERR_INVALID;
  const GridRec *r = Rec(g);
  const int64_t n = g->num_elms, plane = (int64_t)g->dim[0] * g->dim[1];
  const int nz = g->num_dims == 3 ? g->dim[2] : 1;
  if (profile && nz < 2) return PS_HIP_ERR_INVALID;
  const unsigned nb = BlocksFor(n);
  if (r->storage == PS_HIP_STORAGE_BF16)
    hipLaunchKernelGGL(fill_pattern<psk::BF16>, dim3(nb), dim3(256), 0, s, (uint16_t *)g->p0, n, plane, nz, profile,
                       value);
  else if (r->storage == PS_HIP_STORAGE_FP16)
    hipLaunchKernelGGL(fill_pattern<psk::FP16>, dim3(nb), dim3(256), 0, s, (uint16_t *)g->p0, n, plane, nz, profile,
                       value);
  else
    hipLaunchKernelGGL(fill_pattern<psk::Native<float>>, dim3(nb), dim3(256), 0, s, (float *)g->p0, n, plane, nz,
                       profile, value);
  return hipGetLastError() == hipSuccess ? PS_HIP_OK : PS_HIP_ERR_DEVICE;
}

}  // namespace physis_hip

extern "C" int PSHipGridFillUniform(__PSGrid *g, uint64_t seed, float lo, float hi) {
  return FillUniformOn(g, seed, lo, hi, Runtime::Get().stream());
}

extern "C" int PSHipGridFillPattern(__PSGrid *g, int profile, float value) {
  return FillPatternOn(g, profile, value, Runtime::Get().stream());
}

