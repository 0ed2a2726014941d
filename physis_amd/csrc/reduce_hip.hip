// reduce_hip.hip -- whole-grid reductions (PSReduce) on gfx950.
//
// Replaces the CUDA target's thrust reduce / max_element / min_element
// (reference runtime/reduce_cuda.h:25-45, runtime/reduce_cuda.cu:21-29).
// Two passes with a fixed work split: pass 1 gives each of `nb` workgroups
// a fixed interleaved set of 4-element vectors and writes one FP64 partial
// per group; pass 2 folds the partials in index order.  SUM and PROD
// accumulate in FP64; MAX and MIN are exact in FP64.  The split depends only
// on the element count, so a grid reduces to the same bits on every call.
#include "runtime_hip.h"
#include "physis/hip/storage.h"
#include "physis/hip/wave_reduce.h"

#include <cmath>
#include <limits>

namespace physis_hip {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 2048;

template <class Op> struct Identity;
template <> struct Identity<psk::OpSum> { static constexpr double v = 0.0; };
template <> struct Identity<psk::OpProd> { static constexpr double v = 1.0; };
template <> struct Identity<psk::OpMax> { static constexpr double v = -__builtin_inf(); };
template <> struct Identity<psk::OpMin> { static constexpr double v = __builtin_inf(); };

// Loads elements [4*i, 4*i+4) as doubles.
template <class S>
__device__ __forceinline__ void load4(const typename S::stored_t *p, int64_t i, double (&v)[4]) {
  using st = typename S::stored_t;
  if constexpr (sizeof(st) == 4) {
    float4 q = reinterpret_cast<const float4 *>(p)[i];
    v[0] = S::widen(q.x); v[1] = S::widen(q.y); v[2] = S::widen(q.z); v[3] = S::widen(q.w);
  } else if constexpr (sizeof(st) == 8) {
    const double2 *d = reinterpret_cast<const double2 *>(p);
    double2 a = d[2 * i], b = d[2 * i + 1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
  } else {
    uint2 q = reinterpret_cast<const uint2 *>(p)[i];
    v[0] = S::widen((uint16_t)(q.x & 0xffff)); v[1] = S::widen((uint16_t)(q.x >> 16));
    v[2] = S::widen((uint16_t)(q.y & 0xffff)); v[3] = S::widen((uint16_t)(q.y >> 16));
  }
}

template <class S, class Op>
__global__ void __launch_bounds__(kThreads)
reduce_pass1(const typename S::stored_t *p, int64_t n, double *partial) {
  __shared__ double lds[kThreads / 64];
  Op op;
  double acc = Identity<Op>::v;
  const int64_t nvec = n / 4;
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  // Two vectors in flight per iteration.
  for (; i + stride < nvec; i += 2 * stride) {
    double a[4], b[4];
    load4<S>(p, i, a);
    load4<S>(p, i + stride, b);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc = op(acc, a[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc = op(acc, b[k]);
  }
  if (i < nvec) {
    double a[4];
    load4<S>(p, i, a);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc = op(acc, a[k]);
  }
  // Tail elements (n not a multiple of 4) go to block 0.
  if (blockIdx.x == 0 && threadIdx.x < (n & 3))
    acc = op(acc, (double)S::load(p + nvec * 4 + threadIdx.x));
  double r = psk::block_reduce<kThreads>(acc, op, lds, threadIdx.x);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

template <class Op>
__global__ void __launch_bounds__(kThreads)
reduce_pass2(const double *partial, int nb, double *out) {
  __shared__ double lds[kThreads / 64];
  Op op;
  double acc = Identity<Op>::v;
  for (int i = threadIdx.x; i < nb; i += kThreads) acc = op(acc, partial[i]);
  double r = psk::block_reduce<kThreads>(acc, op, lds, threadIdx.x);
  if (threadIdx.x == 0) *out = r;
}

int BlocksFor(int64_t n) {
  int64_t nvec = (n + 3) / 4;
  int64_t b = (nvec + kThreads * 8 - 1) / (kThreads * 8);
  if (b < 1) b = 1;
  if (b > kMaxBlocks) b = kMaxBlocks;
  return (int)b;
}

template <class S, class Op>
void Launch(const void *p, int64_t n, double *partials, double *out, hipStream_t s) {
  int nb = BlocksFor(n);
  hipLaunchKernelGGL((reduce_pass1<S, Op>), dim3(nb), dim3(kThreads), 0, s,
                     static_cast<const typename S::stored_t *>(p), n, partials);
  hipLaunchKernelGGL((reduce_pass2<Op>), dim3(1), dim3(kThreads), 0, s, partials, nb, out);
}

template <class S>
int Dispatch(const void *p, int64_t n, enum PSReduceOp op, double *partials, double *out, hipStream_t s) {
  switch (op) {
    case PS_SUM: Launch<S, psk::OpSum>(p, n, partials, out, s); break;
    case PS_PROD: Launch<S, psk::OpProd>(p, n, partials, out, s); break;
    case PS_MAX: Launch<S, psk::OpMax>(p, n, partials, out, s); break;
    case PS_MIN: Launch<S, psk::OpMin>(p, n, partials, out, s); break;
    default: return PS_HIP_ERR_INVALID;
  }
  return hipGetLastError() == hipSuccess ? PS_HIP_OK : PS_HIP_ERR_DEVICE;
}

// Queues the reduction with the partials at `partials` (kMaxBlocks doubles)
// and the result written to `out` (device memory, or pinned host memory the
// device can address).
int ReduceTo(const void *p, int64_t n, int elm_size, int storage, enum PSReduceOp op, double *partials,
             double *out, hipStream_t s) {
  if (n <= 0) return PS_HIP_ERR_INVALID;
  if (storage == PS_HIP_STORAGE_BF16) return Dispatch<psk::BF16>(p, n, op, partials, out, s);
  if (storage == PS_HIP_STORAGE_FP16) return Dispatch<psk::FP16>(p, n, op, partials, out, s);
  if (elm_size == 4) return Dispatch<psk::Native<float>>(p, n, op, partials, out, s);
  if (elm_size == 8) return Dispatch<psk::Native<double>>(p, n, op, partials, out, s);
  return PS_HIP_ERR_UNSUPPORTED;
}

}  // namespace

// Queues the reduction of n elements at p (storage kind `storage`, element
// size elm_size) and leaves the FP64 result in device memory at *result_dev
// (partials in result_dev[1 .. ReduceScratchDoubles())).
int ReduceAsync(const void *p, int64_t n, int elm_size, int storage, enum PSReduceOp op,
                double *result_dev, double *partials_dev, hipStream_t s) {
  (void)partials_dev;
  return ReduceTo(p, n, elm_size, storage, op, result_dev + 1, result_dev, s);
}

// Scratch doubles a reduction of any size needs: result + partials.
size_t ReduceScratchDoubles() { return 1 + kMaxBlocks; }

}  // namespace physis_hip

using namespace physis_hip;

extern "C" {

int PSHipReduceGrid(double *out, enum PSReduceOp op, __PSGrid *g) {
  if (!out || !g || !g->p0) return PS_HIP_ERR_INVALID;
  if (op < PS_MAX || op > PS_PROD) return PS_HIP_ERR_INVALID;
  GridRec *r = Rec(g);
  Runtime &rt = Runtime::Get();
  hipStream_t s = rt.stream();
  double *partials = static_cast<double *>(rt.Scratch(ReduceScratchDoubles() * sizeof(double)));
  // The folding kernel stores the result straight into pinned host memory
  // (no copy to queue); it is visible once the stream has synchronised.
  double *host = static_cast<double *>(rt.PinnedScratch(sizeof(double)));
  double *host_dev = nullptr;
  if (hipHostGetDevicePointer(reinterpret_cast<void **>(&host_dev), host, 0) != hipSuccess)
    return PS_HIP_ERR_DEVICE;
  int st = ReduceTo(g->p0, g->num_elms, g->elm_size, r->storage, op, partials, host_dev, s);
  if (st != PS_HIP_OK) return st;
  if (hipStreamSynchronize(s) != hipSuccess) return PS_HIP_ERR_DEVICE;
  *out = *static_cast<volatile double *>(host);
  return PS_HIP_OK;
}

// PSReduce(&v, op, g) on a float grid (reference libphysis_rt_ref.cc:150-153).
void __PSReduceGridFloat(void *buf, enum PSReduceOp op, __PSGrid *g) {
  if (op < PS_MAX || op > PS_PROD) PSAbort(1);  // invalid operator (runtime/reduce.h:49-51)
  if (g->elm_size != 4) Fatal("__PSReduceGridFloat", "grid is not a float grid");
  double v = 0.0;
  if (PSHipReduceGrid(&v, op, g) != PS_HIP_OK) Fatal("__PSReduceGridFloat", "reduction failed");
  *static_cast<float *>(buf) = (float)v;
}

void __PSReduceGridDouble(void *buf, enum PSReduceOp op, __PSGrid *g) {
  if (op < PS_MAX || op > PS_PROD) PSAbort(1);
  if (g->elm_size != 8) Fatal("__PSReduceGridDouble", "grid is not a double grid");
  double v = 0.0;
  if (PSHipReduceGrid(&v, op, g) != PS_HIP_OK) Fatal("__PSReduceGridDouble", "reduction failed");
  *static_cast<double *>(buf) = v;
}

}  // extern "C"
