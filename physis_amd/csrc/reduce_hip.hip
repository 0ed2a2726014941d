// reduce_hip.hip -- whole-grid reductions (PSReduce) on gfx950.
//
// Replaces the CUDA target's thrust reduce / max_element / min_element
// (reference runtime/reduce_cuda.h:25-45, runtime/reduce_cuda.cu:21-29).
// Two passes with a fixed work split: the grid is cut into chunks of
// kThreads x kVec(S) four-element vectors (128 bytes per thread); pass 1
// gives workgroup b the chunks b, b + nb, ... (one chunk each up to
// kMaxBlocks chunks) and writes one FP64 partial per group; pass 2 folds the
// partials in index order.  SUM and PROD accumulate in FP64; MAX and MIN are
// exact in FP64.  The split depends only on the element count, so a grid
// reduces to the same bits on every call.
//
// Each thread issues all its loads of a chunk before using any (128 bytes in
// flight per lane), with non-temporal hints: the data is read once, and on
// MI355X a 512 MiB pass then streams at ~7.1 TB/s instead of ~5.7 TB/s with
// a grid-stride loop of two plain 16-byte loads in flight.
#include "runtime_hip.h"
#include "physis/hip/storage.h"
#include "physis/hip/wave_reduce.h"

#include <cmath>
#include <cstdio>
#include <limits>
#include <mutex>

namespace physis_hip {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 16384;

template <class Op> struct Identity;
template <> struct Identity<psk::OpSum> { static constexpr double v = 0.0; };
template <> struct Identity<psk::OpProd> { static constexpr double v = 1.0; };
template <> struct Identity<psk::OpMax> { static constexpr double v = -__builtin_inf(); };
template <> struct Identity<psk::OpMin> { static constexpr double v = __builtin_inf(); };

// Four stored elements as one vector of the storage's width.
template <class S, size_t kBytes = sizeof(typename S::stored_t)> struct Raw4;
template <class S> struct Raw4<S, 2> {
  typedef unsigned int v_t __attribute__((ext_vector_type(2)));
  v_t q;
  __device__ void load(const typename S::stored_t *p, int64_t i) {
    q = __builtin_nontemporal_load(reinterpret_cast<const v_t *>(p) + i);
  }
  __device__ double get(int k) const {
    const unsigned int w = k < 2 ? q.x : q.y;
    return S::widen((uint16_t)((k & 1) ? (w >> 16) : (w & 0xffff)));
  }
};
template <class S> struct Raw4<S, 4> {
  typedef float v_t __attribute__((ext_vector_type(4)));
  v_t q;
  __device__ void load(const typename S::stored_t *p, int64_t i) {
    q = __builtin_nontemporal_load(reinterpret_cast<const v_t *>(p) + i);
  }
  __device__ double get(int k) const { return S::widen(q[k]); }
};
template <class S> struct Raw4<S, 8> {
  typedef double v_t __attribute__((ext_vector_type(2)));
  v_t a, b;
  __device__ void load(const typename S::stored_t *p, int64_t i) {
    a = __builtin_nontemporal_load(reinterpret_cast<const v_t *>(p) + 2 * i);
    b = __builtin_nontemporal_load(reinterpret_cast<const v_t *>(p) + 2 * i + 1);
  }
  __device__ double get(int k) const { return k < 2 ? a[k] : b[k - 2]; }
};

// Four-element vectors per thread per chunk: 128 bytes of loads in flight.
template <class S>
constexpr int kVec = 128 / (4 * (int)sizeof(typename S::stored_t));
template <class S>
constexpr int64_t kChunkVec = (int64_t)kThreads * kVec<S>;

template <class S>
int64_t ChunksFor(int64_t n) {
  const int64_t nvec = n / 4;
  return (nvec + kChunkVec<S> - 1) / kChunkVec<S>;
}

template <class S, class Op>
__global__ void __launch_bounds__(kThreads)
reduce_pass1(const typename S::stored_t *p, int64_t n, double *partial) {
  __shared__ double lds[kThreads / 64];
  constexpr int K = kVec<S>;
  Op op;
  double acc = Identity<Op>::v;
  const int64_t nvec = n / 4;
  const int64_t nchunks = (nvec + kChunkVec<S> - 1) / kChunkVec<S>;
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int64_t base = c * kChunkVec<S> + threadIdx.x;
    Raw4<S> v[K];
    // Lanes past the last vector load vector 0 (in bounds) and skip it.
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t i = base + (int64_t)k * kThreads;
      v[k].load(p, i < nvec ? i : 0);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (base + (int64_t)k * kThreads < nvec) {
#pragma unroll
        for (int e = 0; e < 4; ++e) acc = op(acc, v[k].get(e));
      }
    }
  }
  // Tail elements (n not a multiple of 4) go to block 0.
  if (blockIdx.x == 0 && threadIdx.x < (n & 3))
    acc = op(acc, (double)S::load(p + nvec * 4 + threadIdx.x));
  double r = psk::block_reduce<kThreads>(acc, op, lds, threadIdx.x);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// Pass 2: one workgroup of kThreads2 threads folds the nb <= kMaxBlocks
// partials; thread t holds partials t, t + kThreads2, ... (all its loads in
// flight at once) and folds them in that order before the workgroup tree.
constexpr int kThreads2 = 1024;
static_assert(kMaxBlocks % kThreads2 == 0, "partials per thread");

template <class Op>
__global__ void __launch_bounds__(kThreads2)
reduce_pass2(const double *partial, int nb, double *out) {
  __shared__ double lds[kThreads2 / 64];
  constexpr int K = kMaxBlocks / kThreads2;
  Op op;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * kThreads2;
    v[k] = i < nb ? partial[i] : Identity<Op>::v;
  }
  double acc = Identity<Op>::v;
#pragma unroll
  for (int k = 0; k < K; ++k) acc = op(acc, v[k]);
  double r = psk::block_reduce<kThreads2>(acc, op, lds, threadIdx.x);
  if (threadIdx.x == 0) *out = r;
}

// One workgroup per chunk up to kMaxBlocks (so the partials fit the
// scratch), and at least one for the tail of a grid smaller than a vector.
template <class S>
int BlocksFor(int64_t n) {
  int64_t b = ChunksFor<S>(n);
  if (b < 1) b = 1;
  if (b > kMaxBlocks) b = kMaxBlocks;
  return (int)b;
}

template <class S, class Op>
void Launch(const void *p, int64_t n, double *partials, double *out, hipStream_t s) {
  int nb = BlocksFor<S>(n);
  hipLaunchKernelGGL((reduce_pass1<S, Op>), dim3(nb), dim3(kThreads), 0, s,
                     static_cast<const typename S::stored_t *>(p), n, partials);
  hipLaunchKernelGGL((reduce_pass2<Op>), dim3(1), dim3(kThreads2), 0, s, partials, nb, out);
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

// ---- the reduction without allocation (PSHipReduceGridNoAlloc) -----------
// Its result and partials live in `scratch`, a file-scope device array of
// 1 + kMaxBlocks doubles (128 KiB + 8 bytes, held by every device the
// library runs on until the process ends).  Host code never uses the
// array's name as an address: NoAllocScratchInit resolves its device
// address once per device, with that device current, and the entry point
// passes that address to the kernels and the copy.
constexpr size_t kNoAllocDoubles = 1 + kMaxBlocks;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
namespace {
constexpr int kMaxDevices = 64;
double *g_scratch_addr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Workgroups (partials) of pass 1 for n elements of this storage.
int ReduceBlocks(int64_t n, int elm_size, int storage) {
  if (storage == PS_HIP_STORAGE_BF16) return BlocksFor<psk::BF16>(n);
  if (storage == PS_HIP_STORAGE_FP16) return BlocksFor<psk::FP16>(n);
  if (elm_size == 8) return BlocksFor<psk::Native<double>>(n);
  return BlocksFor<psk::Native<float>>(n);
}
}  // namespace

int NoAllocScratchInit() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return PS_HIP_ERR_DEVICE;
  if (dev < 0 || dev >= kMaxDevices) return PS_HIP_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_addr[dev]) return PS_HIP_OK;
  void *addr = nullptr;
  if (hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)) != hipSuccess || !addr) {
    (void)hipGetLastError();
    return PS_HIP_ERR_DEVICE;
  }
  g_scratch_addr[dev] = static_cast<double *>(addr);
  return PS_HIP_OK;
}

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

// PSHipReduceGrid without allocating: the partials and the result go to the
// device's `scratch` array; the result is copied to *out.  Same kernels and
// work split as PSHipReduceGrid, so the same bits.  Calls that overlap on one
// device race on the array.
int PSHipReduceGridNoAlloc(double *out, enum PSReduceOp op, __PSGrid *g) {
  if (!out || !g || !g->p0) return PS_HIP_ERR_INVALID;
  if (op < PS_MAX || op > PS_PROD) return PS_HIP_ERR_INVALID;
  GridRec *r = Rec(g);
  Runtime &rt = Runtime::Get();
  hipStream_t s = rt.stream();
  const int dev = rt.device();
  if (dev < 0 || dev >= kMaxDevices) return PS_HIP_ERR_UNSUPPORTED;
  if (!g_scratch_addr[dev]) {
    const int st = NoAllocScratchInit();  // the runtime's device is current
    if (st != PS_HIP_OK) return st;
  }
  if (1 + (size_t)ReduceBlocks(g->num_elms, g->elm_size, r->storage) > kNoAllocDoubles) {
    fprintf(stderr, "physis-hip: PSHipReduceGridNoAlloc: %lld elements need more than its %zu-double scratch\n",
            (long long)g->num_elms, kNoAllocDoubles);
    return PS_HIP_ERR_UNSUPPORTED;
  }
  double *buf = g_scratch_addr[dev];
  int st = ReduceTo(g->p0, g->num_elms, g->elm_size, r->storage, op, buf + 1, buf, s);
  if (st != PS_HIP_OK) return st;
  if (hipMemcpyAsync(out, buf, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) return PS_HIP_ERR_DEVICE;
  return hipStreamSynchronize(s) == hipSuccess ? PS_HIP_OK : PS_HIP_ERR_DEVICE;
}

// Partials of a kernel reduction (PSReduce(&v, op, kernel, dom, ...), lowered
// by physisc-hip: include/physis/hip/generated.h RunReduceMap): n FP64 values
// in the runtime's scratch buffer, used in stream order.
double *__PSHipReduceMapPartials(int n) {
  if (n < 1 || n > kMaxBlocks) return nullptr;
  return static_cast<double *>(Runtime::Get().Scratch((size_t)n * sizeof(double)));
}

// Folds the n partials a kernel reduction queued on the runtime stream, in
// index order (the pass 2 of PSHipReduceGrid), waits for it and stores the
// FP64 result in *out.
int __PSHipReduceMapFinish(double *out, enum PSReduceOp op, const double *partials, int n) {
  if (!out || !partials || n < 1 || n > kMaxBlocks) return PS_HIP_ERR_INVALID;
  Runtime &rt = Runtime::Get();
  hipStream_t s = rt.stream();
  double *host = static_cast<double *>(rt.PinnedScratch(sizeof(double)));
  double *host_dev = nullptr;
  if (hipHostGetDevicePointer(reinterpret_cast<void **>(&host_dev), host, 0) != hipSuccess)
    return PS_HIP_ERR_DEVICE;
  switch (op) {
    case PS_SUM: hipLaunchKernelGGL((reduce_pass2<psk::OpSum>), dim3(1), dim3(kThreads2), 0, s, partials, n, host_dev); break;
    case PS_PROD: hipLaunchKernelGGL((reduce_pass2<psk::OpProd>), dim3(1), dim3(kThreads2), 0, s, partials, n, host_dev); break;
    case PS_MAX: hipLaunchKernelGGL((reduce_pass2<psk::OpMax>), dim3(1), dim3(kThreads2), 0, s, partials, n, host_dev); break;
    case PS_MIN: hipLaunchKernelGGL((reduce_pass2<psk::OpMin>), dim3(1), dim3(kThreads2), 0, s, partials, n, host_dev); break;
    default: return PS_HIP_ERR_INVALID;
  }
  if (hipGetLastError() != hipSuccess) return PS_HIP_ERR_DEVICE;
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
