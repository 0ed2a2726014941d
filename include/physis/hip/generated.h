// generated.h -- support code included by every file physisc-hip emits.
//
// Grid references passed to kernels, offset helpers with the reference's
// indexing rules (x fastest; periodic wrap (i + n) % n, reference
// include/physis/physis_cuda.h:158-231), storage dispatch, launch error
// checks and the ping-pong runner that fuses two maps per HBM pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <type_traits>

#include "physis/physis_hip.h"
#include "physis/physis_hip_ext.h"
#include "physis/runtime.h"
#include "physis/hip/skeleton.h"
#include "physis/hip/skeleton_star.h"
#include "physis/hip/skeleton_pack.h"
#include "physis/hip/skeleton_point_tb.h"
#include "physis/hip/skeleton_star_tb.h"
#include "physis/hip/skeleton_stride.h"
#include "physis/hip/storage.h"
#include "physis/hip/wave_reduce.h"

namespace psk {

// A grid as a kernel sees it: device pointer and extents (unused axes 1).
template <class S>
struct GridRef {
  using value_t = typename S::value_t;
  typename S::stored_t *p;
  int dim[3];
  __device__ __forceinline__ value_t load(int64_t off) const { return S::load(p + off); }
  __device__ __forceinline__ void store(int64_t off, value_t v) const { S::store(p + off, v); }
  // Element idx of a grid below 4 GiB: the uniform base plus a 32-bit byte
  // offset (one VGPR per access instead of a 64-bit address).
  __device__ __forceinline__ value_t load_u(uint32_t idx) const {
    return S::load(reinterpret_cast<const typename S::stored_t *>(reinterpret_cast<const char *>(p) +
                                                                  idx * (uint32_t)sizeof(typename S::stored_t)));
  }
  __device__ __forceinline__ void store_u(uint32_t idx, value_t v) const {
    S::store(reinterpret_cast<typename S::stored_t *>(reinterpret_cast<char *>(p) +
                                                      idx * (uint32_t)sizeof(typename S::stored_t)),
             v);
  }
};

template <class S>
inline GridRef<S> MakeRef(const __PSGrid *g) {
  GridRef<S> r;
  r.p = static_cast<typename S::stored_t *>(g->p0);
  for (int d = 0; d < 3; ++d) r.dim[d] = d < g->num_dims ? g->dim[d] : 1;
  return r;
}

template <class R>
__device__ __forceinline__ int64_t Off1(const R &, int i) { return i; }
template <class R>
__device__ __forceinline__ int64_t Off2(const R &g, int i, int j) { return i + (int64_t)j * g.dim[0]; }
template <class R>
__device__ __forceinline__ int64_t Off3(const R &g, int i, int j, int k) {
  return i + (int64_t)j * g.dim[0] + (int64_t)k * g.dim[0] * g.dim[1];
}
template <class R>
__device__ __forceinline__ int64_t OffP1(const R &g, int i) { return (i + g.dim[0]) % g.dim[0]; }
template <class R>
__device__ __forceinline__ int64_t OffP2(const R &g, int i, int j) {
  return (i + g.dim[0]) % g.dim[0] + (int64_t)((j + g.dim[1]) % g.dim[1]) * g.dim[0];
}
template <class R>
__device__ __forceinline__ int64_t OffP3(const R &g, int i, int j, int k) {
  return (i + g.dim[0]) % g.dim[0] + (int64_t)((j + g.dim[1]) % g.dim[1]) * g.dim[0] +
         (int64_t)((k + g.dim[2]) % g.dim[2]) * g.dim[0] * g.dim[1];
}

// Accesses of rank-3 grids in a 3-D plan.  A plan instantiated with kU (its
// launch checked UniformExtents) indexes every grid by ref's extents in 32
// bits, so the grids read at one point share one element index, computed
// once; otherwise each grid's own extents give a 64-bit offset (Off3).
template <bool kU, class R, class Ref>
__device__ __forceinline__ typename R::value_t Get3(const R &g, const Ref &ref, int i, int j, int k) {
  if constexpr (kU) {
    const uint32_t nx = (uint32_t)ref.dim[0], nxy = nx * (uint32_t)ref.dim[1];
    return g.load_u((uint32_t)i + (uint32_t)j * nx + (uint32_t)k * nxy);
  } else {
    return g.load(Off3(g, i, j, k));
  }
}
template <bool kU, class R, class Ref>
__device__ __forceinline__ void Put3(const R &g, const Ref &ref, int i, int j, int k, typename R::value_t v) {
  if constexpr (kU) {
    const uint32_t nx = (uint32_t)ref.dim[0], nxy = nx * (uint32_t)ref.dim[1];
    g.store_u((uint32_t)i + (uint32_t)j * nx + (uint32_t)k * nxy, v);
  } else {
    g.store(Off3(g, i, j, k), v);
  }
}

// Whether every grid has ref's extents and fits 32-bit byte offsets.
template <class Ref, class... R>
inline bool UniformExtents(const Ref &ref, const R &...g) {
  auto fits = [&](const auto &x) {
    return x.dim[0] == ref.dim[0] && x.dim[1] == ref.dim[1] && x.dim[2] == ref.dim[2] &&
           (uint64_t)x.dim[0] * (uint64_t)x.dim[1] * (uint64_t)x.dim[2] * sizeof(*x.p) <= 0xffffffffull;
  };
  return fits(ref) && (fits(g) && ...);
}

inline void CheckLaunch(hipError_t e, const char *kernel) {
  if (e != hipSuccess) {
    fprintf(stderr, "physis-hip: launch of %s failed: %s\n", kernel, hipGetErrorString(e));
    PSAbort(1);
  }
}

// Storage shared by the float grids of a map (PSHipStorage); aborts on a mix.
inline int FloatStorage(const __PSGrid *const *grids, int n) {
  int st = PS_HIP_STORAGE_NATIVE;
  bool first = true;
  for (int i = 0; i < n; ++i) {
    int s = PSHipGridStorage(grids[i]);
    if (first) st = s;
    else if (s != st) {
      fprintf(stderr, "physis-hip: float grids of one map must share a storage kind\n");
      PSAbort(1);
    }
    first = false;
  }
  return st;
}

// Launch shape for the one-step LDS-ring skeleton by map rank.
// parity: -1 for every point, 0/1 for the red/black points of a red-black map.
// 3-D maps of large boxes may run two tile rows per thread (LaunchZMarch3D);
// plans with many grids read outside the ring, or whose two-row ring would
// not fit 64 KiB of LDS, keep one row per thread.
// Bytes of an element of a plan's staged grid (4 when it stages none).
template <class P, class = void>
struct PlanElemBytes {
  static constexpr int value = 4;
};
template <class P>
struct PlanElemBytes<P, std::void_t<typename P::Storage>> {
  static constexpr int value = (int)sizeof(typename P::Storage::stored_t);
};

// 3-D plans that read four or more grids straight from HBM at each point
// (the Himeno Jacobi reads twelve): tiles of 256 threads, the widest of
// 256, 128 and 64 points that covers the box's rows with at most 10% of the
// lanes idle (each step then reads longer runs of every grid), and the
// longest z chunk of 4, 2 planes (4-byte elements) or 16, 8, 4, 2 (2-byte)
// that still gives 1024 workgroups.  The same rule as the built-in Himeno
// launch (himeno19.hip AutoTile), which timed it against 35 tiles.
template <class P>
inline hipError_t LaunchZMarchWide(const typename P::Args &a, const Box &box, hipStream_t s, int parity) {
  const int ex = std::max(box.hi[0] - box.lo[0], 1);
  int bx = 64;
  for (int w : {256, 128}) {
    if (10 * ex >= 9 * ((ex + w - 1) / w * w)) {
      bx = w;
      break;
    }
  }
  static const int kChunks4[] = {4, 2}, kChunks2[] = {16, 8, 4, 2};
  const bool narrow = PlanElemBytes<P>::value == 2;
  const int *chunks = narrow ? kChunks2 : kChunks4;
  const int n = narrow ? 4 : 2;
  int zc = chunks[n - 1];
  for (int i = 0; i < n; ++i) {
    const int blocks = bx == 256 ? ZMarchBlocks<256, 1>(box, chunks[i])
                       : bx == 128 ? ZMarchBlocks<128, 2>(box, chunks[i])
                                   : ZMarchBlocks<64, 4>(box, chunks[i]);
    if (blocks >= 1024) {
      zc = chunks[i];
      break;
    }
  }
  if (bx == 256) return LaunchZMarch<P, 256, 1>(a, box, zc, true, s, parity);
  if (bx == 128) return LaunchZMarch<P, 128, 2>(a, box, zc, true, s, parity);
  return LaunchZMarch<P, 64, 4>(a, box, zc, true, s, parity);
}

template <class P, int RANK>
inline hipError_t LaunchRing(const typename P::Args &a, const Box &box, hipStream_t s, int parity = -1) {
  if (RANK == 1) return LaunchZMarch<P, 64, 1>(a, box, 0, true, s, parity);
  if (RANK == 2) return LaunchZMarch<P, 64, 4>(a, box, 0, true, s, parity);
  if constexpr (RANK == 3 && P::kDirectGrids >= 4) {
    return LaunchZMarchWide<P>(a, box, s, parity);
  } else {
    constexpr size_t kRing2 = (size_t)(P::NS > 0 ? P::NS : 1) * (2 * P::RZ + 2) * (64 + 2 * P::RX) *
                              (16 + 2 * P::RY) * sizeof(typename P::SV);
    return LaunchZMarch3D<P, (P::kDirectGrids <= 2 && kRing2 <= 64 * 1024)>(a, box, 64, true, s, parity);
  }
}

// One step of a star plan: the packed-row tile for whole FP32-valued grids
// whose rows do not fill 256-point waves, the vectorised skeleton when rows are
// whole 16-byte vectors, else the stride skeleton or the LDS ring.
// (nx, ny, nz: extents of the staged grid.)
template <class P>
inline hipError_t LaunchStarStep(const typename P::Args &a, const Box &box, int nx, int ny, int nz, hipStream_t s) {
  const bool whole = box.lo[0] == 0 && box.hi[0] == nx && box.lo[1] == 0 && box.hi[1] == ny && box.lo[2] == 0 &&
                     box.hi[2] == nz;
  // Whole-grid steps of FP32-valued grids whose rows do not fill whole
  // 256-point star waves: the packed-row tile.
  if (whole && nx % 256 != 0) {
    const hipError_t e = LaunchPackStepGrid<P>(a, nx, ny, nz, s);
    if (e != hipErrorNotSupported) return e;
  }
  if (StarSupported(nx)) return LaunchZMarchStar<P, 8>(a, box, 128, true, s);
  // Whole-grid steps with rows of up to 256 points: the stride skeleton
  // (several times the LDS ring on small and mid-size grids).
  if (nx <= 256 && whole) return LaunchStrideStepGrid<P>(a, nx, ny, nz, s);
  return LaunchRing<P, 3>(a, box, s);
}

// Whether two steps of a star plan can be fused on a grid: rows of whole
// 16-byte vectors up to 1024 points (vectorised skeleton), or any rows up
// to 512 points (stride skeleton).
inline bool FusedPairSupported(int nx, int ny, int nz) {
  return StarTB2Supported<1>(nx, ny, nz) || StarTB2Supported<2>(nx, ny, nz) || StarTB2Supported<4>(nx, ny, nz) ||
         (nx >= 1 && nx <= 512 && ny >= 1 && nz >= 1);
}

// Two fused steps src -> dst for a star plan.  mid: also store the
// intermediate step through out_ptr(a1).  Returns the tile that ran: 3 the
// packed-row tile (FP32 values stored in 32 or 16 bits, rows of 4 to 512
// points, unless they fill whole 256-point star waves), 1 the vectorised
// tile (rows of whole 16-byte vectors), 2 the stride tile (other rows), 0
// none (the row width is not supported).
template <class P>
inline int LaunchFusedPair(const typename P::Args &a1, const typename P::Args &a2, int nx, int ny, int nz,
                           hipStream_t s, bool mid) {
  hipError_t e = hipErrorNotSupported;
  int tile = 3;
  if (nx % 256 != 0) e = LaunchPackTB2Grid<P>(a1, a2, nx, ny, nz, s, mid);
  if (e == hipErrorNotSupported) {
    e = LaunchZMarchStarTB2Auto<P>(a1, a2, nx, ny, nz, s, mid);
    tile = 1;
  }
  if (e == hipErrorNotSupported) {
    e = LaunchStrideTB2Grid<P>(a1, a2, nx, ny, nz, s, mid);
    tile = 2;
  }
  if (e == hipErrorNotSupported) return 0;
  CheckLaunch(e, "fused pair");
  return tile;
}

// PSStencilRun(map(k, A -> B), map(k, B -> A), iter) for a star plan whose
// two maps differ only in the swapped grids, over the whole grid.  With
// iter >= 2 every pair of steps is one fused pass: the passes run along the
// buffer path of FusedPathDst (A, a third buffer and B), and the last pass
// writes u_{2 iter} to A and its intermediate u_{2 iter - 1} to B, so both
// grids end exactly as with 2*iter separate steps.  iter == 1 runs the two
// steps unfused (a single fused pass would have to write A while reading
// it).  ab: args of map 1 (A -> B); ba: args of map 2 (B -> A).
template <class P>
inline void RunPingPong(typename P::Args ab, typename P::Args ba, int iter, int nx, int ny, int nz,
                        size_t grid_bytes, hipStream_t s) {
  using st = typename P::Storage::stored_t;
  Box box;
  const int ext[3] = {nx, ny, nz};
  for (int d = 0; d < 3; ++d) {
    box.lo[d] = 0;
    box.hi[d] = ext[d];
  }
  st *A = P::out_ptr(ba);
  st *B = P::out_ptr(ab);
  if (iter >= 2 && FusedPairSupported(nx, ny, nz)) {
    st *aux = static_cast<st *>(__PSHipAuxBuffer(grid_bytes));
    if (!aux) {
      fprintf(stderr, "physis-hip: no memory for the fused-step buffer\n");
      PSAbort(1);
    }
    const st *src = A;
    int tile = 0;
    for (int j = 0; j < iter; ++j) {
      st *dst = FusedPathDst(j, iter, A, B, aux);
      typename P::Args a1 = ab, a2 = ba;
      P::set_in(a1, src);
      P::set_out(a2, dst);
      tile = LaunchFusedPair<P>(a1, a2, nx, ny, nz, s, j == iter - 1);
      if (!tile) {
        fprintf(stderr, "physis-hip: fused pair does not fit the grid\n");
        PSAbort(1);
      }
      src = dst;
    }
    if (__ps_trace)
      fprintf(__ps_trace, "Physis: fused pairs: %d passes on the %s tile\n", iter,
              tile == 1 ? "vectorised" : tile == 3 ? "packed" : "stride");
    return;
  }
  for (int i = 0; i < iter; ++i) {
    CheckLaunch(LaunchStarStep<P>(ab, box, nx, ny, nz, s), "step");
    CheckLaunch(LaunchStarStep<P>(ba, box, nx, ny, nz, s), "step");
  }
}

// ---- kernel reductions: PSReduce(&v, op, kernel, dom, args...) -------------
// (docs/programming.md "Reduction" of the reference; its translators abort on
// it, reference_translator.cc:1357-1360.)  RP::value(a, x, y, z) is the
// reduction kernel's body evaluated at a point of the domain.  The domain is
// cut into items of M x 256 consecutive x points of one row (M = 1, 2 or 4,
// the fewest that cover the row's width up to 1024); workgroup b takes items
// b, b + nb, ... in order, 8 / M at a time, each thread folding the values of
// its points in FP64 (exact for MAX and MIN), and writes one partial; the
// runtime then folds the partials in index order.  The split depends only on
// the domain's extents, so a domain reduces to the same bits on every call.
constexpr int kReduceMapThreads = 256;
constexpr int kReduceMapPoints = 8;  // points per thread in flight
constexpr int kReduceMapMaxGroups = 16384;

template <class Op> struct ReduceIdentity;
template <> struct ReduceIdentity<OpSum> { static constexpr double v = 0.0; };
template <> struct ReduceIdentity<OpProd> { static constexpr double v = 1.0; };
template <> struct ReduceIdentity<OpMax> { static constexpr double v = -__builtin_inf(); };
template <> struct ReduceIdentity<OpMin> { static constexpr double v = __builtin_inf(); };

// I: the item index type (32 bits when the domain has fewer than 2^31
// items, which spares 64-bit divisions).
template <class RP, class Op, class I, int M>
__global__ void __launch_bounds__(kReduceMapThreads)
ReduceMapKernel(typename RP::Args a, Box box, int row_items, int64_t items, double *partials) {
  constexpr int U = kReduceMapPoints / M;
  __shared__ double lds[kReduceMapThreads / 64];
  Op op;
  const I w1 = (I)(box.hi[1] - box.lo[1]), ri = (I)row_items, n = (I)items, step = (I)gridDim.x;
  double acc = ReduceIdentity<Op>::v;
  for (I j0 = (I)blockIdx.x; j0 < n; j0 += U * step) {
    // Items past the end, and lanes past a row's end, evaluate a point of
    // the domain (the last item, the row's last point) and drop the value:
    // no branch around the loads, so all the points' loads are in flight
    // together.
    double v[U][M];
    bool ok[U][M];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const I j = j0 + (I)u * step;
      const I jj = j < n ? j : n - 1;
      const I row = jj / ri;
      const int xc = (int)(jj - row * ri);
      const int y = box.lo[1] + (int)(row % w1), z = box.lo[2] + (int)(row / w1);
      const int x0 = box.lo[0] + xc * (M * kReduceMapThreads) + (int)threadIdx.x;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int x = x0 + m * kReduceMapThreads;
        v[u][m] = (double)RP::value(a, min(x, box.hi[0] - 1), y, z);
        ok[u][m] = j < n && x < box.hi[0];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int m = 0; m < M; ++m)
        if (ok[u][m]) acc = op(acc, v[u][m]);
  }
  const double r = block_reduce<kReduceMapThreads>(acc, op, lds, (int)threadIdx.x);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

template <class RP, class Op, int M>
inline void LaunchReduceMap(const typename RP::Args &a, const Box &box, int row_items, int64_t items, int nb,
                            double *partials, hipStream_t s) {
  const dim3 grid(nb), block(kReduceMapThreads);
  if (items < ((int64_t)1 << 31))
    hipLaunchKernelGGL((ReduceMapKernel<RP, Op, uint32_t, M>), grid, block, 0, s, a, box, row_items, items, partials);
  else
    hipLaunchKernelGGL((ReduceMapKernel<RP, Op, int64_t, M>), grid, block, 0, s, a, box, row_items, items, partials);
}

template <class RP, class Op>
inline void LaunchReduceMapOp(int M, const typename RP::Args &a, const Box &box, int row_items, int64_t items,
                              int nb, double *partials, hipStream_t s) {
  if (M == 1) LaunchReduceMap<RP, Op, 1>(a, box, row_items, items, nb, partials, s);
  else if (M == 2) LaunchReduceMap<RP, Op, 2>(a, box, row_items, items, nb, partials, s);
  else LaunchReduceMap<RP, Op, 4>(a, box, row_items, items, nb, partials, s);
}

// Reduces RP's values over the box with op; the FP64 result goes to *out
// (the operator's identity for an empty box).  Waits for the result.
template <class RP>
inline hipError_t RunReduceMap(const typename RP::Args &a, const Box &box, enum PSReduceOp op, double *out,
                               hipStream_t s) {
  if (op < PS_MAX || op > PS_PROD) return hipErrorInvalidValue;
  const int64_t w0 = box.hi[0] - box.lo[0], w1 = box.hi[1] - box.lo[1], w2 = box.hi[2] - box.lo[2];
  if (w0 <= 0 || w1 <= 0 || w2 <= 0) {
    *out = op == PS_SUM ? 0.0 : op == PS_PROD ? 1.0 : op == PS_MAX ? -__builtin_inf() : __builtin_inf();
    return hipSuccess;
  }
  const int M = w0 <= kReduceMapThreads ? 1 : w0 <= 2 * kReduceMapThreads ? 2 : 4;
  const int row_items = (int)((w0 + M * kReduceMapThreads - 1) / (M * kReduceMapThreads));
  const int64_t items = (int64_t)row_items * w1 * w2;
  const int nb = (int)std::min<int64_t>(items, kReduceMapMaxGroups);
  double *partials = __PSHipReduceMapPartials(nb);
  if (!partials) return hipErrorOutOfMemory;
  switch (op) {
    case PS_SUM: LaunchReduceMapOp<RP, OpSum>(M, a, box, row_items, items, nb, partials, s); break;
    case PS_PROD: LaunchReduceMapOp<RP, OpProd>(M, a, box, row_items, items, nb, partials, s); break;
    case PS_MAX: LaunchReduceMapOp<RP, OpMax>(M, a, box, row_items, items, nb, partials, s); break;
    default: LaunchReduceMapOp<RP, OpMin>(M, a, box, row_items, items, nb, partials, s); break;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return __PSHipReduceMapFinish(out, op, partials, nb) == PS_HIP_OK ? hipSuccess : hipErrorUnknown;
}

}  // namespace psk

// PSGrid{1,2,3}D<Type>New(n0[, n1[, n2]][, attr]) as physisc-hip lowers it.
static inline __PSGrid *__PSGridNewHip(int elm_size, int rank, int attr, int d0, int d1, int d2) {
  PSVectorInt d = {d0, d1, d2};
  if (attr & (PS_GRID_STORAGE_BF16 | PS_GRID_STORAGE_FP16)) return __PSGridNewAttr(elm_size, rank, d, attr);
  return __PSGridNew(elm_size, rank, d, NULL);
}
