// skeleton.h -- the gfx950 kernel skeleton that runs physisc-hip stencil bodies.
//
// One workgroup owns a BX x BY tile of the xy plane and marches a chunk of z.
// Grids the body reads at non-zero offsets ("staged" grids) are streamed
// through an LDS ring of 2*RZ+2 planes, each plane a (BX+2RX) x (BY+2RY)
// tile with its halo.  Each z step
//   1. moves the plane fetched during the previous step from registers into
//      the ring slot of plane z+RZ,
//   2. passes one workgroup barrier,
//   3. issues the global loads of plane z+RZ+1 (consumed next step), and
//   4. calls the body for (x, y, z); the body reads staged grids from LDS
//      through Ctx::at(s, dx, dy, dz) and everything else straight from HBM.
// The ring has one spare slot, so one barrier per step is enough: the slot
// written at step z held plane z-RZ-2, last read at step z-2.
//
// Loads along x are contiguous across the 64 lanes of a wave (coalesced).
// Halo coordinates outside the grid are wrapped (periodic plans) or clamped
// into the grid (all other plans: the DSL body guards such reads with its
// own boundary conditionals, and out-of-grid gets are undefined in the
// reference).  Wrapping is exact for any offset, matching (i + n) % n of the
// reference for i >= -n (include/physis/physis_ref.h:69-81).
//
// Workgroups are numbered so that consecutive xy tiles run on the same XCD
// (the hardware deals workgroups to the 8 XCDs round robin); neighbouring
// tiles then share halo lines in that XCD's L2.
//
// A Plan (emitted by physisc-hip, or hand-written for built-in programs)
// provides:
//   using SV;                         value type of staged grids
//   static constexpr int NS;          number of staged grids (may be 0)
//   static constexpr int RX, RY, RZ;  staging radius per axis
//   static constexpr bool kWrap;      periodic halo
//   static constexpr bool kReduce;    point() returns a double that is summed
//                                     (FP64, fixed order) into
//                                     a.partials[workgroup]; workgroup 0 of
//                                     the launch first calls
//                                     Plan::fold_previous<NT>(a, lds, tid),
//                                     which may sum the partials of the
//                                     previous launch (complete by then)
//   struct Args;                      by-value kernel arguments
//   static void stage_dims(const Args&, int&, int&, int&);   staged grid extents
//   static SV stage_load(const Args&, int s, int64_t off);    element of staged grid s
//   template <class C> static void point(const C&, const Args&, int x, int y, int z);
// and optionally, with NS == 1,
//   using Storage;                    storage of the staged grid
//   static const Storage::stored_t *stage_ptr(const Args&);   its elements
// in which case the next plane is held in registers as stored elements and
// widened only when written to LDS (see Stager).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "physis/physis_hip.h"
#include "physis/hip/storage.h"
#include "physis/hip/wave_reduce.h"

namespace psk {

// Calls f(std::integral_constant<int, J>{}) for J = 0 .. N-1 in order.  The
// z-march loops are unrolled with it by the depth of their register queues:
// a plane then stays in the registers it was loaded into (renamed at
// compile time), because a register copy of a load still in flight would
// make the wave wait for that load.
template <class F, int... J>
__device__ __forceinline__ void StaticFor(F &&f, std::integer_sequence<int, J...>) {
  (f(std::integral_constant<int, J>{}), ...);
}

// Half-open box [lo, hi) per axis.
struct Box {
  int lo[3];
  int hi[3];
};

// Box of a domain for a map of `nd` dimensions; unused axes are [0, 1).
inline Box BoxFromDomain(const __PSDomain &d, int nd) {
  Box b;
  for (int i = 0; i < 3; ++i) {
    b.lo[i] = i < nd ? d.local_min[i] : 0;
    b.hi[i] = i < nd ? d.local_max[i] : 1;
  }
  return b;
}

struct Geometry {
  int tiles_x;
  int tiles_y;
  int zchunk;  // output planes per workgroup
  int remap;   // XCD-aware tile numbering
  int parity;  // -1: every point; 0 (red) / 1 (black): points with (x+y+z+parity) even
};

constexpr int kNumXcd = 8;

// Tile id for workgroup b of n so that the tiles each XCD receives are
// contiguous (bijective on [0, n)).
__device__ __forceinline__ int XcdTile(int b, int n) {
  const int xcd = b % kNumXcd, local = b / kNumXcd;
  const int q = n / kNumXcd, r = n % kNumXcd;
  return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// Index fix-up of halo coordinates.
template <bool kWrap>
__device__ __forceinline__ int FixIndex(int i, int n) {
  if (kWrap) {
    int m = i % n;
    return m < 0 ? m + n : m;
  }
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// How ZMarch loads the next plane into registers.  A plan that names its
// staged grid's Storage and pointer is staged raw: registers hold the stored
// elements until the LDS write, where they are widened.  Widening at the
// load instead would make the wave wait for the load right there (a 16-bit
// element is converted by a shift of the loaded register), so the next
// plane's latency would not overlap the current plane's compute.
template <class Plan, class = void>
struct Stager {
  using raw_t = typename Plan::SV;
  __device__ __forceinline__ static raw_t load(const typename Plan::Args &a, int s, int64_t off) {
    return Plan::stage_load(a, s, off);
  }
  __device__ __forceinline__ static typename Plan::SV widen(raw_t v) { return v; }
};
template <class Plan>
struct Stager<Plan, std::void_t<typename Plan::Storage, std::enable_if_t<Plan::NS == 1>,
                                decltype(Plan::stage_ptr(std::declval<const typename Plan::Args &>()))>> {
  using S = typename Plan::Storage;
  using raw_t = typename S::stored_t;
  __device__ __forceinline__ static raw_t load(const typename Plan::Args &a, int, int64_t off) {
    return Plan::stage_ptr(a)[off];
  }
  __device__ __forceinline__ static typename Plan::SV widen(raw_t v) { return (typename Plan::SV)S::widen(v); }
};

// Per-point view of the staged planes handed to Plan::point.
template <class SV, int NSLOT, int PLANE, int TW, int RZ>
struct Ctx {
  const SV *lds;          // [NS][NSLOT][PLANE]
  int slot[2 * RZ + 1];   // ring slot of plane z+dz, dz = -RZ..RZ
  int center;             // LDS index of (x, y) inside a plane

  __device__ __forceinline__ SV at(int s, int dx, int dy, int dz) const {
    return lds[(s * NSLOT + slot[dz + RZ]) * PLANE + center + dy * TW + dx];
  }
};

// R is the number of tile rows per thread: a workgroup of BX x BY threads
// owns a BX x (BY * R) tile, thread (tx, ty) computing rows ty, ty + BY, ...
// The R points are independent, so more loads and LDS reads are in flight
// per barrier, and the y halo is shared by more rows.
template <class Plan, int BX, int BY, int R = 1>
__global__ void __launch_bounds__(BX * BY)
ZMarch(typename Plan::Args a, Box box, Geometry geo) {
  using SV = typename Plan::SV;
  constexpr int RX = Plan::RX, RY = Plan::RY, RZ = Plan::RZ, NS = Plan::NS;
  constexpr int TY = BY * R;  // tile rows
  constexpr int TW = BX + 2 * RX, TH = TY + 2 * RY, PLANE = TW * TH;
  constexpr int NSLOT = 2 * RZ + 2;
  constexpr int NT = BX * BY;
  constexpr int PER = (PLANE + NT - 1) / NT;
  constexpr int NSA = NS > 0 ? NS : 1;
  __shared__ SV lds[NSA * NSLOT * PLANE];

  const int ntiles = geo.tiles_x * geo.tiles_y;
  const int b = geo.remap ? XcdTile(blockIdx.x, ntiles) : (int)blockIdx.x;
  const int tile_x = b % geo.tiles_x, tile_y = b / geo.tiles_x;
  const int x0 = box.lo[0] + tile_x * BX, y0 = box.lo[1] + tile_y * TY;
  const int zb = box.lo[2] + (int)blockIdx.y * geo.zchunk;
  const int ze = min(zb + geo.zchunk, box.hi[2]);
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
  const int x = x0 + tx;

  using Stage = Stager<Plan>;
  typename Stage::raw_t buf[NSA][PER];
  int goff[PER];
  int nx = 1, ny = 1, nz = 1;
  if constexpr (NS > 0) {
    Plan::stage_dims(a, nx, ny, nz);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int e = tid + j * NT;
      const int lx = e % TW, ly = e / TW;
      const int gx = FixIndex<Plan::kWrap>(x0 - RX + lx, nx);
      const int gy = FixIndex<Plan::kWrap>(y0 - RY + ly, ny);
      goff[j] = gx + gy * nx;
    }
  }
  const int64_t plane_elems = (int64_t)nx * ny;

  auto load_plane = [&](int p) {
    if constexpr (NS > 0) {
      const int64_t zo = (int64_t)FixIndex<Plan::kWrap>(p, nz) * plane_elems;
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int j = 0; j < PER; ++j)
          if (PLANE % NT == 0 || tid + j * NT < PLANE) buf[s][j] = Stage::load(a, s, goff[j] + zo);
    }
  };
  auto store_plane = [&](int p) {
    if constexpr (NS > 0) {
      const int sl = (p - zb + RZ) % NSLOT;
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int j = 0; j < PER; ++j) {
          const int e = tid + j * NT;
          if (PLANE % NT == 0 || e < PLANE) lds[(s * NSLOT + sl) * PLANE + e] = Stage::widen(buf[s][j]);
        }
    }
  };

  if constexpr (NS > 0) {
    for (int k = -RZ; k < RZ; ++k) {
      load_plane(zb + k);
      store_plane(zb + k);
    }
    load_plane(zb + RZ);
  }

  if constexpr (Plan::kReduce) {
    __shared__ double fold[NT / 64 > 0 ? NT / 64 : 1];
    if (blockIdx.x == 0 && blockIdx.y == 0) Plan::template fold_previous<NT>(a, fold, tid);
  }

  Ctx<SV, NSLOT, PLANE, TW, RZ> c[R];
  int y[R];
  bool in_box[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    y[r] = y0 + ty + r * BY;
    in_box[r] = x < box.hi[0] && y[r] < box.hi[1];
    c[r].lds = lds;
    c[r].center = (ty + r * BY + RY) * TW + tx + RX;
  }
  double acc = 0.0;
  for (int z = zb; z < ze; ++z) {
    if constexpr (NS > 0) {
      store_plane(z + RZ);
      __syncthreads();
      if (z + 1 < ze) load_plane(z + RZ + 1);
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = -RZ; k <= RZ; ++k) c[r].slot[k + RZ] = (z + k - zb + RZ) % NSLOT;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      // Red-black maps (reference translator/reference_translator.cc:462-477)
      // update only the points of one colour.
      const bool active = in_box[r] && (geo.parity < 0 || ((x + y[r] + z + geo.parity) & 1) == 0);
      if constexpr (Plan::kReduce) {
        if (active) acc += Plan::point(c[r], a, x, y[r], z);
      } else {
        if (active) Plan::point(c[r], a, x, y[r], z);
      }
    }
  }
  if constexpr (Plan::kReduce) {
    __shared__ double red[NT / 64 > 0 ? NT / 64 : 1];
    const double r = block_reduce<NT>(acc, OpSum(), red, tid);
    if (tid == 0) a.partials[blockIdx.x + blockIdx.y * gridDim.x] = r;
  }
}

// Number of workgroups LaunchZMarch uses for a box (partials needed by a
// reducing plan).
template <int BX, int BY, int R = 1>
inline int ZMarchBlocks(const Box &box, int zchunk) {
  const int ex = box.hi[0] - box.lo[0], ey = box.hi[1] - box.lo[1], ez = box.hi[2] - box.lo[2];
  if (ex <= 0 || ey <= 0 || ez <= 0) return 0;
  if (zchunk <= 0 || zchunk > ez) zchunk = ez;
  return ((ex + BX - 1) / BX) * ((ey + BY * R - 1) / (BY * R)) * ((ez + zchunk - 1) / zchunk);
}

// Host-side launch of a plan over a box with BX x BY threads, each
// computing R rows of a BX x (BY * R) tile.
template <class Plan, int BX, int BY, int R = 1>
inline hipError_t LaunchZMarch(const typename Plan::Args &a, const Box &box, int zchunk,
                               bool remap, hipStream_t stream, int parity = -1) {
  const int ex = box.hi[0] - box.lo[0], ey = box.hi[1] - box.lo[1], ez = box.hi[2] - box.lo[2];
  if (ex <= 0 || ey <= 0 || ez <= 0) return hipSuccess;  // empty domain: nothing to run
  Geometry geo;
  geo.tiles_x = (ex + BX - 1) / BX;
  geo.tiles_y = (ey + BY * R - 1) / (BY * R);
  if (zchunk <= 0 || zchunk > ez) zchunk = ez;
  geo.zchunk = zchunk;
  geo.remap = remap ? 1 : 0;
  geo.parity = parity;
  const int nchunks = (ez + zchunk - 1) / zchunk;
  dim3 grid(geo.tiles_x * geo.tiles_y, nchunks, 1);
  dim3 block(BX, BY, 1);
  hipLaunchKernelGGL((ZMarch<Plan, BX, BY, R>), grid, block, 0, stream, a, box, geo);
  return hipGetLastError();
}

// Workgroups at which a launch fills the MI355X (256 CUs, two per CU).
constexpr int kFillWorkgroups = 512;

// One step of a plan over a 3-D box: 64 x 8 threads with two tile rows each
// when the box still yields enough of those workgroups to fill the device
// (measured 35-40% faster than 64 x 4 threads with one row on 512^3), else
// 64 x 4 with one row (small grids are latency-bound: more, shorter
// workgroups win; two rows per thread made 32-256^3 grids up to 1.5x
// slower).  kTwoRowsOk lets the caller rule the wide tile out (plans with
// many directly read grids, or rings too large for LDS).
template <class Plan, bool kTwoRowsOk>
inline hipError_t LaunchZMarch3D(const typename Plan::Args &a, const Box &box, int zchunk, bool remap,
                                 hipStream_t stream, int parity = -1) {
  if constexpr (kTwoRowsOk) {
    if (ZMarchBlocks<64, 8, 2>(box, zchunk) >= kFillWorkgroups)
      return LaunchZMarch<Plan, 64, 8, 2>(a, box, zchunk, remap, stream, parity);
  }
  return LaunchZMarch<Plan, 64, 4>(a, box, zchunk, remap, stream, parity);
}

}  // namespace psk
