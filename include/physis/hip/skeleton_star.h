// skeleton_star.h -- vectorised z-march skeleton for radius-1 "star" stencils.
//
// Used when the stencil-range analysis finds that the body reads one input
// grid only at the seven offsets (0,0,0), (+-1,0,0), (0,+-1,0), (0,0,+-1) and
// emits one value per point (7-point diffusion, Jacobi, the test corpus's
// 7-point kernels).  Each lane owns four consecutive x points (one 16-byte
// load or store per row), a wave owns 256 x points of one row, and a
// workgroup owns BY rows.  Per z step a lane
//   - keeps planes z-1, z, z+1 of its four points in registers and has the
//     load of plane z+2 in flight (register queue, nothing re-read from HBM),
//   - publishes its plane-z row to LDS (double-buffered by step parity, so
//     one barrier per step) and reads rows y-1 and y+1 back,
//   - gets x-1 / x+4 from the neighbouring lanes by lane shuffles; the two
//     wave-edge values come from loads issued one step ahead.
// Rows y0-1 and y0+BY of the workgroup are loaded by its first and last
// wave.  Coordinates outside the grid are wrapped (periodic) or clamped.
//
// Extra Plan members:
//   using Storage;                        storage kind of the in/out grids
//   static const stored_t *stage_ptr(const Args&);  the input grid
//   static stored_t *out_ptr(const Args&);          the output grid
//   template <class C> static float compute(const C&, const Args&, int x, int y, int z);
// Requirements checked by the launcher: nx % 4 == 0 (16-byte rows).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "physis/hip/skeleton.h"
#include "physis/hip/storage.h"

namespace psk {

typedef float v4f_t __attribute__((ext_vector_type(4)));
typedef uint32_t v2u_t __attribute__((ext_vector_type(2)));

// Four consecutive elements from HBM as FP32.
template <class S>
__device__ __forceinline__ void Load4(const typename S::stored_t *p, float (&v)[4]) {
  if constexpr (sizeof(typename S::stored_t) == 4) {
    const v4f_t q = GlobalLoad(reinterpret_cast<const v4f_t *>(p));
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  } else {
    const v2u_t q = GlobalLoad(reinterpret_cast<const v2u_t *>(p));
    v[0] = S::widen((uint16_t)(q.x & 0xffffu));
    v[1] = S::widen((uint16_t)(q.x >> 16));
    v[2] = S::widen((uint16_t)(q.y & 0xffffu));
    v[3] = S::widen((uint16_t)(q.y >> 16));
  }
}

// Four consecutive 16-bit elements as loaded (two packed pairs), and their
// widening: lets a loop carry a prefetched plane in storage form and widen it
// where it is first used, so no wait for the load falls at the loop's end.
template <class S>
__device__ __forceinline__ v2u_t LoadRaw4(const typename S::stored_t *p) {
  static_assert(sizeof(typename S::stored_t) == 2, "16-bit storage");
  return GlobalLoad(reinterpret_cast<const v2u_t *>(p));
}
template <class S>
__device__ __forceinline__ void Widen4(const v2u_t q, float (&v)[4]) {
  v[0] = S::widen_lo(q.x);
  v[1] = S::widen_hi(q.x);
  v[2] = S::widen_lo(q.y);
  v[3] = S::widen_hi(q.y);
}

// Four consecutive elements to HBM (streaming store).
template <class S>
__device__ __forceinline__ void Store4(typename S::stored_t *p, const float (&v)[4]) {
  if constexpr (sizeof(typename S::stored_t) == 4) {
    v4f_t q = {v[0], v[1], v[2], v[3]};
    GlobalStoreNT(reinterpret_cast<v4f_t *>(p), q);
  } else {
    v2u_t q;
    q.x = S::narrow2(v[0], v[1]);
    q.y = S::narrow2(v[2], v[3]);
    GlobalStoreNT(reinterpret_cast<v2u_t *>(p), q);
  }
}

// A value as it would read back after a store and a load in storage S: the
// on-chip intermediate of fused steps goes through the same rounding as an
// intermediate written to HBM, so fused and unfused runs agree bitwise.
template <class S>
__device__ __forceinline__ float RoundTrip(float v) {
  if constexpr (sizeof(typename S::stored_t) == 4) return v;
  else return S::widen(S::narrow(v));
}

// RoundTrip of N values, converted in pairs (one instruction per pair).
template <class S, int N>
__device__ __forceinline__ void RoundTripN(float (&v)[N]) {
  if constexpr (sizeof(typename S::stored_t) == 2) {
#pragma unroll
    for (int j = 0; j + 1 < N; j += 2) {
      const uint32_t r = S::narrow2(v[j], v[j + 1]);
      v[j] = S::widen_lo(r);
      v[j + 1] = S::widen_hi(r);
    }
    if constexpr (N % 2 == 1) v[N - 1] = RoundTrip<S>(v[N - 1]);
  }
}

// The seven taps of one point; at() is resolved at compile time once the
// body is inlined (its offsets are literals).
struct StarCtx {
  float c, w, e, n, s, b, t;
  __device__ __forceinline__ float at(int, int dx, int dy, int dz) const {
    if (dz < 0) return b;
    if (dz > 0) return t;
    if (dy < 0) return n;
    if (dy > 0) return s;
    if (dx < 0) return w;
    if (dx > 0) return e;
    return c;
  }
};

// Whether Plan provides compute_exact<kX, kY, kZ>: compute() for a point
// whose x (kX), y (kY) and z (kZ) neighbours are the elements at the clamped
// or wrapped coordinates.
template <class Plan, class = void>
struct ExactNeighbours : std::false_type {};
template <class Plan>
struct ExactNeighbours<Plan, std::void_t<decltype(Plan::kExactNeighbours)>>
    : std::integral_constant<bool, Plan::kExactNeighbours> {};

template <class Plan, bool kX, bool kY, bool kZ, class C>
__device__ __forceinline__ float ComputeAt(const C &c, const typename Plan::Args &a, int x, int y, int z) {
  if constexpr (ExactNeighbours<Plan>::value)
    return Plan::template compute_exact<kX, kY, kZ>(c, a, x, y, z);
  else
    return Plan::compute(c, a, x, y, z);
}

struct StarGeometry {
  int x_base;   // first x of tile column 0 (box.lo[0] rounded down to 4)
  int tiles_x;  // 256-wide tiles
  int tiles_y;  // BY-row tiles
  int zchunk;
  int remap;
};

template <class Plan, int BY>
__global__ void __launch_bounds__(64 * BY)
ZMarchStar4(typename Plan::Args a, Box box, StarGeometry geo) {
  using S = typename Plan::Storage;
  using st = typename S::stored_t;
  constexpr bool W = Plan::kWrap;
  __shared__ float rows[2][BY + 2][256];

  const int lane = threadIdx.x, wy = threadIdx.y;
  const int ntiles = geo.tiles_x * geo.tiles_y;
  const int tid = geo.remap ? XcdTile(blockIdx.x, ntiles) : (int)blockIdx.x;
  const int tile_x = tid % geo.tiles_x, tile_y = tid / geo.tiles_x;
  const int xs = geo.x_base + tile_x * 256 + lane * 4;
  const int y0 = box.lo[1] + tile_y * BY;
  const int y = y0 + wy;
  const int zb = box.lo[2] + (int)blockIdx.y * geo.zchunk;
  const int ze = min(zb + geo.zchunk, box.hi[2]);
  int nx, ny, nz;
  Plan::stage_dims(a, nx, ny, nz);
  const st *in = Plan::stage_ptr(a);
  st *out = Plan::out_ptr(a);
  const int64_t plane = (int64_t)nx * ny;
  const bool has_data = xs < nx;  // nx % 4 == 0: then xs + 3 < nx as well
  const int64_t row = (int64_t)FixIndex<W>(y, ny) * nx;
  const bool halo_lo = wy == 0, halo_hi = wy == BY - 1;
  const int64_t hrow_lo = (int64_t)FixIndex<W>(y0 - 1, ny) * nx;
  const int64_t hrow_hi = (int64_t)FixIndex<W>(y0 + BY, ny) * nx;
  const bool need_w = lane == 0 && has_data;
  const bool need_e = has_data && (lane == 63 || xs + 4 >= nx);
  const int xw = FixIndex<W>(xs - 1, nx), xe = FixIndex<W>(xs + 4, nx);
  const bool row_active = y < box.hi[1];

  auto zoff = [&](int z) { return (int64_t)FixIndex<W>(z, nz) * plane; };
  auto ld = [&](int64_t off, float (&v)[4]) {
    if (has_data) {
      Load4<S>(in + off, v);
    } else {
      v[0] = v[1] = v[2] = v[3] = 0.f;
    }
  };

  float bq[4], cq[4], tq[4], pq[4] = {}, hlo[4] = {}, hhi[4] = {}, nlo[4] = {}, nhi[4] = {};
  ld(zoff(zb - 1) + row + xs, bq);
  ld(zoff(zb) + row + xs, cq);
  ld(zoff(zb + 1) + row + xs, tq);
  if (halo_lo) ld(zoff(zb) + hrow_lo + xs, hlo);
  if (halo_hi) ld(zoff(zb) + hrow_hi + xs, hhi);
  float wv = need_w ? S::load(in + zoff(zb) + row + xw) : 0.f;
  float ev = need_e ? S::load(in + zoff(zb) + row + xe) : 0.f;

  for (int z = zb; z < ze; ++z) {
    const bool more = z + 1 < ze;
    float wn = 0.f, en = 0.f;
    if (more) {
      const int64_t z1 = zoff(z + 1);
      ld(zoff(z + 2) + row + xs, pq);
      if (halo_lo) ld(z1 + hrow_lo + xs, nlo);
      if (halo_hi) ld(z1 + hrow_hi + xs, nhi);
      if (need_w) wn = S::load(in + z1 + row + xw);
      if (need_e) en = S::load(in + z1 + row + xe);
    }
    float (*R)[256] = rows[z & 1];
    *reinterpret_cast<v4f_t *>(&R[wy + 1][lane * 4]) = v4f_t{cq[0], cq[1], cq[2], cq[3]};
    if (halo_lo) *reinterpret_cast<v4f_t *>(&R[0][lane * 4]) = v4f_t{hlo[0], hlo[1], hlo[2], hlo[3]};
    if (halo_hi) *reinterpret_cast<v4f_t *>(&R[BY + 1][lane * 4]) = v4f_t{hhi[0], hhi[1], hhi[2], hhi[3]};
    __syncthreads();
    const v4f_t nn = *reinterpret_cast<const v4f_t *>(&R[wy][lane * 4]);
    const v4f_t ss = *reinterpret_cast<const v4f_t *>(&R[wy + 2][lane * 4]);
    float wl = __shfl_up(cq[3], 1, 64);
    float er = __shfl_down(cq[0], 1, 64);
    if (lane == 0) wl = wv;
    if (need_e) er = ev;
    if (row_active && has_data) {
      const float nv[4] = {nn.x, nn.y, nn.z, nn.w};
      const float sv[4] = {ss.x, ss.y, ss.z, ss.w};
      float v[4];
      bool all = xs >= box.lo[0] && xs + 3 < box.hi[0];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        StarCtx c;
        c.c = cq[k];
        c.w = k > 0 ? cq[k - 1] : wl;
        c.e = k < 3 ? cq[k + 1] : er;
        c.n = nv[k];
        c.s = sv[k];
        c.b = bq[k];
        c.t = tq[k];
        const int x = xs + k;
        v[k] = (x >= box.lo[0] && x < box.hi[0]) ? Plan::compute(c, a, x, y, z) : 0.f;
      }
      st *o = out + (int64_t)z * plane + (int64_t)y * nx + xs;
      if (all) {
        Store4<S>(o, v);
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (xs + k >= box.lo[0] && xs + k < box.hi[0]) S::store(o + k, v[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      bq[k] = cq[k];
      cq[k] = tq[k];
      tq[k] = pq[k];
      hlo[k] = nlo[k];
      hhi[k] = nhi[k];
    }
    wv = wn;
    ev = en;
  }
}

// Whether the star skeleton can run a grid of extent nx.
inline bool StarSupported(int nx) { return nx % 4 == 0; }

template <class Plan, int BY>
inline hipError_t LaunchZMarchStar(const typename Plan::Args &a, const Box &box, int zchunk, bool remap,
                                   hipStream_t stream) {
  const int ex = box.hi[0] - box.lo[0], ey = box.hi[1] - box.lo[1], ez = box.hi[2] - box.lo[2];
  if (ex <= 0 || ey <= 0 || ez <= 0) return hipSuccess;
  StarGeometry geo;
  geo.x_base = box.lo[0] & ~3;
  geo.tiles_x = (box.hi[0] - geo.x_base + 255) / 256;
  geo.tiles_y = (ey + BY - 1) / BY;
  if (zchunk <= 0 || zchunk > ez) zchunk = ez;
  geo.zchunk = zchunk;
  geo.remap = remap ? 1 : 0;
  dim3 grid(geo.tiles_x * geo.tiles_y, (ez + zchunk - 1) / zchunk, 1);
  hipLaunchKernelGGL((ZMarchStar4<Plan, BY>), grid, dim3(64, BY, 1), 0, stream, a, box, geo);
  return hipGetLastError();
}

}  // namespace psk
