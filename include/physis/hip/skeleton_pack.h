// skeleton_pack.h -- star-stencil steps (two fused, or one) on rows of any
// length, packed across lanes (FP32 values stored in 32 or 16 bits).
//
// The row-tile skeletons give a wave whole rows of a fixed width: a row of
// nx points occupies 64*NP lane slots (stride tiles) or 256 (star tiles), so
// rows whose length is not a multiple of that width leave lanes idle, and
// the stride tiles move one element per lane and instruction.  Here a row
// takes V = ceil(nx/4) lanes of four consecutive points each (one 16-byte
// access, 8 bytes for 16-bit storage), and a workgroup of T lanes holds
// R = floor(T/V) consecutive rows of one z plane, packed lane after lane:
// lane L holds row L / V, vector vi = L % V, points xs = min(4 vi, nx - 4)
// .. xs + 3.  The last vector of a row is aligned to the row's end, so it
// overlaps its neighbour when nx is not a multiple of four: both lanes
// compute the overlapped points from the same values and write the same
// bits.  Every access stays inside its row; vector accesses need only
// element alignment on this target (the runtime runs in its unaligned access
// mode, which the compiler's own code for unaligned copies relies on), so
// rows need not start on 16 bytes.  Lanes past R*V shadow a lane of the last
// row (same loads, same LDS words, no stores).
//
// LDS holds a plane slot-major: point j of every vector in array j, at the
// vector's position (row-major over rows and vectors).  A lane's y
// neighbours are then the words V positions away in the same four arrays,
// read across the wave as consecutive words (free of bank conflicts, two
// slots per instruction); its x neighbours inside the vector are its own
// registers, and the two across the vector's ends are one word each whose
// array and position are fixed per lane: the neighbouring vector's adjacent
// point, or at a row end the point itself (clamped) or the row's other end
// (periodic).  Halo rows and planes are loaded at wrapped or clamped
// coordinates, so every neighbour a u1 value reads is the element at the
// wrapped/clamped coordinate (compute_exact).
//
// As in skeleton_star_tb.h the workgroup computes u1 = body(u0) on its R
// rows and writes u2 = body(u1) on the inner R-2 rows, marching z with the
// planes of its four points in register queues rotated over an unrolled
// period of three steps, double-buffered LDS (one barrier per step).  u2
// reads u1 neighbours; at a clamped y face it reads its own row instead of
// the neighbour row (positions chosen once per lane) and at a z face it
// keeps the plan's select.  The same FP32 body evaluates both levels, so the
// result is bitwise equal to two separate steps.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "physis/hip/skeleton.h"
#include "physis/hip/skeleton_star.h"
#include "physis/hip/skeleton_stride.h"

namespace psk {

// Rows of a workgroup of 64*nw lanes on rows of nx points (nx >= 4).
__host__ __device__ constexpr int PackRows(int nx, int nw) { return 64 * nw / ((nx + 3) / 4); }

// LDS of one packed tile for rows of up to NXMAX points: four slot arrays
// per plane of u0 (the T lane positions and two halo rows of NXMAX/4
// vectors) and of u1 (the T lane positions), double-buffered by step
// parity.  Array strides are multiples of 64 words (two arrays per
// stride-64 read).
template <int NW, int NXMAX>
struct PackTB2Lds {
  static constexpr int kT = 64 * NW;
  static constexpr int kS0 = (kT + 2 * (NXMAX / 4) + 63) / 64 * 64;
  static constexpr int kS1 = kT;
  float u0[2][4 * kS0];
  float u1[2][4 * kS1];
};

// Whether a packed tile of nw waves serves rows of nx points (at least one
// output row per tile).
__host__ __device__ inline bool PackSupported(int nx, int nw, int nxmax) {
  return nx >= 4 && nx <= nxmax && PackRows(nx, nw) >= 3;
}

// Whether a plane of nx x ny elements of elem bytes fits the tiles' 32-bit
// byte offsets, with room for the out-of-range store offset (kPackNoStore).
__host__ __device__ inline bool PackPlaneFits(int nx, int ny, int elem) {
  return (int64_t)nx * ny * elem < 0x7ffffff0LL;
}

// Storage kinds the packed tiles serve: FP32 values stored in 32 or 16 bits.
template <class S>
constexpr bool kPackStorage = std::is_same<typename S::value_t, float>::value;

// One plane of a grid as a buffer resource of P elements (the same for the
// whole wave: broadcast from the first lane).
template <class st>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t PackPlane(const st *grid, int64_t zoff, int P) {
  const uint64_t a = reinterpret_cast<uint64_t>(grid + zoff);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void *>(((uint64_t)hi << 32) | lo), (short)0,
                                           __builtin_amdgcn_readfirstlane(P * (int)sizeof(st)), kBufferWord3);
}

// Four consecutive points (16 bytes, or 8 for 16-bit storage, at any
// element-aligned offset) as FP32.
template <class S>
__device__ __forceinline__ void PackLoad(__amdgpu_buffer_rsrc_t r, int byte_off, float (&v)[4]) {
  if constexpr (sizeof(typename S::stored_t) == 4) {
    const v4f_t q = __builtin_bit_cast(v4f_t, __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 0));
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  } else {
    Widen4<S>(__builtin_bit_cast(v2u_t, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0)), v);
  }
}

// One point as FP32.
template <class S>
__device__ __forceinline__ float PackLoad1(__amdgpu_buffer_rsrc_t r, int byte_off) {
  if constexpr (sizeof(typename S::stored_t) == 4)
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
  else
    return S::widen(__builtin_amdgcn_raw_buffer_load_b16(r, byte_off, 0, 0));
}

// Four points, non-temporal (16-bit storage: rounded to nearest even in
// pairs); dwords past the resource's range are dropped.
template <class S>
__device__ __forceinline__ void PackStore(__amdgpu_buffer_rsrc_t r, int byte_off, const float (&v)[4]) {
  typedef uint32_t v4u_t __attribute__((ext_vector_type(4)));
  if constexpr (sizeof(typename S::stored_t) == 4) {
    const v4f_t q = {v[0], v[1], v[2], v[3]};
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u_t, q), r, byte_off, 0, 2);
  } else {
    v2u_t q;
    q.x = S::narrow2(v[0], v[1]);
    q.y = S::narrow2(v[2], v[3]);
    __builtin_amdgcn_raw_buffer_store_b64(q, r, byte_off, 0, 2);
  }
}

// Where lane `lane` of a packed workgroup of 64*nw lanes works on rows of
// nx points (V = ceil(nx/4) lanes per row, R = floor(64 nw / V) rows):
// row r (a lane past the R rows shadows lane (R-1, vi)), vector vi, first
// point xs; and the LDS words of the x neighbours across its vector's ends,
// each as (slot array, position offset from the vector's own position):
// west of point xs (wj, wd) and east of point xs + 3 (ej, ed) -- the
// neighbouring vector's adjacent point, or at a row end the point itself
// (clamped) or the row's other end (wrap).
struct PackLane {
  int r, vi, xs, shadow;
  int wj, wd, ej, ed;
};
__host__ __device__ inline PackLane PackLaneOf(int lane, int nx, int nw, bool wrap) {
  const int V = (nx + 3) >> 2, R = 64 * nw / V;
  PackLane p;
  int r = lane / V;
  p.vi = lane - r * V;
  p.shadow = r >= R;
  p.r = r < R ? r : R - 1;
  p.xs = 4 * p.vi < nx - 4 ? 4 * p.vi : nx - 4;
  const bool last = p.vi == V - 1;
  if (p.vi == 0) {
    p.wj = wrap ? 3 : 0;
    p.wd = wrap ? V - 1 : 0;  // wrap: the row's last point; clamped: itself
  } else {
    p.wj = last ? nx - 4 * V + 3 : 3;  // point xs - 1 in vector vi - 1
    p.wd = -1;
  }
  if (last) {
    p.ej = wrap ? 0 : 3;
    p.ed = wrap ? -p.vi : 0;  // wrap: the row's first point; clamped: itself
  } else {
    p.ej = p.vi == V - 2 ? 4 * V - nx : 0;  // point xs + 4 in vector vi + 1
    p.ed = 1;
  }
  return p;
}

// The two LDS places (slot array, position) of point x of a halo row whose
// vectors start at position `base` (rows of nx points, V vectors): its
// vector's slot, and where a row's last vector (aligned to the row's end)
// holds it too; the same place twice if the point is in one vector only.
struct PackHalo {
  int ja, pa, jb, pb;
};
__host__ __device__ inline PackHalo PackHaloOf(int x, int nx, int base) {
  const int V = (nx + 3) >> 2;
  PackHalo h;
  const int lastslot = x - (nx - 4);
  if (lastslot >= 0) {
    h.jb = lastslot;
    h.pb = base + V - 1;
  } else {
    h.jb = x & 3;
    h.pb = base + (x >> 2);
  }
  if ((x >> 2) < V - 1) {
    h.ja = x & 3;
    h.pa = base + (x >> 2);
  } else {
    h.ja = h.jb;
    h.pa = h.pb;
  }
  return h;
}

// A byte offset past any plane's range (planes here hold < 2^31 bytes): a
// lane that writes nothing stores there, and the range check drops it, so
// the store needs no branch.
constexpr int kPackNoStore = 0x7ffffff0;

// One fused tile: u1 on rows [ya, ya+R) and u2 on rows [ya+1, ya+R-1)
// (those inside the grid), planes [zb, ze), of an FP32 grid nx x ny x nz
// (4 <= nx <= NXMAX).  kMid: u1 of the output rows and planes is also
// stored through out_ptr(a1).  Block: 64*NW threads (one dimension).
template <class Plan, int NW, int NXMAX, bool kMid = false>
__device__ __forceinline__ void PackTB2Tile(const typename Plan::Args &a1, const typename Plan::Args &a2, int nx,
                                            int ny, int nz, int ya, int zb, int ze, PackTB2Lds<NW, NXMAX> &L) {
  using S = typename Plan::Storage;
  using st = typename S::stored_t;
  static_assert(kPackStorage<S>, "FP32 values");
  constexpr bool W = Plan::kWrap;
  constexpr int kT = 64 * NW;
  constexpr int kS0 = PackTB2Lds<NW, NXMAX>::kS0, kS1 = PackTB2Lds<NW, NXMAX>::kS1;
  nx = __builtin_amdgcn_readfirstlane(nx);
  ny = __builtin_amdgcn_readfirstlane(ny);
  nz = __builtin_amdgcn_readfirstlane(nz);
  ya = __builtin_amdgcn_readfirstlane(ya);
  const int P = nx * ny;
  const int64_t plane = P;
  const int V = (nx + 3) >> 2, R = kT / V;
  const st *in = Plan::stage_ptr(a1);
  st *out = Plan::out_ptr(a2);
  const int lane = threadIdx.x;
  const PackLane pl = PackLaneOf(lane, nx, NW, W);
  const int r = pl.r, vi = pl.vi, xs = pl.xs;
  const int y = ya + r;
  const int yf = FixIndex<W>(y, ny);
  const int goff = (yf * nx + xs) * (int)sizeof(st);  // byte offset in a plane
  const bool out_row = r >= 1 && r <= R - 2 && y < ny;
  const int soff = out_row ? goff : kPackNoStore;
  // Positions: u0 rows at (r + 1) V (the halo rows at 0 and (R + 1) V), u1
  // rows at r V.
  const int q0 = (r + 1) * V + vi, q1 = r * V + vi;
  // u2's y neighbours: at a clamped face the row's own u1; lanes that write
  // nothing read their own row too (in range).
  const int qn = (!out_row || (!W && y == 0)) ? q1 : q1 - V;
  const int qs = (!out_row || (!W && y == ny - 1)) ? q1 : q1 + V;
  const int wj = pl.wj, wd = pl.wd, ej = pl.ej, ed = pl.ed;
  const int aw0 = wj * kS0 + q0 + wd, ae0 = ej * kS0 + q0 + ed;
  const int aw1 = wj * kS1 + q1 + wd, ae1 = ej * kS1 + q1 + ed;

  // Halo rows ya - 1 (positions [0, V)) and ya + R (positions [(R+1) V,
  // (R+2) V)): element t of the 2 nx per lane and round, loaded and
  // published without a branch (a lane past them repeats element t mod 2 nx:
  // the same load, the same words).  A point in the overlap of a row's last
  // two vectors is written to both.
  constexpr int kHN = (2 * NXMAX + kT - 1) / kT;
  int hoff[kHN], ha[kHN], hb[kHN];
  {
    const int hy_lo = FixIndex<W>(ya - 1, ny), hy_hi = FixIndex<W>(ya + R, ny);
#pragma unroll
    for (int j = 0; j < kHN; ++j) {
      int t = lane + j * kT;
      t = t < 2 * nx ? t : t % (2 * nx);
      const bool hi = t >= nx;
      const int x = hi ? t - nx : t;
      hoff[j] = ((hi ? hy_hi : hy_lo) * nx + x) * (int)sizeof(st);
      const PackHalo h = PackHaloOf(x, nx, hi ? (R + 1) * V : 0);
      ha[j] = h.ja * kS0 + h.pa;
      hb[j] = h.jb * kS0 + h.pb;
    }
  }
  float hraw[kHN];
  auto ldh = [&](int zf) {
    const __amdgpu_buffer_rsrc_t rs = PackPlane(in, (int64_t)zf * plane, P);
#pragma unroll
    for (int j = 0; j < kHN; ++j)
      hraw[j] = PackLoad1<S>(rs, hoff[j]);
  };
  auto puth = [&](float *R0) {
#pragma unroll
    for (int j = 0; j < kHN; ++j) {
      R0[ha[j]] = hraw[j];
      R0[hb[j]] = hraw[j];
    }
  };
  auto ld = [&](int zf, float (&v)[4]) { PackLoad<S>(PackPlane(in, (int64_t)zf * plane, P), goff, v); };
  // The lane's four points into the slot arrays (stride S) at position q.
  auto put = [&](float *Rb, int q, int S, const float (&v)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) Rb[j * S + q] = v[j];
  };
  auto get = [&](const float *Rb, int q, int S, float (&v)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = Rb[j * S + q];
  };

  // u1 of the lane's points at plane k (index kf) from u0 planes b, c, t.
  auto level1 = [&](const float *R0, int kf, const float (&b)[4], const float (&c)[4], const float (&t)[4],
                    float (&u)[4]) {
    float nv[4], sv[4];
    get(R0, q0 - V, kS0, nv);
    get(R0, q0 + V, kS0, sv);
    const float wl = R0[aw0], er = R0[ae0];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      StarCtx cx;
      cx.c = c[j];
      cx.w = j > 0 ? c[j - 1] : wl;
      cx.e = j < 3 ? c[j + 1] : er;
      cx.n = nv[j];
      cx.s = sv[j];
      cx.b = b[j];
      cx.t = t[j];
      u[j] = ComputeAt<Plan, true, true, true>(cx, a1, xs + j, yf, kf);
    }
    // 16-bit grids: the intermediate level holds what a separate step would
    // store and load back.
    RoundTripN<S>(u);
  };
  // u2 of the lane's points at plane k - 1 from u1 planes b, c, t.
  auto level2 = [&](const float *R1, int k, const float (&b)[4], const float (&c)[4], const float (&t)[4]) {
    float nv[4], sv[4], o[4];
    get(R1, qn, kS1, nv);
    get(R1, qs, kS1, sv);
    const float wl = R1[aw1], er = R1[ae1];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      StarCtx cx;
      cx.c = c[j];
      cx.w = j > 0 ? c[j - 1] : wl;
      cx.e = j < 3 ? c[j + 1] : er;
      cx.n = nv[j];
      cx.s = sv[j];
      cx.b = b[j];
      cx.t = t[j];
      o[j] = ComputeAt<Plan, true, true, false>(cx, a2, xs + j, y, k - 1);
    }
    PackStore<S>(PackPlane(out, (int64_t)(k - 1) * plane, P), soff, o);
  };

  float Q0[3][4], Q1[3][4] = {};
  // Issued in the order the loop leaves its loads in flight (the halo plane
  // before the newest u0 plane; the scheduling barriers keep the order).
  ld(FixIndex<W>(zb - 2, nz), Q0[0]);
  ld(FixIndex<W>(zb - 1, nz), Q0[1]);
  __builtin_amdgcn_sched_barrier(0);
  ldh(FixIndex<W>(zb - 1, nz));
  __builtin_amdgcn_sched_barrier(0);
  ld(FixIndex<W>(zb, nz), Q0[2]);
  __builtin_amdgcn_sched_barrier(0);
  // Wrapped/clamped indices of planes k, k + 1, k + 2.
  int z0 = FixIndex<W>(zb - 1, nz), z1 = FixIndex<W>(zb, nz), z2 = FixIndex<W>(zb + 1, nz);
  auto step = [&](int k, auto rc) {
    constexpr int rq = decltype(rc)::value;
    constexpr int j0 = rq, j1 = (rq + 1) % 3, j2 = (rq + 2) % 3;
    const int buf = (k + 2) & 1;
    float *R0 = L.u0[buf];
    float *R1 = L.u1[buf];
    put(R0, q0, kS0, Q0[j1]);
    puth(R0);
    put(R1, q1, kS1, Q1[j1]);
    // z1 is an in-grid index for every k: at k == ze the load is unused.
    ldh(z1);
    __syncthreads();
    level1(R0, z0, Q0[j0], Q0[j1], Q0[j2], Q1[j2]);
    if constexpr (kMid) {
      if (k >= zb && k < ze) PackStore<S>(PackPlane(Plan::out_ptr(a1), (int64_t)k * plane, P), soff, Q1[j2]);
    }
    ld(z2, Q0[j0]);
    if (k - 1 >= zb) level2(R1, k, Q1[j0], Q1[j1], Q1[j2]);
    z0 = z1;
    z1 = z2;
    z2 = NextPlane<W>(z2, nz);
  };
  using C0 = std::integral_constant<int, 0>;
  using C1 = std::integral_constant<int, 1>;
  using C2 = std::integral_constant<int, 2>;
  int k = zb - 1;
  for (; k + 2 <= ze; k += 3) {
    step(k, C0{});
    step(k + 1, C1{});
    step(k + 2, C2{});
  }
  if (k <= ze) step(k, C0{});
  if (k + 1 <= ze) step(k + 1, C1{});
}

// One map on the packed tile: out = body(in) on rows [ya, ya+R) (those
// inside the grid), planes [zb, ze), of an FP32 grid (4 <= nx <= NXMAX):
// the level-1 half of PackTB2Tile, whose values at in-grid points are the
// map's, with every row an output row.  The LDS layout is PackTB2Tile's u0.
template <int NW, int NXMAX>
struct PackStepLds {
  static constexpr int kS0 = PackTB2Lds<NW, NXMAX>::kS0;
  float u0[2][4 * kS0];
};

template <class Plan, int NW, int NXMAX>
__device__ __forceinline__ void PackStepTile(const typename Plan::Args &a, int nx, int ny, int nz, int ya, int zb,
                                             int ze, PackStepLds<NW, NXMAX> &L) {
  using S = typename Plan::Storage;
  using st = typename S::stored_t;
  static_assert(kPackStorage<S>, "FP32 values");
  constexpr bool W = Plan::kWrap;
  constexpr int kT = 64 * NW;
  constexpr int kS0 = PackStepLds<NW, NXMAX>::kS0;
  nx = __builtin_amdgcn_readfirstlane(nx);
  ny = __builtin_amdgcn_readfirstlane(ny);
  nz = __builtin_amdgcn_readfirstlane(nz);
  ya = __builtin_amdgcn_readfirstlane(ya);
  const int P = nx * ny;
  const int64_t plane = P;
  const int V = (nx + 3) >> 2, R = kT / V;
  const st *in = Plan::stage_ptr(a);
  st *out = Plan::out_ptr(a);
  const int lane = threadIdx.x;
  const PackLane pl = PackLaneOf(lane, nx, NW, W);
  const int r = pl.r, vi = pl.vi, xs = pl.xs;
  const int y = ya + r;
  const int yf = FixIndex<W>(y, ny);
  const int goff = (yf * nx + xs) * (int)sizeof(st);
  const int soff = !pl.shadow && y < ny ? goff : kPackNoStore;
  const int q0 = (r + 1) * V + vi;
  const int wj = pl.wj, wd = pl.wd, ej = pl.ej, ed = pl.ed;
  const int aw0 = wj * kS0 + q0 + wd, ae0 = ej * kS0 + q0 + ed;
  constexpr int kHN = (2 * NXMAX + kT - 1) / kT;
  int hoff[kHN], ha[kHN], hb[kHN];
  {
    const int hy_lo = FixIndex<W>(ya - 1, ny), hy_hi = FixIndex<W>(ya + R, ny);
#pragma unroll
    for (int j = 0; j < kHN; ++j) {
      int t = lane + j * kT;
      t = t < 2 * nx ? t : t % (2 * nx);
      const bool hi = t >= nx;
      const int x = hi ? t - nx : t;
      hoff[j] = ((hi ? hy_hi : hy_lo) * nx + x) * (int)sizeof(st);
      const PackHalo h = PackHaloOf(x, nx, hi ? (R + 1) * V : 0);
      ha[j] = h.ja * kS0 + h.pa;
      hb[j] = h.jb * kS0 + h.pb;
    }
  }
  float hraw[kHN];
  auto ldh = [&](int zf) {
    const __amdgpu_buffer_rsrc_t rs = PackPlane(in, (int64_t)zf * plane, P);
#pragma unroll
    for (int j = 0; j < kHN; ++j)
      hraw[j] = PackLoad1<S>(rs, hoff[j]);
  };
  auto ld = [&](int zf, float (&v)[4]) { PackLoad<S>(PackPlane(in, (int64_t)zf * plane, P), goff, v); };

  float Q0[3][4];
  // Loads in flight at the loop head, oldest first: the halo rows of plane
  // z, then plane z + 1.
  ld(FixIndex<W>(zb - 1, nz), Q0[0]);
  ld(FixIndex<W>(zb, nz), Q0[1]);
  __builtin_amdgcn_sched_barrier(0);
  ldh(FixIndex<W>(zb, nz));
  __builtin_amdgcn_sched_barrier(0);
  ld(FixIndex<W>(zb + 1, nz), Q0[2]);
  __builtin_amdgcn_sched_barrier(0);
  // Wrapped/clamped indices of planes z + 1 and z + 2.
  int z1 = FixIndex<W>(zb + 1, nz), z2 = FixIndex<W>(zb + 2, nz);
  // Step z with role rotation rq = (z - zb) % 3: planes z-1, z, z+1 in
  // Q0[rq], Q0[rq+1], Q0[rq+2]; plane z+2 is loaded into Q0[rq] once the
  // point is computed.
  auto step = [&](int z, auto rc) {
    constexpr int rq = decltype(rc)::value;
    constexpr int j0 = rq, j1 = (rq + 1) % 3, j2 = (rq + 2) % 3;
    float *R0 = L.u0[z & 1];
#pragma unroll
    for (int j = 0; j < 4; ++j) R0[j * kS0 + q0] = Q0[j1][j];
#pragma unroll
    for (int j = 0; j < kHN; ++j) {
      R0[ha[j]] = hraw[j];
      R0[hb[j]] = hraw[j];
    }
    // z1 is an in-grid index for every z: at the chunk's end the load is
    // unused.
    ldh(z1);
    __syncthreads();
    float nv[4], sv[4], o[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      nv[j] = R0[j * kS0 + q0 - V];
      sv[j] = R0[j * kS0 + q0 + V];
    }
    const float wl = R0[aw0], er = R0[ae0];
    const float (&b)[4] = Q0[j0];
    const float (&c)[4] = Q0[j1];
    const float (&t)[4] = Q0[j2];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      StarCtx cx;
      cx.c = c[j];
      cx.w = j > 0 ? c[j - 1] : wl;
      cx.e = j < 3 ? c[j + 1] : er;
      cx.n = nv[j];
      cx.s = sv[j];
      cx.b = b[j];
      cx.t = t[j];
      o[j] = ComputeAt<Plan, true, true, true>(cx, a, xs + j, y, z);
    }
    ld(z2, Q0[j0]);
    PackStore<S>(PackPlane(out, (int64_t)z * plane, P), soff, o);
    z1 = z2;
    z2 = NextPlane<W>(z2, nz);
  };
  using C0 = std::integral_constant<int, 0>;
  using C1 = std::integral_constant<int, 1>;
  using C2 = std::integral_constant<int, 2>;
  int z = zb;
  for (; z + 3 <= ze; z += 3) {
    step(z, C0{});
    step(z + 1, C1{});
    step(z + 2, C2{});
  }
  if (z < ze) step(z, C0{});
  if (z + 1 < ze) step(z + 1, C1{});
}

// One map over one whole FP32 grid on the packed tile: workgroup
// (blockIdx.x, blockIdx.y) = (tile of PackRows rows, z chunk); tiles are
// renumbered per XCD.
template <class Plan, int NW, int NXMAX>
__global__ void __launch_bounds__(64 * NW, 8)
PackStepGrid(typename Plan::Args a, int nx, int ny, int nz, int zchunk) {
  __shared__ PackStepLds<NW, NXMAX> L;
  const int ty = XcdTile(blockIdx.x, gridDim.x);
  const int zb = (int)blockIdx.y * zchunk;
  PackStepTile<Plan, NW, NXMAX>(a, nx, ny, nz, ty * PackRows(nx, NW), zb, min(zb + zchunk, nz), L);
}

// Two fused maps over one whole FP32 grid on the packed tile: workgroup
// (blockIdx.x, blockIdx.y) = (tile of PackRows - 2 output rows, z chunk);
// tiles are renumbered per XCD.
template <class Plan, int NW, int NXMAX, bool kMid>
__global__ void __launch_bounds__(64 * NW, 8)
PackTB2Grid(typename Plan::Args a1, typename Plan::Args a2, int nx, int ny, int nz, int zchunk) {
  __shared__ PackTB2Lds<NW, NXMAX> L;
  const int rows_out = PackRows(nx, NW) - 2;
  const int ty = XcdTile(blockIdx.x, gridDim.x);
  const int zb = (int)blockIdx.y * zchunk;
  PackTB2Tile<Plan, NW, NXMAX, kMid>(a1, a2, nx, ny, nz, ty * rows_out - 1, zb, min(zb + zchunk, nz), L);
}

// Compute units of the current device (queried once per device).
inline int PackDeviceCUs() {
  static int cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (dev >= 0 && dev < 64 && cache[dev] > 0) return cache[dev];
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  if (dev >= 0 && dev < 64) cache[dev] = cus;
  return cus;
}

// z chunk of a whole-grid packed launch: the one in [4, 64] that minimises
// rounds x (chunk + 4), where a round is one workgroup per slot of the
// device (four 8-wave or two 16-wave workgroups per CU, the LDS and
// register limits) and the 4 planes stand for a chunk's halo planes and
// pipeline fill.  A rule that only asks for enough waves leaves a last round
// nearly empty when the workgroup count falls just past a multiple of the
// slots (300^3: 8-plane chunks fill 2.08 rounds).
inline int PackZChunk(int tiles_y, int nz, int nw) {
  const int64_t slots = (int64_t)PackDeviceCUs() * (nw == 8 ? 4 : 2);
  int best = 4;
  int64_t best_cost = -1;
  for (int z = 4; z <= 64; ++z) {
    const int64_t groups = (int64_t)tiles_y * ((nz + z - 1) / z);
    const int64_t cost = (groups + slots - 1) / slots * (std::min(z, nz) + 4);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = z;
    }
  }
  return best;
}

// Host launch of PackTB2Grid for a plan of FP32 values (stored in 32 or 16
// bits) and rows of 4 to 512 points
// (8 waves per workgroup up to 256 points, else 16), z chunk by PackZChunk;
// hipErrorNotSupported for other grids and storage kinds.
// mid: also store the intermediate level through out_ptr(a1).
template <class Plan>
inline hipError_t LaunchPackTB2Grid(const typename Plan::Args &a1, const typename Plan::Args &a2, int nx, int ny,
                                    int nz, hipStream_t stream, bool mid = false) {
  if constexpr (!kPackStorage<typename Plan::Storage>) {
    return hipErrorNotSupported;
  } else {
    if (nx < 4 || nx > 512 || ny < 1 || nz < 1) return hipErrorNotSupported;
    if (!PackPlaneFits(nx, ny, (int)sizeof(typename Plan::Storage::stored_t))) return hipErrorNotSupported;
    const bool wide = nx > 256;
    const int nw = wide ? 16 : 8;
    if (!PackSupported(nx, nw, wide ? 512 : 256)) return hipErrorNotSupported;
    const int rows_out = PackRows(nx, nw) - 2;
    const int tiles_y = (ny + rows_out - 1) / rows_out;
    const int zchunk = PackZChunk(tiles_y, nz, nw);
    const dim3 grid(tiles_y, (nz + zchunk - 1) / zchunk, 1), block(64 * nw, 1, 1);
#define PSK_PACK(NW, NXMAX)                                                                                       \
  if (mid)                                                                                                        \
    hipLaunchKernelGGL((PackTB2Grid<Plan, NW, NXMAX, true>), grid, block, 0, stream, a1, a2, nx, ny, nz, zchunk);  \
  else                                                                                                            \
    hipLaunchKernelGGL((PackTB2Grid<Plan, NW, NXMAX, false>), grid, block, 0, stream, a1, a2, nx, ny, nz, zchunk);
    if (wide) {
      PSK_PACK(16, 512)
    } else {
      PSK_PACK(8, 256)
    }
#undef PSK_PACK
    return hipGetLastError();
  }
}

// Host launch of PackStepGrid for a plan of FP32 values (stored in 32 or 16
// bits) and rows of 4 to 512 points
// (8 waves per workgroup up to 256 points, else 16; z chunk by
// PackZChunk); hipErrorNotSupported for other grids and storage kinds.
template <class Plan>
inline hipError_t LaunchPackStepGrid(const typename Plan::Args &a, int nx, int ny, int nz, hipStream_t stream) {
  if constexpr (!kPackStorage<typename Plan::Storage>) {
    return hipErrorNotSupported;
  } else {
    if (nx < 4 || nx > 512 || ny < 1 || nz < 1) return hipErrorNotSupported;
    if (!PackPlaneFits(nx, ny, (int)sizeof(typename Plan::Storage::stored_t))) return hipErrorNotSupported;
    const bool wide = nx > 256;
    const int nw = wide ? 16 : 8;
    const int rows = PackRows(nx, nw);
    const int tiles_y = (ny + rows - 1) / rows;
    const int zchunk = PackZChunk(tiles_y, nz, nw);
    const dim3 grid(tiles_y, (nz + zchunk - 1) / zchunk, 1), block(64 * nw, 1, 1);
    if (wide)
      hipLaunchKernelGGL((PackStepGrid<Plan, 16, 512>), grid, block, 0, stream, a, nx, ny, nz, zchunk);
    else
      hipLaunchKernelGGL((PackStepGrid<Plan, 8, 256>), grid, block, 0, stream, a, nx, ny, nz, zchunk);
    return hipGetLastError();
  }
}

}  // namespace psk
