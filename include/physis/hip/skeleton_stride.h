// skeleton_stride.h -- star-stencil tiles for rows of any length.
//
// A wave owns one whole row of up to 64*NP points: lane i holds x = i + 64k
// (k < NP), so every load and store instruction touches 64 consecutive
// elements whatever the row's alignment.  A workgroup holds BY rows of one
// z plane.  As in the 16-byte skeletons, z-1/z/z+1 of a lane's points are a
// register queue with the next plane's load in flight, and each step
// publishes the row to an LDS buffer (double-buffered by step parity, four
// pad floats at each end) from which x-1/x+1 and y-1/y+1 are read.  Periodic
// plans write the wrapped row ends into the pads; clamped plans never select
// them.  Halo rows and planes are loaded at wrapped/clamped coordinates, so
// the y and z neighbours a one-step point or a fused step's u1 point reads
// are the elements there: plans with kExactNeighbours drop their y/z
// clamped-face selects (ComputeAt).  A fused step's u2 reads its own row at
// a clamped y face (intermediate halo rows are not copies of in-grid ones)
// and keeps the z-face selects.  (Filling clamped pads per row too, to drop
// the x selects, made the four- and eight-point kernels spill.)
//
// These are tile-level device functions: the single-grid launchers and the
// batched launchers (many grids in one launch) call them per work item.
//   StrideStep   one map:  out(z) = body(in planes z-1, z, z+1)
//   StrideTB2    two maps fused: u1 = body(u0) on chip, out = body(u1)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "physis/hip/skeleton.h"
#include "physis/hip/skeleton_star.h"

namespace psk {

template <int NP, int BY>
struct StrideLds {
  static constexpr int kRow = 64 * NP + 8;
  float step[2][BY + 2][kRow];   // StrideStep rows
};

template <int NP, int BY>
struct StrideTB2Lds {
  static constexpr int kRow = 64 * NP + 8;
  float u0[2][BY + 2][kRow];
  float u1[2][BY][kRow];
};

// Word 3 of a GFX9 buffer resource for raw (untyped, unswizzled) access.
constexpr int kBufferWord3 = 0x00020000;

// One grid row of nx elements addressed through a buffer resource: the
// hardware's range check returns 0 for loads past the row's end and drops
// stores there, so lanes need no guard, branch or exec-mask update, and
// their offsets are 32-bit constants (no 64-bit address arithmetic per
// access).  Stores are non-temporal, like Storage::store.
//
// The row must be the same for every lane of the wave (a wave owns whole
// rows): its address and length are broadcast from the first lane, so the
// resource sits in scalar registers even where the compiler cannot prove the
// values uniform (pointers a batched kernel reads from its argument table);
// otherwise it would loop over the distinct values lane by lane.
template <class S>
struct BufRow {
  using st = typename S::stored_t;
  static_assert(sizeof(st) == 4 || sizeof(st) == 2, "float grids, stored in 32 or 16 bits");
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ static st *Uniform(const st *p) {
    const uint64_t a = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return reinterpret_cast<st *>(((uint64_t)hi << 32) | lo);
  }
  __device__ __forceinline__ BufRow(const st *row, int nx)
      : r(__builtin_amdgcn_make_buffer_rsrc(Uniform(row), (short)0,
                                            __builtin_amdgcn_readfirstlane(nx * (int)sizeof(st)), kBufferWord3)) {}
  __device__ __forceinline__ float load(int x) const {
    if constexpr (sizeof(st) == 4)
      return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, x * 4, 0, 0));
    else
      return S::widen(__builtin_amdgcn_raw_buffer_load_b16(r, x * 2, 0, 0));
  }
  __device__ __forceinline__ void store(int x, float v) const {
    if constexpr (sizeof(st) == 4)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, x * 4, 0, 2);
    else
      __builtin_amdgcn_raw_buffer_store_b16(S::narrow(v), r, x * 2, 0, 2);
  }
};

// Publishes the NP points of every lane of a row into LDS row r (kRow =
// 64*NP + 8 floats: lanes past the row's end write inside it, never read as
// in-row points).  Wrap plans then fill the pads with the wrapped row ends;
// those writes come after, so they win over a lane past the end that wrote
// the right pad's position.  Clamped plans never select the pads.
template <bool W, int NP>
__device__ __forceinline__ void StridePut(float *r, const float (&v)[NP], int lane, int nx) {
#pragma unroll
  for (int k = 0; k < NP; ++k) r[lane + 64 * k + 4] = v[k];
  if (W) {
#pragma unroll
    for (int k = 0; k < NP; ++k)
      if (lane + 64 * k == nx - 1) r[3] = v[k];
    if (lane == 0) r[nx + 4] = v[0];
  }
}

// Wrapped/clamped index of plane z + 1 given that of plane z (no integer
// modulo, which costs ~25 VALU instructions on this target).
template <bool W>
__device__ __forceinline__ int NextPlane(int z, int nz) {
  return W ? (z + 1 == nz ? 0 : z + 1) : min(z + 1, nz - 1);
}

// The NP points of every lane of a row (0 past the row's end).
template <class S, int NP>
__device__ __forceinline__ void StrideLoad(const typename S::stored_t *row, float (&v)[NP], int lane, int nx) {
  const BufRow<S> br(row, nx);
#pragma unroll
  for (int k = 0; k < NP; ++k) v[k] = br.load(lane + 64 * k);
}

// One map over rows [y0, y0+BY) x planes [zb, ze) of a grid nx x ny x nz
// (nx <= 64*NP), x over the whole row.  Out-of-range rows are skipped.
template <class Plan, int NP, int BY>
__device__ __forceinline__ void StrideStep(const typename Plan::Args &a, int nx, int ny, int nz, int y0, int zb,
                                           int ze, StrideLds<NP, BY> &L) {
  using S = typename Plan::Storage;
  using st = typename S::stored_t;
  constexpr bool W = Plan::kWrap;
  nx = __builtin_amdgcn_readfirstlane(nx);
  ny = __builtin_amdgcn_readfirstlane(ny);
  nz = __builtin_amdgcn_readfirstlane(nz);
  const int lane = threadIdx.x, wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int y = y0 + wy;
  const int64_t plane = (int64_t)nx * ny;
  const st *in = BufRow<S>::Uniform(Plan::stage_ptr(a));
  st *out = BufRow<S>::Uniform(Plan::out_ptr(a));
  const int64_t row = (int64_t)FixIndex<W>(y, ny) * nx;
  const bool halo_lo = wy == 0, halo_hi = wy == BY - 1;
  const int64_t hlo_row = (int64_t)FixIndex<W>(y0 - 1, ny) * nx;
  const int64_t hhi_row = (int64_t)FixIndex<W>(y0 + BY, ny) * nx;
  const bool active = y >= 0 && y < ny;
  auto zo = [&](int z) { return (int64_t)FixIndex<W>(z, nz) * plane; };

  float b[NP], c[NP], t[NP], p[NP] = {}, hlo[NP] = {}, hhi[NP] = {}, nlo[NP] = {}, nhi[NP] = {};
  StrideLoad<S, NP>(in + zo(zb - 1) + row, b, lane, nx);
  StrideLoad<S, NP>(in + zo(zb) + row, c, lane, nx);
  StrideLoad<S, NP>(in + zo(zb + 1) + row, t, lane, nx);
  if (halo_lo) StrideLoad<S, NP>(in + zo(zb) + hlo_row, hlo, lane, nx);
  if (halo_hi) StrideLoad<S, NP>(in + zo(zb) + hhi_row, hhi, lane, nx);
  int z1 = FixIndex<W>(zb + 1, nz), z2 = FixIndex<W>(zb + 2, nz);  // planes z+1, z+2
  for (int z = zb; z < ze; ++z) {
    if (z + 1 < ze) {
      StrideLoad<S, NP>(in + (int64_t)z2 * plane + row, p, lane, nx);
      if (halo_lo) StrideLoad<S, NP>(in + (int64_t)z1 * plane + hlo_row, nlo, lane, nx);
      if (halo_hi) StrideLoad<S, NP>(in + (int64_t)z1 * plane + hhi_row, nhi, lane, nx);
    }
    z1 = z2;
    z2 = NextPlane<W>(z2, nz);
    auto &R = L.step[z & 1];
    StridePut<W, NP>(R[wy + 1], c, lane, nx);
    if (halo_lo) StridePut<W, NP>(R[0], hlo, lane, nx);
    if (halo_hi) StridePut<W, NP>(R[BY + 1], hhi, lane, nx);
    __syncthreads();
    if (active) {
      // Points past the row's end compute on in-bounds LDS values; their
      // stores are dropped by the row's range.
      const BufRow<S> o(out + (int64_t)z * plane + (int64_t)y * nx, nx);
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const int x = lane + 64 * k;
        StarCtx cx;
        cx.c = c[k];
        cx.w = R[wy + 1][x + 3];
        cx.e = R[wy + 1][x + 5];
        cx.n = R[wy][x + 4];
        cx.s = R[wy + 2][x + 4];
        cx.b = b[k];
        cx.t = t[k];
        o.store(x, ComputeAt<Plan, false, true, true>(cx, a, x, y, z));
      }
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      b[k] = c[k];
      c[k] = t[k];
      t[k] = p[k];
      hlo[k] = nlo[k];
      hhi[k] = nhi[k];
    }
  }
}

// Two fused maps: u1 = body(u0 from stage_ptr(a1)) on rows [y0-1+1 ...]:
// the tile computes u1 on rows [ya, ya+BY) and writes u2 = body(u1) through
// out_ptr(a2) on rows [ya+1, ya+BY-1), planes [zb, ze).  kMid: u1 of those
// rows and planes is also stored through out_ptr(a1).
template <class Plan, int NP, int BY, bool kMid = false>
__device__ __forceinline__ void StrideTB2(const typename Plan::Args &a1, const typename Plan::Args &a2, int nx,
                                          int ny, int nz, int ya, int zb, int ze, StrideTB2Lds<NP, BY> &L) {
  using S = typename Plan::Storage;
  using st = typename S::stored_t;
  constexpr bool W = Plan::kWrap;
  // wy is the same for every lane of a wave (blockDim.x == 64): broadcast,
  // it keeps row offsets scalar and row-role branches free of exec masks.
  // Grid pointers and extents are the same for the whole workgroup: broadcast
  // so that row addresses are computed in scalar registers.
  nx = __builtin_amdgcn_readfirstlane(nx);
  ny = __builtin_amdgcn_readfirstlane(ny);
  nz = __builtin_amdgcn_readfirstlane(nz);
  const int lane = threadIdx.x;
  const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int y = ya + wy;
  const int yf = FixIndex<W>(y, ny);
  const int64_t plane = (int64_t)nx * ny;
  const st *in = BufRow<S>::Uniform(Plan::stage_ptr(a1));
  st *out = BufRow<S>::Uniform(Plan::out_ptr(a2));
  const int64_t row = (int64_t)yf * nx;
  // The two halo rows (ya - 1 into slot 0, ya + BY into slot BY + 1) are
  // loaded by all lanes of the workgroup together, kHN elements per lane of
  // the 2 * 64 * NP row elements, without a branch (a load inside a branch
  // makes the compiler wait for every load in flight where the paths join).
  // Lanes past a row end, or past the rows, read an in-bounds element and
  // publish nothing.  Elements stay in storage form until published.
  constexpr int kW = 64 * NP, kT = 64 * BY;
  constexpr int kHN = (2 * kW + kT - 1) / kT;
  const int tid = lane + wy * 64;
  int hoff[kHN], hx[kHN];
  bool hhi[kHN], hpub[kHN];
  {
    const int hy_lo = FixIndex<W>(ya - 1, ny), hy_hi = FixIndex<W>(ya + BY, ny);
#pragma unroll
    for (int j = 0; j < kHN; ++j) {
      const int t = tid + j * kT;
      const int e = t % (2 * kW);
      hhi[j] = e >= kW;
      hx[j] = hhi[j] ? e - kW : e;
      hpub[j] = t < 2 * kW && hx[j] < nx;
      hoff[j] = (hhi[j] ? hy_hi : hy_lo) * nx + min(hx[j], nx - 1);
    }
  }
  st hraw[kHN];
  auto ldh = [&](int64_t zo) {
#pragma unroll
    for (int j = 0; j < kHN; ++j) hraw[j] = GlobalLoad(in + zo + hoff[j]);
  };
  // Clamped plans never read the pads; wrap plans read the wrapped row ends.
  auto puth = [&](float (*R0)[StrideTB2Lds<NP, BY>::kRow]) {
#pragma unroll
    for (int j = 0; j < kHN; ++j) {
      if (!hpub[j]) continue;
      float *r = R0[hhi[j] ? BY + 1 : 0];
      const float v = S::widen(hraw[j]);
      r[hx[j] + 4] = v;
      if (W) {
        if (hx[j] == nx - 1) r[3] = v;
        if (hx[j] == 0) r[nx + 4] = v;
      }
    }
  };
  const bool out_row = wy >= 1 && wy <= BY - 2 && y >= 0 && y < ny;
  auto zo = [&](int z) { return (int64_t)FixIndex<W>(z, nz) * plane; };
  // Every lane loads without a branch (points past the row end read 0
  // through the row's range and are never stored).
  auto ld = [&](const st *r, float (&v)[NP]) { StrideLoad<S, NP>(r, v, lane, nx); };
  // u1 at plane k (this wave's row) from u0 planes b, c, t.
  auto level1 = [&](float (*R0)[StrideTB2Lds<NP, BY>::kRow], int kf, const float (&b)[NP], const float (&c)[NP],
                    const float (&t)[NP], float (&u)[NP]) {
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const int x = lane + 64 * q;
      StarCtx cx;
      cx.c = c[q];
      cx.w = R0[wy + 1][x + 3];
      cx.e = R0[wy + 1][x + 5];
      cx.n = R0[wy][x + 4];
      cx.s = R0[wy + 2][x + 4];
      cx.b = b[q];
      cx.t = t[q];
      u[q] = ComputeAt<Plan, false, true, true>(cx, a1, x, yf, kf);
    }
    RoundTripN<S>(u);
  };
  // u2 output row at plane k - 1 from u1 planes b, c, t.
  auto level2 = [&](float (*R1)[StrideTB2Lds<NP, BY>::kRow], int k, const float (&b)[NP], const float (&c)[NP],
                    const float (&t)[NP]) {
    const BufRow<S> o(out + (int64_t)(k - 1) * plane + (int64_t)y * nx, nx);
    const int rn = !W && y == 0 ? wy : wy - 1, rs = !W && y == ny - 1 ? wy : wy + 1;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const int x = lane + 64 * q;
      StarCtx cx;
      cx.c = c[q];
      cx.w = R1[wy][x + 3];
      cx.e = R1[wy][x + 5];
      cx.n = R1[rn][x + 4];
      cx.s = R1[rs][x + 4];
      cx.b = b[q];
      cx.t = t[q];
      o.store(x, ComputeAt<Plan, false, true, false>(cx, a2, x, y, k - 1));
    }
  };

  // The queues hold three planes each and rotate their roles over an
  // unrolled period of three steps (no register moves): at step k with
  // r = (k - zb + 1) % 3, u0 planes k-1, k, k+1 are Q0[r], Q0[r+1], Q0[r+2]
  // and u1 planes k-2, k-1 are Q1[r], Q1[r+1]; u1 plane k goes to Q1[r+2]
  // and u0 plane k+2 is loaded into Q0[r] once u1 no longer needs it.
  // Issued in the order the loop leaves its loads in flight (the halo plane
  // before the newest u0 plane; the scheduling barriers keep the order), so
  // that the waits at the loop head are the loop's own.
  float Q0[3][NP], Q1[3][NP] = {};
  ld(in + zo(zb - 2) + row, Q0[0]);
  ld(in + zo(zb - 1) + row, Q0[1]);
  __builtin_amdgcn_sched_barrier(0);
  ldh(zo(zb - 1));
  __builtin_amdgcn_sched_barrier(0);
  ld(in + zo(zb) + row, Q0[2]);
  __builtin_amdgcn_sched_barrier(0);
  // Wrapped/clamped indices of planes k, k+1, k+2.
  int z0 = FixIndex<W>(zb - 1, nz), z1 = FixIndex<W>(zb, nz), z2 = FixIndex<W>(zb + 1, nz);
  auto step = [&](int k, auto rc) {
    constexpr int r = decltype(rc)::value;
    constexpr int i0 = r, i1 = (r + 1) % 3, i2 = (r + 2) % 3;
    const int buf = (k + 2) & 1;
    auto &R0 = L.u0[buf];
    auto &R1 = L.u1[buf];
    StridePut<W, NP>(R0[wy + 1], Q0[i1], lane, nx);
    puth(R0);
    StridePut<W, NP>(R1[wy], Q1[i1], lane, nx);
    // z1 is an in-grid index for every k: at k == ze the load is unused.
    ldh((int64_t)z1 * plane);
    __syncthreads();
    level1(R0, z0, Q0[i0], Q0[i1], Q0[i2], Q1[i2]);
    if constexpr (kMid) {
      if (out_row && k >= zb && k < ze) {
        const BufRow<S> m(BufRow<S>::Uniform(Plan::out_ptr(a1)) + (int64_t)k * plane + (int64_t)y * nx, nx);
#pragma unroll
        for (int q = 0; q < NP; ++q) m.store(lane + 64 * q, Q1[i2][q]);
      }
    }
    ld(in + (int64_t)z2 * plane + row, Q0[i0]);
    if (out_row && k - 1 >= zb) level2(R1, k, Q1[i0], Q1[i1], Q1[i2]);
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

// One map over one whole grid: workgroup (blockIdx.x, blockIdx.y) = (y tile
// of BY rows, z chunk); tiles are renumbered per XCD.
template <class Plan, int NP, int BY>
__global__ void __launch_bounds__(64 * BY)
StrideStepGrid(typename Plan::Args a, int nx, int ny, int nz, int zchunk) {
  __shared__ StrideLds<NP, BY> L;
  const int ty = XcdTile(blockIdx.x, gridDim.x);
  const int zb = (int)blockIdx.y * zchunk;
  StrideStep<Plan, NP, BY>(a, nx, ny, nz, ty * BY, zb, min(zb + zchunk, nz), L);
}

// z-chunk of a whole-grid stride launch: the longest of 32, 16 and 8 planes
// that gives at least 16384 waves, else 4096, else 8.
inline int StrideZChunk(int tiles_y, int nz, int waves_per_group) {
  auto waves = [&](int z) { return (int64_t)tiles_y * ((nz + z - 1) / z) * waves_per_group; };
  for (int want : {16384, 4096})
    for (int z : {32, 16, 8})
      if (waves(z) >= want) return z;
  return 8;
}

// Points per lane of a row of nx points: the fewest of 1, 2, 3, 4, 6 and 8
// that cover it (a lane computes all NP of its points whether or not they are
// in the row, so the wave's width follows the row's).
inline int StrideNp(int nx) {
  return nx <= 64 ? 1 : nx <= 128 ? 2 : nx <= 192 ? 3 : nx <= 256 ? 4 : nx <= 384 ? 6 : 8;
}

// Host launch of StrideStepGrid for rows of any length up to 512 points (NP
// = StrideNp(nx), 8 rows per workgroup); hipErrorNotSupported for wider
// rows.
template <class Plan>
inline hipError_t LaunchStrideStepGrid(const typename Plan::Args &a, int nx, int ny, int nz, hipStream_t stream) {
  constexpr int BY = 8;
  if (nx < 1 || nx > 512 || ny < 1 || nz < 1) return hipErrorNotSupported;
  const int tiles_y = (ny + BY - 1) / BY;
  const int zchunk = StrideZChunk(tiles_y, nz, BY);
  const dim3 grid(tiles_y, (nz + zchunk - 1) / zchunk, 1), block(64, BY, 1);
  switch (StrideNp(nx)) {
#define PSK_STEP(NP) \
  case NP: hipLaunchKernelGGL((StrideStepGrid<Plan, NP, BY>), grid, block, 0, stream, a, nx, ny, nz, zchunk); break;
    PSK_STEP(1) PSK_STEP(2) PSK_STEP(3) PSK_STEP(4) PSK_STEP(6) PSK_STEP(8)
#undef PSK_STEP
  }
  return hipGetLastError();
}

// Two fused maps over one whole grid: workgroup (blockIdx.x, blockIdx.y) =
// (y tile of BY - 2 output rows, z chunk); tiles are renumbered per XCD.
template <class Plan, int NP, int BY, bool kMid = false>
__global__ void __launch_bounds__(64 * BY)
StrideTB2Grid(typename Plan::Args a1, typename Plan::Args a2, int nx, int ny, int nz, int zchunk) {
  __shared__ StrideTB2Lds<NP, BY> L;
  const int ty = XcdTile(blockIdx.x, gridDim.x);
  const int zb = (int)blockIdx.y * zchunk;
  StrideTB2<Plan, NP, BY, kMid>(a1, a2, nx, ny, nz, ty * (BY - 2) - 1, zb, min(zb + zchunk, nz), L);
}

// Host launch of StrideTB2Grid for rows of any length up to 512 points:
// NP = StrideNp(nx), 8 rows per workgroup, and the z-chunk of StrideZChunk
// (the rule of ChooseTB2Shape).
// hipErrorNotSupported for wider rows.  mid: also store the intermediate
// level through out_ptr(a1) (kMid above).
template <class Plan>
inline hipError_t LaunchStrideTB2Grid(const typename Plan::Args &a1, const typename Plan::Args &a2, int nx, int ny,
                                      int nz, hipStream_t stream, bool mid = false) {
  constexpr int BY = 8;
  if (nx < 1 || nx > 512 || ny < 1 || nz < 1) return hipErrorNotSupported;
  const int tiles_y = (ny + BY - 3) / (BY - 2);
  const int zchunk = StrideZChunk(tiles_y, nz, BY);
  const dim3 grid(tiles_y, (nz + zchunk - 1) / zchunk, 1), block(64, BY, 1);
  switch (StrideNp(nx)) {
#define PSK_TB2(NP)                                                                                                 \
  case NP:                                                                                                          \
    if (mid)                                                                                                        \
      hipLaunchKernelGGL((StrideTB2Grid<Plan, NP, BY, true>), grid, block, 0, stream, a1, a2, nx, ny, nz, zchunk);  \
    else                                                                                                            \
      hipLaunchKernelGGL((StrideTB2Grid<Plan, NP, BY, false>), grid, block, 0, stream, a1, a2, nx, ny, nz, zchunk); \
    break;
    PSK_TB2(1) PSK_TB2(2) PSK_TB2(3) PSK_TB2(4) PSK_TB2(6) PSK_TB2(8)
#undef PSK_TB2
  }
  return hipGetLastError();
}

}  // namespace psk
