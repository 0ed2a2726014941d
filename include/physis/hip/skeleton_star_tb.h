// skeleton_star_tb.h -- two time steps per HBM pass for radius-1 star stencils.
//
// Fuses the two maps of a ping-pong PSStencilRun (map(k, dom, A, B, ...) then
// map(k, dom, B, A, ...), same kernel and scalars, domain = whole grid) into
// one z-march: u1 = step(u0) is kept on chip and only u2 = step(u1) is
// written, so one read and one write of HBM advance the field two steps.
// u2 goes to a third buffer, never back into u0: workgroups march z at
// different paces and still read u0 planes that a faster neighbour has
// passed.
//
// A workgroup spans whole rows (nx <= 256*NWX, one lane per four x points)
// and BY rows; it computes u1 on its BY rows and writes u2 on the inner
// BY-2 rows (tiles overlap by two rows in y).  Per z step k:
//   publish u0 plane k (rows y0-1 .. y0+BY) and u1 plane k-1 to LDS
//   (double-buffered by step parity: one barrier per step), then
//   u1(k)   = body(u0 planes k-1, k, k+1)          registers + LDS rows
//   u2(k-1) = body(u1 planes k-2, k-1, k)          registers + LDS rows
// Rows have four pad floats at each end: periodic plans fill them with the
// wrapped row ends, clamped plans with the replicated end points.  Every
// neighbour a u1 value reads (and every x neighbour a u2 value reads) is
// then the element at the wrapped/clamped coordinate, so a plan that sets
// kExactNeighbours may skip its clamped-face selects there (compute_exact):
// at a clamped face the DSL body selects the point itself, which is the same
// element.  Intermediate points outside the grid are evaluated at their
// wrapped/clamped coordinates but are not copies of in-grid ones, so u2
// must not read them as neighbours: at a clamped y face it reads its own row
// instead (a wave-uniform LDS row choice), and at a z face it keeps the
// plan's select.  Each u1 value is computed by the same
// body in FP32 as in the one-step path, so the fused result is bitwise equal
// to two separate steps.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "physis/hip/skeleton.h"
#include "physis/hip/skeleton_star.h"

namespace psk {

constexpr int kMaxZChunks = 64;
struct TBGeometry {
  int tiles_y;  // y tiles, each writing BY-2 rows
  int zchunk;   // planes per workgroup; -1: the chunks zs (below)
  int remap;
  int zs[kMaxZChunks + 1];  // zchunk == -1: chunk c is planes [zs[c], zs[c+1])
};

// Shared memory of one fused tile: u0 rows (with the two halo rows) and u1
// rows of a plane, double-buffered by step parity; four pad floats at each
// row end.
template <int NWX, int BY>
struct StarTB2Lds {
  static constexpr int kRow = 256 * NWX + 8;
  float u0[2][BY + 2][kRow];
  float u1[2][BY][kRow];
};

// One fused tile: u1 on rows [ya, ya+BY) and u2 on rows [ya+1, ya+BY-1)
// (those inside the grid), planes [zb, ze), of a grid of whole 16-byte rows
// (nx <= 256*NWX); called by the single-grid kernel below and by batched
// launches over many grids.  Block: 64*NWX x BY threads.
//
// kMid: the tile also stores u1 on its output rows and planes through
// out_ptr(a1) (the first map's output grid), so a run's last fused pass
// leaves both grids of the ping-pong as two separate steps leave them (one
// more write per point instead of an unfused pair of steps).
//
// kRenamed selects the loop schedule.  false: the next u0 plane is loaded
// before the barrier into a fourth register set and the queues are shifted
// by register moves at the end of each step.  true: the queues hold three
// planes each and rotate their roles over an unrolled period of three steps
// (no moves); the next u0 plane is loaded, after u1 is computed, into the
// set of the plane u1 no longer needs, and halo rows load straight into
// their own registers once published.
template <class Plan, int NWX, int BY, bool kRenamed, bool kMid = false>
__device__ __forceinline__ void StarTB2Tile(const typename Plan::Args &a1, const typename Plan::Args &a2, int nx,
                                            int ny, int nz, int ya, int zb, int ze, StarTB2Lds<NWX, BY> &L) {
  using S = typename Plan::Storage;
  using st = typename S::stored_t;
  constexpr bool W = Plan::kWrap;
  static_assert(BY >= 3, "need at least one output row");
  static_assert(!kMid || kRenamed, "the intermediate store is part of the rotating-queue schedule");
  auto &L0 = L.u0;
  auto &L1 = L.u1;

  // wx and wy are the same for every lane of a wave (blockDim.x is a
  // multiple of 64); broadcasting them lets the compiler keep row offsets in
  // scalar registers and branch on row roles without exec-mask updates.
  const int lane = threadIdx.x & 63;
  const int wx = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int y = ya + wy;
  const int xs = wx * 256 + lane * 4;
  const bool has = xs < nx;
  const int64_t plane = (int64_t)nx * ny;
  const st *in = Plan::stage_ptr(a1);
  st *out = Plan::out_ptr(a2);
  const int yf = FixIndex<W>(y, ny);
  const int64_t row = (int64_t)yf * nx + xs;
  const bool halo_lo = wy == 0, halo_hi = wy == BY - 1;
  const int64_t hrow_lo = (int64_t)FixIndex<W>(ya - 1, ny) * nx + xs;
  const int64_t hrow_hi = (int64_t)FixIndex<W>(ya + BY, ny) * nx + xs;
  const bool out_row = wy >= 1 && wy <= BY - 2 && y < ny;
  // Pad writers (the pad element next to each row end).  Periodic: the lane
  // holding x = nx-1 fills the left pad, the lane holding x = 0 the right
  // pad.  Clamped: the lane holding x = 0 fills the left pad, the lane
  // holding x = nx-1 the right pad.
  const bool first = xs == 0, last = xs == nx - 4;

  auto zoff = [&](int z) { return (int64_t)FixIndex<W>(z, nz) * plane; };
  auto ld = [&](int64_t off, float (&v)[4]) {
    if (has) {
      Load4<S>(in + off, v);
    } else {
      v[0] = v[1] = v[2] = v[3] = 0.f;
    }
  };
  auto put = [&](float *r, const float (&v)[4]) {
    if (!has) return;
    *reinterpret_cast<v4f_t *>(r + xs + 4) = v4f_t{v[0], v[1], v[2], v[3]};
    // Only the pad element next to the row is ever read.
    if (W) {
      if (last) r[3] = v[3];
      if (first) r[nx + 4] = v[0];
    } else {
      if (first) r[3] = v[0];
      if (last) r[nx + 4] = v[3];
    }
  };
  auto get4 = [&](const float *r, float (&v)[4]) {
    const v4f_t q = *reinterpret_cast<const v4f_t *>(r + xs + 4);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  };

  // One u1 point row at plane k (this wave's row) from u0 planes b, c, t.
  auto level1 = [&](int buf, int kf, const float (&b)[4], const float (&c)[4], const float (&t)[4],
                    float (&u)[4]) {
    float nv[4], sv[4];
    get4(&L0[buf][wy][0], nv);
    get4(&L0[buf][wy + 2][0], sv);
    const float wl = L0[buf][wy + 1][xs + 3], er = L0[buf][wy + 1][xs + 8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      StarCtx q;
      q.c = c[j];
      q.w = j > 0 ? c[j - 1] : wl;
      q.e = j < 3 ? c[j + 1] : er;
      q.n = nv[j];
      q.s = sv[j];
      q.b = b[j];
      q.t = t[j];
      u[j] = ComputeAt<Plan, true, true, true>(q, a1, xs + j, yf, kf);
    }
    RoundTripN<S>(u);
  };
  // One u2 output row at plane k - 1 from u1 planes b, c, t.  Clamped plans:
  // at a y face the row's own u1 values stand in for the missing neighbour
  // row (a wave-uniform choice of LDS row), so only the z faces need selects.
  auto level2 = [&](int buf, int k, const float (&b)[4], const float (&c)[4], const float (&t)[4]) {
    float n1[4], s1[4], v[4];
    const int rn = !W && y == 0 ? wy : wy - 1, rs = !W && y == ny - 1 ? wy : wy + 1;
    get4(&L1[buf][rn][0], n1);
    get4(&L1[buf][rs][0], s1);
    const float w1 = L1[buf][wy][xs + 3], e1 = L1[buf][wy][xs + 8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      StarCtx q;
      q.c = c[j];
      q.w = j > 0 ? c[j - 1] : w1;
      q.e = j < 3 ? c[j + 1] : e1;
      q.n = n1[j];
      q.s = s1[j];
      q.b = b[j];
      q.t = t[j];
      v[j] = ComputeAt<Plan, true, true, false>(q, a2, xs + j, y, k - 1);
    }
    if (has) Store4<S>(out + (int64_t)(k - 1) * plane + (int64_t)y * nx + xs, v);
  };

  if constexpr (kRenamed) {
    // Every lane loads, without a branch: a lane past the row end reads the
    // row's last four points (never stored or published).  A load inside a
    // branch makes the compiler wait for every load in flight where the
    // paths join, since it cannot tell which were issued.
    const int xl = has ? xs : nx - 4;
    const int64_t lrow = (int64_t)yf * nx + xl;
    auto ldu = [&](int64_t off, float (&v)[4]) { Load4<S>(in + off, v); };
    // The two halo rows (ya - 1 into slot 0, ya + BY into slot BY + 1) are
    // loaded by all lanes of the workgroup together, kHN elements per lane
    // of the 2 * kW row elements, also without a branch (lanes past a row
    // end, or past the rows, read an in-bounds element and publish
    // nothing).  They stay in storage form until published, so that no
    // conversion waits on a load before the plane is needed.
    constexpr int kW = 256 * NWX, kT = 64 * NWX * BY;
    constexpr int kHN = (2 * kW + kT - 1) / kT;
    const int tid = threadIdx.x + (int)threadIdx.y * 64 * NWX;
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
    auto puth = [&](int buf) {
#pragma unroll
      for (int j = 0; j < kHN; ++j) {
        if (!hpub[j]) continue;
        float *r = &L0[buf][hhi[j] ? BY + 1 : 0][0];
        const float v = S::widen(hraw[j]);
        r[hx[j] + 4] = v;
        if (W) {
          if (hx[j] == nx - 1) r[3] = v;
          if (hx[j] == 0) r[nx + 4] = v;
        } else {
          if (hx[j] == 0) r[3] = v;
          if (hx[j] == nx - 1) r[nx + 4] = v;
        }
      }
    };
    float Q0[3][4], Q1[3][4] = {};
    // 16-bit grids carry the newest u0 plane in storage form (praw) and
    // widen it into its queue slot after the next barrier, where it is first
    // used: a widened value carried around the loop would make the compiler
    // wait for its load at the loop's end.
    constexpr bool kRaw = sizeof(st) == 2;
    v2u_t praw = {};
    auto ldnext = [&](int64_t off, float (&v)[4]) {
      if constexpr (kRaw)
        praw = LoadRaw4<S>(in + off);
      else
        ldu(off, v);
    };
    // Issued in the order the loop leaves its loads in flight (the halo
    // plane before the newest u0 plane; the scheduling barriers keep the
    // order), so that the waits at the loop head are the loop's own.
    ldu(zoff(zb - 2) + lrow, Q0[0]);
    ldu(zoff(zb - 1) + lrow, Q0[1]);
    __builtin_amdgcn_sched_barrier(0);
    ldh(zoff(zb - 1));
    __builtin_amdgcn_sched_barrier(0);
    ldnext(zoff(zb) + lrow, Q0[2]);
    __builtin_amdgcn_sched_barrier(0);
    // Plane indices of k (z0), k + 1 (z1) and k + 2 (z2), as below.
    int z0 = FixIndex<W>(zb - 1, nz), z1 = FixIndex<W>(zb, nz), z2 = FixIndex<W>(zb + 1, nz);
    auto advance = [&](int z) { return W ? (z + 1 == nz ? 0 : z + 1) : min(z + 1, nz - 1); };
    // Step k with role rotation r = (k - zb + 1) % 3: u0 planes k-1, k, k+1
    // in Q0[r], Q0[r+1], Q0[r+2]; u1 planes k-2, k-1 in Q1[r], Q1[r+1], u1
    // plane k into Q1[r+2] (indices mod 3).
    auto step = [&](int k, auto rc) {
      constexpr int r = decltype(rc)::value;
      constexpr int i0 = r, i1 = (r + 1) % 3, i2 = (r + 2) % 3;
      const int buf = (k + 2) & 1;
      put(&L0[buf][wy + 1][0], Q0[i1]);
      puth(buf);
      put(&L1[buf][wy][0], Q1[i1]);
      // At k == ze the planes loaded below are not used (z1, z2 are in-grid
      // indices whatever k is, so the loads stay in bounds).
      ldh((int64_t)z1 * plane);
      __syncthreads();
      if constexpr (kRaw) Widen4<S>(praw, Q0[i2]);
      // Lanes past the row end compute on in-bounds LDS garbage that is
      // never published or stored.
      level1(buf, z0, Q0[i0], Q0[i1], Q0[i2], Q1[i2]);
      if constexpr (kMid) {
        // u1 of plane k on an output row: rounded to storage already, so
        // the stored bits are those a separate step stores.
        if (has && out_row && k >= zb && k < ze)
          Store4<S>(Plan::out_ptr(a1) + (int64_t)k * plane + (int64_t)y * nx + xs, Q1[i2]);
      }
      // Q0[i0] (plane k-1) is dead: it receives plane k+2.
      ldnext((int64_t)z2 * plane + lrow, Q0[i0]);
      if (out_row && k - 1 >= zb) level2(buf, k, Q1[i0], Q1[i1], Q1[i2]);
      z0 = z1;
      z1 = z2;
      z2 = advance(z2);
    };
    using R0 = std::integral_constant<int, 0>;
    using R1 = std::integral_constant<int, 1>;
    using R2 = std::integral_constant<int, 2>;
    int k = zb - 1;
    for (; k + 2 <= ze; k += 3) {
      step(k, R0{});
      step(k + 1, R1{});
      step(k + 2, R2{});
    }
    if (k <= ze) step(k, R0{});
    if (k + 1 <= ze) step(k + 1, R1{});
    return;
  }

  float b0[4], c0[4], t0[4], p0[4] = {}, hlo[4] = {}, hhi[4] = {}, nlo[4] = {}, nhi[4] = {};
  float b1[4] = {}, c1[4] = {}, t1[4] = {};
  ld(zoff(zb - 2) + row, b0);
  ld(zoff(zb - 1) + row, c0);
  ld(zoff(zb) + row, t0);
  if (halo_lo) ld(zoff(zb - 1) + hrow_lo, hlo);
  if (halo_hi) ld(zoff(zb - 1) + hrow_hi, hhi);
  // Wrapped/clamped plane indices of k, k + 1 and k + 2, advanced by one plane
  // per step (an integer modulo per load would cost ~25 VALU instructions on
  // this target, which has no scalar division).
  int z0 = FixIndex<W>(zb - 1, nz), z1 = FixIndex<W>(zb, nz), z2 = FixIndex<W>(zb + 1, nz);
  auto advance = [&](int z) { return W ? (z + 1 == nz ? 0 : z + 1) : min(z + 1, nz - 1); };

  for (int k = zb - 1; k <= ze; ++k) {
    const int buf = (k + 2) & 1;
    put(&L0[buf][wy + 1][0], c0);
    if (halo_lo) put(&L0[buf][0][0], hlo);
    if (halo_hi) put(&L0[buf][BY + 1][0], hhi);
    put(&L1[buf][wy][0], c1);
    if (k < ze) {
      const int64_t o2 = (int64_t)z2 * plane, o1 = (int64_t)z1 * plane;
      ld(o2 + row, p0);
      if (halo_lo) ld(o1 + hrow_lo, nlo);
      if (halo_hi) ld(o1 + hrow_hi, nhi);
    }
    __syncthreads();
    if (has) {
      level1(buf, z0, b0, c0, t0, t1);  // u1 at plane k, this wave's row
      if (out_row && k - 1 >= zb) level2(buf, k, b1, c1, t1);  // u2 at plane k-1
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      b0[j] = c0[j];
      c0[j] = t0[j];
      t0[j] = p0[j];
      hlo[j] = nlo[j];
      hhi[j] = nhi[j];
      b1[j] = c1[j];
      c1[j] = t1[j];
    }
    z0 = z1;
    z1 = z2;
    z2 = advance(z2);
  }
}

// Workgroup (b, c) takes z chunk c of y tile b (tiles remapped so that each
// XCD takes a contiguous range): planes [c * zchunk, (c + 1) * zchunk) with
// zchunk > 0, else [zs[c], zs[c + 1]) (a guided schedule: long chunks
// dispatched first, short ones last).
template <class Plan, int NWX, int BY, bool kRenamed = false, bool kMid = false>
__global__ void __launch_bounds__(64 * NWX * BY, kRenamed ? 8 : 1)
ZMarchStar4TB2(typename Plan::Args a1, typename Plan::Args a2, TBGeometry geo) {
  __shared__ StarTB2Lds<NWX, BY> L;
  int nx, ny, nz;
  Plan::stage_dims(a1, nx, ny, nz);
  const int tile = geo.remap ? XcdTile(blockIdx.x, geo.tiles_y) : (int)blockIdx.x;
  const int c = (int)blockIdx.y;
  const int zb = geo.zchunk > 0 ? c * geo.zchunk : geo.zs[c];
  const int ze = geo.zchunk > 0 ? min(zb + geo.zchunk, nz) : geo.zs[c + 1];
  StarTB2Tile<Plan, NWX, BY, kRenamed, kMid>(a1, a2, nx, ny, nz, tile * (BY - 2) - 1, zb, ze, L);
}

// Workgroups of `kernel` (blocks of `threads`) the current device keeps
// resident at once, cached per device and per Tag (one Tag per kernel).
template <class Plan, int NWX, int BY, bool kRenamed, bool kMid>
struct TB2KernelId {};
template <class Tag>
inline int ResidentWorkgroups(const void *kernel, int threads) {
  static int cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (dev >= 0 && dev < 64 && cache[dev] > 0) return cache[dev];
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const int n = per_cu * cus;
  if (dev >= 0 && dev < 64) cache[dev] = n;
  return n;
}

// Whether a grid can take the fused two-step path with NWX waves per row.
template <int NWX>
inline bool StarTB2Supported(int nx, int ny, int nz) {
  return nx % 4 == 0 && nx <= 256 * NWX && nx >= 4 && ny >= 1 && nz >= 1;
}

// Launch shape of the fused kernel for a grid: NWX = ceil(nx / 256) waves
// per row (whole rows per workgroup, so fewer idle lanes than a fixed
// width), 8 rows per workgroup (4 with four waves per row), and the largest
// z-chunk of 32, 16 and 8 planes that still gives kTB2WavesFull waves, else
// the largest that gives kTB2Waves, else 8 (a chunk re-reads two halo
// planes, so longer chunks move fewer bytes once the device is full).
// Timed on MI355X with the rotating-queue loop (GLUPS, clamped): 128^3 298,
// 192^3 673, 256^3 1011, 320^3 837, 384^3 1033, 448^3 1104, 512^3 1255-1314,
// each within run-to-run noise (2-4%) of the best 8-, 16- or 32-plane chunk.
// Grids large enough for every resident workgroup to get kResidentPlanes
// planes take the guided schedule instead (kGuidedIfLong).
//
// How the fused kernel's workgroups split the z-march (TB2Shape::sched):
//   kChunked       chunks of zchunk planes per tile;
//   kGuided(2)     chunks that shrink as the march is dispatched: each the
//                  larger of zchunk planes (8 if 0) and the remaining planes'
//                  share of one (two) waves of the workgroups the device keeps
//                  resident, so the last workgroups are short while most
//                  planes are in long chunks that re-read few halo planes;
//   kGuidedIfLong  kGuided when every resident workgroup's share of the whole
//                  grid is at least kResidentPlanes planes, else chunked.
// MI355X, 512^3 clamped, 200 steps, GLUPS (chunks of 16 / 32 planes / guided
// / guided2): FP32 1299 / 1290 / 1332 / 1294; BF16 1556 / 1527 / 1647 /
// 1553.  Measured and dropped: only the resident workgroups, each taking an
// equal share of the (tile, plane) positions in one or two segments: FP32
// 1207 (its passes fetch 0.787 GB against 0.744 with 16-plane chunks), BF16
// 1628; the same with the largest multiple of the tile count, so tiles split
// into the same segments: FP32 1213 (0.648 GB, but 430 workgroups for 512
// slots), BF16 1453.
enum TB2Sched : int {
  kChunked = 0,
  kGuided = 4,
  kGuided2 = 5,
  kGuidedIfLong = 6,
};
struct TB2Shape {
  int nwx, by, zchunk;
  int sched = kGuidedIfLong;
};
constexpr int kResidentPlanes = 32;
constexpr int kTB2Waves = 4096, kTB2WavesFull = 16384;

inline TB2Shape ChooseTB2Shape(int nx, int ny, int nz) {
  TB2Shape sh;
  sh.nwx = nx <= 256 ? 1 : (nx <= 512 ? 2 : 4);
  sh.by = sh.nwx == 4 ? 4 : 8;
  const int tiles_y = (ny + sh.by - 3) / (sh.by - 2);
  auto waves = [&](int z) { return (int64_t)tiles_y * ((nz + z - 1) / z) * sh.nwx * sh.by; };
  sh.zchunk = 8;
  for (int want : {kTB2WavesFull, kTB2Waves}) {
    for (int z : {32, 16, 8}) {
      if (waves(z) >= want) {
        sh.zchunk = z;
        return sh;
      }
    }
  }
  return sh;
}

template <class Plan, int NWX, int BY, bool kRenamed = false, bool kMid = false>
inline hipError_t LaunchZMarchStarTB2(const typename Plan::Args &a1, const typename Plan::Args &a2, int ny,
                                      int nz, int zchunk, bool remap, hipStream_t stream, int sched = kChunked) {
  const auto kernel = ZMarchStar4TB2<Plan, NWX, BY, kRenamed, kMid>;
  const dim3 block(64 * NWX, BY, 1);
  TBGeometry geo;
  geo.tiles_y = (ny + BY - 3) / (BY - 2);
  geo.remap = remap ? 1 : 0;
  if (sched != kChunked) {
    const int slots = ResidentWorkgroups<TB2KernelId<Plan, NWX, BY, kRenamed, kMid>>(
        reinterpret_cast<const void *>(kernel), 64 * NWX * BY);
    const bool is_long = slots > 0 && (int64_t)geo.tiles_y * nz >= (int64_t)kResidentPlanes * slots;
    if (slots > 0 && (sched != kGuidedIfLong || is_long)) {
      const int k = sched == kGuided2 ? 2 : 1;
      const int64_t lo = zchunk > 0 && sched != kGuidedIfLong ? zchunk : 8;
      int n = 0, z = 0;
      while (z < nz && n < kMaxZChunks) {
        const int64_t share = ((int64_t)(nz - z) * geo.tiles_y + (int64_t)k * slots - 1) / ((int64_t)k * slots);
        int c = (int)(share > lo ? share : lo);
        if (n == kMaxZChunks - 1 || c > nz - z) c = nz - z;
        geo.zs[n++] = z;
        z += c;
      }
      geo.zs[n] = nz;
      geo.zchunk = -1;
      hipLaunchKernelGGL(kernel, dim3(geo.tiles_y, n, 1), block, 0, stream, a1, a2, geo);
      return hipGetLastError();
    }
  }
  if (zchunk <= 0 || zchunk > nz) zchunk = nz;
  geo.zchunk = zchunk;
  dim3 grid(geo.tiles_y, (nz + zchunk - 1) / zchunk, 1);
  hipLaunchKernelGGL(kernel, grid, block, 0, stream, a1, a2, geo);
  return hipGetLastError();
}

// Two fused steps with the shape ChooseTB2Shape picks; hipErrorNotSupported
// when the rows are not whole 16-byte vectors or wider than 1024 points.
// mid: also store the intermediate level through out_ptr(a1) (kMid above).
template <class Plan>
inline hipError_t LaunchZMarchStarTB2Auto(const typename Plan::Args &a1, const typename Plan::Args &a2, int nx,
                                          int ny, int nz, hipStream_t stream, bool mid = false) {
  const TB2Shape sh = ChooseTB2Shape(nx, ny, nz);
#define PSK_TB2_AUTO(NWX, BY)                                                                          \
  if (sh.nwx == NWX && StarTB2Supported<NWX>(nx, ny, nz))                                               \
    return mid ? LaunchZMarchStarTB2<Plan, NWX, BY, true, true>(a1, a2, ny, nz, sh.zchunk, true, stream,     \
                                                                 sh.sched)                                 \
               : LaunchZMarchStarTB2<Plan, NWX, BY, true, false>(a1, a2, ny, nz, sh.zchunk, true, stream,    \
                                                                  sh.sched);
  PSK_TB2_AUTO(1, 8)
  PSK_TB2_AUTO(2, 8)
  PSK_TB2_AUTO(4, 4)
#undef PSK_TB2_AUTO
  return hipErrorNotSupported;
}

// Buffer path of a run of `pairs` fused passes (pairs >= 2) that starts from
// and ends in buffer A, the last pass reading a third buffer X (so that it may
// also store its intermediate level into B): pass j reads path[j] and writes
// path[j + 1], where path[0] = path[pairs] = A and path[pairs - 1] = X; the
// passes before alternate between A and X, with B standing in for A as the
// first destination when that would make a pass write its own input.  No
// pass writes the buffer it reads (workgroups march at different paces), and
// the last pass reads neither A nor B.  FusedPathBuffer returns path[i] as
// 0 (A), 1 (B) or 2 (X).
__host__ __device__ inline int FusedPathBuffer(int i, int pairs) {
  if (i == 0 || i == pairs) return 0;
  if (((pairs - 1 - i) & 1) == 0) return 2;
  return i == 1 ? 1 : 0;
}

// The buffer pass j of such a run writes.
template <class T>
__host__ __device__ inline T *FusedPathDst(int j, int pairs, T *A, T *B, T *X) {
  const int b = FusedPathBuffer(j + 1, pairs);
  return b == 0 ? A : (b == 1 ? B : X);
}

}  // namespace psk
