// skeleton_point_tb.h -- T sweeps per HBM pass for a relaxation whose body
// reads one grid (the staged grid) within radius 1 on every axis, diagonals
// included, and any number of other grids (up to kMaxPointGrids) at the
// point itself only, over the interior domain [1, n-1) of every axis: the
// Himeno Jacobi (himeno19.hip) and the maps of that form physisc-hip
// translates (the reference's jacobi_kernel, himenobmtxpa_physis.c:331-361).
//
// The point grids (Himeno's twelve coefficient grids) are read at the point
// itself, so fusing T sweeps into one z-march cuts HBM traffic by close to T:
// each point-grid plane is loaded once into a per-thread register queue and
// used by all T sweeps as the march passes it.  One workgroup is 64 lanes (x)
// by BY waves (y), one point per thread.  Level 0 is the input field u_k;
// level t = 1..T is u_{k+t}.  Level t of plane z is computed at march step
// s = z + t - 1 from the level t-1 planes z-1, z, z+1 (held in an LDS ring of
// four planes per level, one barrier per level and step), so every level
// trails the one below it by one plane.  The xy region a level is needed on
// shrinks by one point per side and level: a tile writes the inner
// (64 - 2(T-1)) x (BY - 2(T-1)) points of level T, and neighbouring tiles
// overlap (their shared point-grid lines meet in L2).
//
// Only interior points [1, n-1) are relaxed (the domain of the reference
// program, himenobmtxpa_physis.c:372-374).  Sweeps never write the boundary
// shell, so the shell of u_j is the shell of the buffer the j-th sweep writes
// into: buffer A for even j, B for odd j (p0 and p1 of the program).  Every
// level takes its shell values from there, which lets the pass read from and
// write into a third buffer.  Each value is computed by the same FP32
// expression as the one-sweep kernel, so results are bitwise equal to T
// separate sweeps.  A plan with kReduce also returns a per-point value ss
// whose FP64 sum of squares per level is accumulated over the tile's own
// points only (each interior point counted once per sweep).
//
// A Plan provides: NG (point grids), kReduce, a Scalars struct (the map's
// scalar arguments) and
//   template <bool kFma, class G, class PF>
//   __device__ static float update(const G &g, const PF &P, const Scalars &s, float *ss);
// where g(n) is point grid n at the point and P(dx, dy, dz) the staged grid
// at an offset; kFma is true for grids stored in 16 bits (a plan may use it
// to pick an evaluation order).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>
#include <vector>

#include "physis/hip/skeleton.h"
#include "physis/hip/storage.h"
#include "physis/hip/wave_reduce.h"

namespace psk {

constexpr int kMaxPointGrids = 16;

template <class Scalars>
struct PointTBArgs {
  const void *in;        // u_k (interior read; shell taken from shell[k0 & 1])
  void *out;             // u_{k+T}, interior points only
  const void *shell[2];  // buffers holding the shell of even / odd iterates
  const void *c[kMaxPointGrids];  // the point grids
  int dim[3];
  int k0;
  int zchunk;
  int tiles_x;
  Scalars sc;
  double *partials;  // [T][blocks] (kReduce)
  // The previous launch's partials ([prev_levels][prev_blocks]), summed by
  // workgroup 0 into prev_out[0 .. prev_levels) (none when prev_out is NULL).
  const double *prev_partials;
  int prev_levels, prev_blocks;
  double *prev_out;
};

// ... of a BY = NW * R row tile (R independent points per thread keep more
// loads and LDS reads in flight between barriers).  Prefetch distance D:
// coefficient and level-0 planes are loaded D march steps before their
// first use.  The register queues hold Q = T + D planes and the march is
// unrolled Q times, so a plane's registers never move (a register copy of a
// load in flight would wait for it).
// Which prefetch loads are issued without a branch (bit 0: level-0 planes,
// bit 1: coefficients) and whether the shell values of intermediate levels
// are prefetched with the coefficients (bit 2) instead of loaded where a
// boundary point needs them.  Branch-free loads keep the compiler's waits
// from cancelling the prefetch distance, but let it hoist every load of an
// unrolled period, which costs registers.  MI355X, ms per sweep, 257x129x129
// (385^3), shapes degree / rows / z-chunk: BF16 2/8/16 0.0412 -> 0.0294
// with all three, 2/8/32 (385^3) 0.388 -> 0.308; BF16 3/16/32 0.0351 ->
// 0.0343 with bits 0 and 2 (all three spill registers: 0.073); FP32
// 2/8/32 0.0320 -> 0.0304 with bits 0 and 2 (0.0371 with all three).
template <class S, int T>
constexpr int kPointTBBranchFree = sizeof(typename S::stored_t) == 2 && T == 2 ? 7 : 5;

template <class Plan, class S, int T, int NW, int R, int D>
__global__ void __launch_bounds__(64 * NW)
PointTB(PointTBArgs<typename Plan::Scalars> a) {
  constexpr int kMode = kPointTBBranchFree<S, T>;
  using st = typename S::stored_t;
  constexpr int BY = NW * R;
  constexpr int TW = 66, TH = BY + 2, PLANE = TW * TH, NT = 64 * NW;
  constexpr int PER = (PLANE + NT - 1) / NT;
  constexpr int OX = 64 - 2 * (T - 1), OY = BY - 2 * (T - 1);
  constexpr int Q = T + D;
  static_assert(OY >= 1, "tile too short for this temporal degree");
  static_assert(Plan::NG <= kMaxPointGrids, "too many point grids");
  __shared__ float ring[T][4][PLANE];

  const int nx = a.dim[0], ny = a.dim[1], nz = a.dim[2];
  const int b = XcdTile(blockIdx.x, gridDim.x);
  const int tile_x = b % a.tiles_x, tile_y = b / a.tiles_x;
  const int ox = 1 + tile_x * OX, oy = 1 + tile_y * OY;  // first owned point
  const int lane = threadIdx.x, w = threadIdx.y, tid = w * 64 + lane;
  const int x = ox - (T - 1) + lane;
  const int zb = 1 + (int)blockIdx.y * a.zchunk;
  const int ze = min(zb + a.zchunk, nz - 1);
  // Byte offsets are 32-bit (the host keeps grids below 4 GiB on this
  // path), so every load is a uniform base address plus one VGPR offset.
  const uint32_t plane = (uint32_t)(nx * ny) * (uint32_t)sizeof(st);
  const char *in = static_cast<const char *>(a.in);
  char *out = static_cast<char *>(a.out);
  const char *shell0 = static_cast<const char *>(a.shell[a.k0 & 1]);
  auto ld = [](const char *base, uint32_t off) { return S::load(reinterpret_cast<const st *>(base + off)); };

  // Per row r of this thread: tile row ty = w + r*NW, grid row y.
  bool col_in[R], col_int[R], owned[R];
  uint32_t col[R];
  int center[R], ty[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    ty[r] = w + r * NW;
    const int y = oy - (T - 1) + ty[r];
    col_in[r] = x >= 0 && x < nx && y >= 0 && y < ny;
    col_int[r] = x >= 1 && x < nx - 1 && y >= 1 && y < ny - 1;
    col[r] = col_in[r] ? (uint32_t)(y * nx + x) * (uint32_t)sizeof(st) : 0u;
    owned[r] = lane >= T - 1 && lane < 64 - (T - 1) && ty[r] >= T - 1 && ty[r] < BY - (T - 1);
    center[r] = (ty[r] + 1) * TW + lane + 1;
  }

  // Level-0 plane staging: element e of the (66 x BY+2) plane.
  uint32_t goff[PER];
  int gkind[PER];  // 0: outside the grid, 1: shell column, 2: interior column
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int e = tid + j * NT;
    const int gx = ox - T + e % TW, gy = oy - T + e / TW;
    const bool inb = e < PLANE && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
    goff[j] = inb ? (uint32_t)(gy * nx + gx) * (uint32_t)sizeof(st) : 0u;
    gkind[j] = !inb ? 0 : (gx >= 1 && gx < nx - 1 && gy >= 1 && gy < ny - 1) ? 2 : 1;
  }
  // A branch-free prefetch (kMode) loads from an in-bounds address (an
  // element outside the grid or the march reads the grid's first element);
  // what it loads for such an element is masked where it is published, or
  // never used.  The coefficient queue holds widened values (16-bit queues
  // cost more registers than they save); level-0 planes stay in storage
  // form until published.  The planes and values of march step s live in
  // slot (s - s0) mod Q.
  st pre[Q][PER];
  float cq[Q][R][Plan::NG];
  // Shell values of the intermediate levels (t = 1 .. T-1) a step needs, at
  // the grid's boundary points; slot of step s as above.
  float shq[Q][T > 1 ? T - 1 : 1][R];
  auto load0 = [&](st (&dst)[PER], int z) {
    const bool zin = z >= 0 && z < nz, zint = z >= 1 && z < nz - 1;
    const uint32_t zo = (uint32_t)(zin ? z : 0) * plane;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const char *base = gkind[j] == 2 && zint ? in : shell0;
      if constexpr (kMode & 1) {
        dst[j] = GlobalLoad(reinterpret_cast<const st *>(base + (gkind[j] != 0 ? zo + goff[j] : 0u)));
      } else {
        st v = 0;
        if (zin && gkind[j] != 0) v = GlobalLoad(reinterpret_cast<const st *>(base + zo + goff[j]));
        dst[j] = v;
      }
    }
  };
  auto store0 = [&](const st (&src)[PER], int z) {
    const bool zin = z >= 0 && z < nz;
    float *rg = ring[0][(z + 8) & 3];
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (PLANE % NT == 0 || tid + j * NT < PLANE) rg[tid + j * NT] = zin && gkind[j] != 0 ? S::widen(src[j]) : 0.f;
  };
  // Coefficients of plane z, used only at interior points of interior
  // planes.
  auto loadc = [&](float (&dst)[R][Plan::NG], int z) {
    const bool zok = z >= 1 && z < nz - 1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t o = col_int[r] && zok ? col[r] + (uint32_t)z * plane : 0u;
#pragma unroll
      for (int n = 0; n < Plan::NG; ++n) {
        if constexpr (kMode & 2)
          dst[r][n] = S::load(reinterpret_cast<const st *>(static_cast<const char *>(a.c[n]) + o));
        else
          dst[r][n] = col_int[r] && zok ? S::load(reinterpret_cast<const st *>(static_cast<const char *>(a.c[n]) + o)) : 0.f;
      }
    }
  };
  // Shell values level t reads at step s (plane s - t + 1), for t < T.
  auto loadsh = [&](float (&dst)[T > 1 ? T - 1 : 1][R], int s) {
#pragma unroll
    for (int t = 1; t < T; ++t) {
      const int z = s - t + 1;
      const bool zin = z >= 0 && z < nz, zint = z >= 1 && z < nz - 1;
      const char *sh = static_cast<const char *>(a.shell[(a.k0 + t) & 1]);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint32_t o = zin && col_in[r] && !(col_int[r] && zint) ? col[r] + (uint32_t)z * plane : 0u;
        dst[t - 1][r] = S::load(reinterpret_cast<const st *>(sh + o));
      }
    }
  };

  const int s0 = zb - (T - 1);        // first march step (level 1 at plane s0)
  const int send = ze + T - 1;        // one past the last step
  {
    st tmp[PER];
    load0(tmp, s0 - 1);
    store0(tmp, s0 - 1);
    load0(tmp, s0);
    store0(tmp, s0);
  }
  StaticFor([&](auto dc) {
    constexpr int d = decltype(dc)::value;
    load0(pre[(d + 1) % Q], s0 + 1 + d);  // planes s0+1 .. s0+D
    loadc(cq[d % Q], s0 + d);             // planes s0 .. s0+D-1
    if constexpr (T > 1 && (kMode & 4)) loadsh(shq[d % Q], s0 + d);
  }, std::make_integer_sequence<int, D>{});

  double acc[T];
#pragma unroll
  for (int t = 0; t < T; ++t) acc[t] = 0.0;

  __shared__ double red[NT / 64];
  if (blockIdx.x == 0 && blockIdx.y == 0 && a.prev_out) {
    for (int t = 0; t < a.prev_levels; ++t) {
      const double r = block_sum<NT>(a.prev_partials + (int64_t)t * a.prev_blocks, a.prev_blocks, red, tid);
      if (tid == 0) a.prev_out[t] = r;
    }
  }

  // One march step s (J = (s - s0) mod Q).
  auto step = [&](auto jc, const int s) {
    constexpr int J = decltype(jc)::value;
    store0(pre[(J + 1) % Q], s + 1);
    __syncthreads();
    // Past the last step the planes loaded are not used.
    load0(pre[(J + 1 + D) % Q], s + 1 + D);
    loadc(cq[(J + D) % Q], s + D);
    if constexpr (T > 1 && (kMode & 4)) loadsh(shq[(J + D) % Q], s + D);
    StaticFor([&](auto tc) {
      constexpr int t = decltype(tc)::value + 1;
      const int z = s - t + 1;
      // Needed on this level: lanes/rows [t-1, 64-(t-1)) x [t-1, BY-(t-1)),
      // planes [zb - (T-t), ze + (T-t)).
      const bool zneed = z >= zb - (T - t) && z < ze + (T - t) && z >= 0 && z < nz;
      const bool zint = z >= 1 && z < nz - 1;
      const float *r0 = ring[t - 1][(z + 7) & 3], *r1 = ring[t - 1][(z + 8) & 3], *r2 = ring[t - 1][(z + 9) & 3];
      float v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const bool need = zneed && col_in[r] && lane >= t - 1 && lane < 64 - (t - 1) && ty[r] >= t - 1 &&
                          ty[r] < BY - (t - 1);
        v[r] = 0.f;
        if (need) {
          if (col_int[r] && zint) {
            const float *g = cq[(J - t + 1 + Q) % Q][r];
            const int c = center[r];
            auto P = [&](int dx, int dy, int dz) {
              const float *rg = dz < 0 ? r0 : dz > 0 ? r2 : r1;
              return rg[c + dy * TW + dx];
            };
            float ss;
            v[r] = Plan::template update<sizeof(st) == 2>([&](int n) { return g[n]; }, P, a.sc, &ss);
            if constexpr (Plan::kReduce)
              if (owned[r] && z >= zb && z < ze) acc[t - 1] += (double)ss * (double)ss;
            if (t == T && owned[r]) S::store(reinterpret_cast<st *>(out + ((uint32_t)z * plane + col[r])), v[r]);
          } else if constexpr (t < T) {
            if constexpr (kMode & 4)
              v[r] = shq[J][t - 1][r];
            else
              v[r] = ld(static_cast<const char *>(a.shell[(a.k0 + t) & 1]), (uint32_t)z * plane + col[r]);
          }
        }
      }
      if constexpr (t < T) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          // Intermediate levels hold what a single sweep would have stored.
          if constexpr (sizeof(st) == 2) v[r] = S::widen(S::narrow(v[r]));
          ring[t][(z + 8) & 3][center[r]] = v[r];
        }
        __syncthreads();
      }
    }, std::make_integer_sequence<int, T>{});
  };

  // Whole periods of Q steps without a guard, then the remaining steps.
  int s = s0;
  for (; s + Q <= send; s += Q) {
    StaticFor([&](auto jc) { step(jc, s + decltype(jc)::value); }, std::make_integer_sequence<int, Q>{});
  }
  StaticFor([&](auto jc) {
    constexpr int J = decltype(jc)::value;
    if (s + J < send) step(jc, s + J);
  }, std::make_integer_sequence<int, Q>{});

  if constexpr (Plan::kReduce) {
    const int blocks = gridDim.x * gridDim.y;
    const int bid = blockIdx.x + blockIdx.y * gridDim.x;
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const double r = block_reduce<NT>(acc[t], OpSum(), red, tid);
      if (tid == 0) a.partials[(int64_t)t * blocks + bid] = r;
    }
  }
}

// The passes of a run of `sweeps` sweeps with degree T: pass i is
// out[4i .. 4i+3] = {first iterate k, sweeps in the pass, source buffer,
// destination buffer} with buffers 0 = A (p0), 1 = B (p1), 2 = the third
// buffer.  Fused passes (T sweeps) run while at least T + 2 sweeps remain,
// writing to the third buffer when reading A or B, else to the buffer of the
// output's parity; every other pass is one sweep into the buffer of its
// output's parity.  Returns the number of passes, or -1 if max is too small.
inline int TBSchedule(int sweeps, int T, int *out, int max) {
  auto parity = [](int j) { return j & 1; };
  int cur = 0, k = 0, n = 0;
  auto emit = [&](int t, int dst) {
    if (n >= max) return false;
    out[4 * n] = k;
    out[4 * n + 1] = t;
    out[4 * n + 2] = cur;
    out[4 * n + 3] = dst;
    ++n;
    cur = dst;
    k += t;
    return true;
  };
  if (T > 1)
    while (sweeps - k >= T + 2)
      if (!emit(T, cur != 2 ? 2 : parity(k + T))) return -1;
  while (k < sweeps)
    if (!emit(1, parity(k + 1))) return -1;
  return n;
}

// `sweeps` sweeps of a point plan without reduction (kReduce false) from u_0
// in A over the interior of a grid of a.dim, leaving A and B as single
// sweeps leave them (u_n in A for even n, else in B; u_{n-1} in the other):
// passes of two sweeps through the third buffer X while at least four
// sweeps remain (TBSchedule), then one sweep per pass, on 64 x 8 tiles (the
// fused Himeno shape fastest for FP32 storage on 257x129x129 and 384^3).  a.c, a.sc and a.dim are the caller's; the rest is
// filled here.
template <class Plan, class S>
inline hipError_t RunPointTB(PointTBArgs<typename Plan::Scalars> a, void *A, void *B, void *X, int sweeps,
                             hipStream_t stream) {
  static_assert(!Plan::kReduce, "plans with a reduction need the partial sums of their own runner");
  const int ix = a.dim[0] - 2, iy = a.dim[1] - 2, iz = a.dim[2] - 2;
  if (ix <= 0 || iy <= 0 || iz <= 0 || sweeps <= 0) return hipSuccess;
  std::vector<int> passes(4 * ((size_t)sweeps + 1));
  const int n = TBSchedule(sweeps, 2, passes.data(), sweeps + 1);
  if (n < 0) return hipErrorInvalidValue;
  void *buf[3] = {A, B, X};
  a.shell[0] = A;
  a.shell[1] = B;
  a.partials = nullptr;
  a.prev_partials = nullptr;
  a.prev_levels = a.prev_blocks = 0;
  a.prev_out = nullptr;
  for (int i = 0; i < n; ++i) {
    const int T = passes[4 * i + 1];
    a.k0 = passes[4 * i];
    a.in = buf[passes[4 * i + 2]];
    a.out = buf[passes[4 * i + 3]];
    const int ox = 64 - 2 * (T - 1), oy = 8 - 2 * (T - 1);
    a.tiles_x = (ix + ox - 1) / ox;
    const int tiles_y = (iy + oy - 1) / oy;
    // Fused passes take 32-plane chunks; the one-sweep passes at the end the
    // longest of 16, 8, 4, 2 planes that still gives 1024 workgroups.
    int zc = 32;
    if (T == 1)
      for (zc = 16; zc > 2 && a.tiles_x * tiles_y * ((iz + zc - 1) / zc) < 1024; zc /= 2) {
      }
    a.zchunk = iz < zc ? iz : zc;
    const int chunks = (iz + a.zchunk - 1) / a.zchunk;
    const void *k = T == 2 ? reinterpret_cast<const void *>(&PointTB<Plan, S, 2, 8, 1, 1>)
                           : reinterpret_cast<const void *>(&PointTB<Plan, S, 1, 8, 1, 1>);
    void *args[] = {&a};
    const hipError_t e = hipLaunchKernel(k, dim3(a.tiles_x * tiles_y, chunks), dim3(64, 8), args, 0, stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
}  // namespace psk
