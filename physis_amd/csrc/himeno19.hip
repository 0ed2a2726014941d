// himeno19.hip -- the Himeno 19-point Jacobi program, as physisc-hip lowers
// reference examples/himeno/himenobmtxpa_physis.c:331-361 (jacobi_kernel)
// over the interior domain of :372-374, with a fused FP64 epilogue that sums
// ss*ss per sweep (the reference port leaves gosa at 0, :387-392; the
// original accumulates it in float, himenobmtxpa_original.c:334).
//
// Grids stored in 16 bits evaluate the update with fused multiply-adds
// (HimenoUpdate below); FP32 grids evaluate it as written.
//
// p0 is the only grid read at offsets: it is staged in the skeleton's LDS
// ring (radius 1 in x, y and z, including the xy/yz/xz diagonals).  The 12
// coefficient grids are read at the point itself, straight from HBM.
#include "runtime_hip.h"
#include "physis/hip/skeleton.h"
#include "physis/hip/skeleton_point_tb.h"
#include "physis/hip/storage.h"

#include <algorithm>
#include <vector>
#include <type_traits>
#include <utility>

namespace physis_hip {
namespace {

enum { A0, A1, A2, A3, B0, B1, B2, C0, C1, C2, BND, WRK1, NCOEF };

// One point of jacobi_kernel (himenobmtxpa_physis.c:331-361): G(n) is
// coefficient grid n at the point, P(dx, dy, dz) the pressure at an offset;
// returns the new pressure and leaves ss in *ss.  kFma: the same terms in
// the same order with each product after the first fused into the running
// sum (and omega*ss into the update), used for grids stored in 16 bits,
// which are judged against FP64 within the storage's rounding; FP32 grids
// evaluate the expression as written, as the sequential reference does.
template <bool kFma, class G, class PF>
__device__ __forceinline__ float HimenoUpdate(const G &g, const PF &P, float omega, float *ss_out) {
  float s0, ss;
  if constexpr (kFma) {
    s0 = g(A0) * P(0, 0, 1);
    s0 = __builtin_fmaf(g(A1), P(0, 1, 0), s0);
    s0 = __builtin_fmaf(g(A2), P(1, 0, 0), s0);
    s0 = __builtin_fmaf(g(B0), P(0, 1, 1) - P(0, -1, 1) - P(0, 1, -1) + P(0, -1, -1), s0);
    s0 = __builtin_fmaf(g(B1), P(1, 1, 0) - P(1, -1, 0) - P(-1, 1, 0) + P(-1, -1, 0), s0);
    s0 = __builtin_fmaf(g(B2), P(1, 0, 1) - P(1, 0, -1) - P(-1, 0, 1) + P(-1, 0, -1), s0);
    s0 = __builtin_fmaf(g(C0), P(0, 0, -1), s0);
    s0 = __builtin_fmaf(g(C1), P(0, -1, 0), s0);
    s0 = __builtin_fmaf(g(C2), P(-1, 0, 0), s0);
    s0 = s0 + g(WRK1);
    ss = __builtin_fmaf(s0, g(A3), -P(0, 0, 0)) * g(BND);
    *ss_out = ss;
    return __builtin_fmaf(omega, ss, P(0, 0, 0));
  } else {
    s0 = g(A0) * P(0, 0, 1) + g(A1) * P(0, 1, 0) + g(A2) * P(1, 0, 0) +
         g(B0) * (P(0, 1, 1) - P(0, -1, 1) - P(0, 1, -1) + P(0, -1, -1)) +
         g(B1) * (P(1, 1, 0) - P(1, -1, 0) - P(-1, 1, 0) + P(-1, -1, 0)) +
         g(B2) * (P(1, 0, 1) - P(1, 0, -1) - P(-1, 0, 1) + P(-1, 0, -1)) + g(C0) * P(0, 0, -1) +
         g(C1) * P(0, -1, 0) + g(C2) * P(-1, 0, 0) + g(WRK1);
    ss = (s0 * g(A3) - P(0, 0, 0)) * g(BND);
    *ss_out = ss;
    return P(0, 0, 0) + omega * ss;
  }
}

// kOff32: every grid is below 4 GiB, so coefficient addresses are a uniform
// base plus a 32-bit byte offset (one VGPR per load instead of a 64-bit add).
template <class S, bool kOff32>
struct HimenoPlan {
  using SV = float;
  using Storage = S;
  using st = typename S::stored_t;
  static constexpr int NS = 1, RX = 1, RY = 1, RZ = 1;
  static constexpr bool kWrap = false;
  static constexpr bool kReduce = true;
  struct Args {
    const st *p;
    st *q;
    const st *c[NCOEF];
    int dim[3];
    float omega;
    double *partials;
    // The previous sweep's partials, summed by workgroup 0 into *prev_out
    // (none when prev_out is NULL).
    const double *prev_partials;
    int prev_blocks;
    double *prev_out;
  };
  template <int NT>
  __device__ static void fold_previous(const Args &a, double *lds, int tid) {
    if (!a.prev_out) return;
    const double r = psk::block_sum<NT>(a.prev_partials, a.prev_blocks, lds, tid);
    if (tid == 0) *a.prev_out = r;
  }
  __device__ static void stage_dims(const Args &a, int &nx, int &ny, int &nz) {
    nx = a.dim[0];
    ny = a.dim[1];
    nz = a.dim[2];
  }
  __device__ static SV stage_load(const Args &a, int, int64_t off) { return S::load(a.p + off); }
  __host__ __device__ static const st *stage_ptr(const Args &a) { return a.p; }
  template <class C>
  __device__ static double point(const C &P, const Args &a, const int i, const int j, const int k) {
    const int64_t o = i + (int64_t)j * a.dim[0] + (int64_t)k * a.dim[0] * a.dim[1];
    const uint32_t ob = (uint32_t)(i + j * a.dim[0] + k * a.dim[0] * a.dim[1]) * (uint32_t)sizeof(st);
    auto g = [&](int n) {
      if constexpr (kOff32)
        return S::load(reinterpret_cast<const st *>(reinterpret_cast<const char *>(a.c[n]) + ob));
      else
        return S::load(a.c[n] + o);
    };
    float ss;
    const float v = HimenoUpdate<sizeof(st) == 2>(
        g, [&](int dx, int dy, int dz) { return P.at(0, dx, dy, dz); }, a.omega, &ss);
    if constexpr (kOff32)
      S::store(reinterpret_cast<st *>(reinterpret_cast<char *>(a.q) + ob), v);
    else
      S::store(a.q + o, v);
    return (double)ss * (double)ss;
  }
};

// ---------------------------------------------------------------------------
// T sweeps per HBM pass (temporal blocking): psk::PointTB
// (physis/hip/skeleton_point_tb.h) with the twelve coefficient grids as its
// point grids and the FP64 sum of ss^2 of every level.
struct HimenoTBPlan {
  static constexpr int NG = NCOEF;
  static constexpr bool kReduce = true;
  struct Scalars {
    float omega;
  };
  template <bool kFma, class G, class PF>
  __device__ static float update(const G &g, const PF &P, const Scalars &s, float *ss) {
    return HimenoUpdate<kFma>(g, P, s.omega, ss);
  }
};
using TBArgs = psk::PointTBArgs<HimenoTBPlan::Scalars>;

// Sums the per-workgroup partials of one sweep in index order.
__global__ void __launch_bounds__(256) SumPartials(const double *partials, int n, double *out) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += partials[i];
  const double r = psk::block_reduce<256>(acc, psk::OpSum(), red, threadIdx.x);
  if (threadIdx.x == 0) *out = r;
}

// One workgroup per level: out[t] = sum of partials[t*n .. t*n + n).
__global__ void __launch_bounds__(256) SumPartialsLevels(const double *partials, int n, double *out) {
  __shared__ double red[4];
  const double *p = partials + (int64_t)blockIdx.x * n;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += p[i];
  const double r = psk::block_reduce<256>(acc, psk::OpSum(), red, threadIdx.x);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

struct Variant {
  const char *name;
  int bx, by, zchunk;
  int rows = 1;  // tile rows per thread
};
// Two or more rows per thread (which the generic skeleton supports) were
// 4-9x slower for this plan on MI355X: its twelve coefficient loads per point
// are issued in the point body, and R points' worth of them exhaust the
// registers.
const Variant kVariants[] = {
    {"skeleton64x4_z32", 64, 4, 32},
    {"skeleton64x8_z32", 64, 8, 32},
    {"skeleton64x4_zall", 64, 4, 0},
    {"skeleton64x4_z16", 64, 4, 16},
    {"skeleton64x4_z64", 64, 4, 64},
    {"skeleton64x2_z32", 64, 2, 32},
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

psk::Box InteriorBox(const __PSGrid *p0) {
  psk::Box box;
  for (int d = 0; d < 3; ++d) {
    box.lo[d] = 1;
    box.hi[d] = p0->dim[d] - 1;
  }
  return box;
}

int Blocks(const Variant &v, const psk::Box &box) {
  const int tile_rows = v.by * v.rows;
  const int ex = box.hi[0] - box.lo[0], ey = box.hi[1] - box.lo[1], ez = box.hi[2] - box.lo[2];
  if (ex <= 0 || ey <= 0 || ez <= 0) return 0;
  const int zc = v.zchunk <= 0 || v.zchunk > ez ? ez : v.zchunk;
  return ((ex + v.bx - 1) / v.bx) * ((ey + tile_rows - 1) / tile_rows) * ((ez + zc - 1) / zc);
}

template <class P>
hipError_t LaunchVariant(const Variant &v, const typename P::Args &a, const psk::Box &box, hipStream_t s) {
#define PSH_RING(BX, BY, R) \
  if (v.bx == BX && v.by == BY && v.rows == R) return psk::LaunchZMarch<P, BX, BY, R>(a, box, v.zchunk, true, s);
  PSH_RING(64, 4, 1) PSH_RING(64, 8, 1) PSH_RING(64, 2, 1) PSH_RING(128, 2, 1) PSH_RING(128, 4, 1) PSH_RING(256, 1, 1)
  PSH_RING(256, 2, 1) PSH_RING(256, 1, 2) PSH_RING(128, 2, 2) PSH_RING(64, 4, 2)
#undef PSH_RING
  return hipErrorInvalidValue;
}

// Queues `sweeps` sweeps p0 -> p1 -> p0 ... on stream s.  gosa_dev (may be
// NULL) receives the FP64 sum of ss^2 of each sweep; partials needs
// Blocks(v, box) doubles.
template <class S, bool kOff32>
int EnqueuePlan(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids, float omega, int sweeps,
                double *gosa_dev, double *partials, const Variant &v, hipStream_t s) {
  using P = HimenoPlan<S, kOff32>;
  using st = typename S::stored_t;
  const psk::Box box = InteriorBox(p0);
  const int nblocks = Blocks(v, box);
  typename P::Args a;
  for (int n = 0; n < NCOEF; ++n) a.c[n] = static_cast<const st *>(grids[n]->p0);
  for (int d = 0; d < 3; ++d) a.dim[d] = p0->dim[d];
  a.omega = omega;
  a.prev_blocks = nblocks;
  // Partials alternate between two halves of the buffer; each sweep's
  // workgroup 0 sums the previous sweep's half, and one kernel sums the last.
  for (int i = 0; i < sweeps; ++i) {
    a.p = static_cast<const st *>((i & 1) ? p1->p0 : p0->p0);
    a.q = static_cast<st *>((i & 1) ? p0->p0 : p1->p0);
    a.partials = partials + (size_t)(i & 1) * nblocks;
    a.prev_partials = partials + (size_t)((i + 1) & 1) * nblocks;
    a.prev_out = gosa_dev && i > 0 ? gosa_dev + i - 1 : nullptr;
    hipError_t e = LaunchVariant<P>(v, a, box, s);
    if (e != hipSuccess) return PS_HIP_ERR_DEVICE;
  }
  if (gosa_dev && nblocks > 0 && sweeps > 0) {
    hipLaunchKernelGGL(SumPartials, dim3(1), dim3(256), 0, s, partials + (size_t)((sweeps - 1) & 1) * nblocks, nblocks,
                       gosa_dev + sweeps - 1);
    if (hipGetLastError() != hipSuccess) return PS_HIP_ERR_DEVICE;
  }
  return PS_HIP_OK;
}

// The 14 grids of a run have one shape and storage, so p0's size decides.
template <class S>
int EnqueueTyped(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids, float omega, int sweeps,
                 double *gosa_dev, double *partials, const Variant &v, hipStream_t s) {
  if (Rec(p0)->bytes <= 0xffffffffull)
    return EnqueuePlan<S, true>(p0, p1, grids, omega, sweeps, gosa_dev, partials, v, s);
  return EnqueuePlan<S, false>(p0, p1, grids, omega, sweeps, gosa_dev, partials, v, s);
}

template <class S>
int RunTyped(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids, float omega, int sweeps,
             double *gosa, const Variant &v) {
  Runtime &rt = Runtime::Get();
  hipStream_t s = rt.stream();
  const int nblocks = Blocks(v, InteriorBox(p0));
  const size_t nres = sweeps > 0 ? (size_t)sweeps : 1;
  double *scratch = static_cast<double *>(rt.Scratch((nres + 2 * (size_t)nblocks + 1) * sizeof(double)));
  double *gosa_dev = scratch;
  double *partials = scratch + nres;
  int st = EnqueueTyped<S>(p0, p1, grids, omega, sweeps, gosa ? gosa_dev : nullptr, partials, v, s);
  if (st != PS_HIP_OK) return st;
  if (gosa && sweeps > 0) {
    if (nblocks == 0) {
      for (int i = 0; i < sweeps; ++i) gosa[i] = 0.0;
      return PS_HIP_OK;
    }
    if (hipMemcpyAsync(gosa, gosa_dev, sweeps * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
      return PS_HIP_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return PS_HIP_ERR_DEVICE;
  }
  return PS_HIP_OK;
}

// ---- temporal-blocking launches -------------------------------------------

// Launch shape of the fused kernel: T sweeps per pass, tiles of 64 x by
// points, rpw rows per wave (by / rpw waves per workgroup), zchunk output
// planes per workgroup (0: all), prefetch distance pf in march steps (0:
// the shape's default).
struct TBConfig {
  int T, by, rpw, zchunk, pf = 0;
};

struct TBGrid {
  int tiles_x = 0, tiles_y = 0, chunks = 0;
  int Blocks() const { return tiles_x * tiles_y * chunks; }
};

TBGrid TBGridFor(const int *dim, const TBConfig &c) {
  TBGrid g;
  const int ix = dim[0] - 2, iy = dim[1] - 2, iz = dim[2] - 2;
  if (ix <= 0 || iy <= 0 || iz <= 0) return g;
  const int ox = 64 - 2 * (c.T - 1), oy = c.by - 2 * (c.T - 1);
  const int zc = c.zchunk <= 0 || c.zchunk > iz ? iz : c.zchunk;
  g.tiles_x = (ix + ox - 1) / ox;
  g.tiles_y = (iy + oy - 1) / oy;
  g.chunks = (iz + zc - 1) / zc;
  return g;
}

// Compiled shapes (T, waves, rows per wave); the one-sweep shapes of a
// tile (T = 1) exist for every tile the fused shapes use.
template <class S>
const void *TBKernel(int T, int nw, int r, int pf) {
  if (pf == 0) pf = 1;
#define PSH_TB(TT, NW, RR, DD) \
  if (T == TT && nw == NW && r == RR && pf == DD)                                                        \
    return reinterpret_cast<const void *>(&psk::PointTB<HimenoTBPlan, S, TT, NW, RR, DD>);
  PSH_TB(1, 8, 1, 1) PSH_TB(1, 8, 1, 2) PSH_TB(1, 8, 1, 3) PSH_TB(1, 8, 1, 4) PSH_TB(1, 4, 1, 1) PSH_TB(1, 4, 1, 2)
  PSH_TB(1, 4, 2, 1) PSH_TB(1, 8, 2, 1) PSH_TB(1, 16, 1, 1)
  PSH_TB(2, 8, 1, 1) PSH_TB(2, 8, 1, 2) PSH_TB(2, 8, 1, 3) PSH_TB(2, 4, 2, 1) PSH_TB(2, 8, 2, 1) PSH_TB(2, 16, 1, 1)
  PSH_TB(3, 8, 1, 1) PSH_TB(3, 4, 2, 1) PSH_TB(3, 8, 2, 1) PSH_TB(3, 16, 1, 1) PSH_TB(3, 16, 1, 2)
  PSH_TB(4, 8, 2, 1) PSH_TB(4, 16, 1, 1)
#undef PSH_TB
  return nullptr;
}

const void *TBKernelFor(int storage, int T, int by, int rpw, int pf) {
  if (rpw <= 0 || by % rpw != 0) return nullptr;
  const int nw = by / rpw;
  switch (storage) {
    case PS_HIP_STORAGE_NATIVE: return TBKernel<psk::Native<float>>(T, nw, rpw, pf);
    case PS_HIP_STORAGE_BF16: return TBKernel<psk::BF16>(T, nw, rpw, pf);
    case PS_HIP_STORAGE_FP16: return TBKernel<psk::FP16>(T, nw, rpw, pf);
  }
  return nullptr;
}

// The fused shape and the one-sweep shape of the same tile (for the last
// sweeps of a run) must both be compiled.
bool TBSupported(const TBConfig &c) {
  return c.zchunk >= 0 && c.T >= 1 && TBKernelFor(PS_HIP_STORAGE_NATIVE, c.T, c.by, c.rpw, c.pf) != nullptr &&
         TBKernelFor(PS_HIP_STORAGE_NATIVE, 1, c.by, c.rpw, 0) != nullptr;
}

// Scratch doubles a run with configuration c needs for its partials: two
// halves, each large enough for one launch.
int TBHalf(const __PSGrid *p0, const TBConfig &c) {
  const TBConfig one{1, c.by, c.rpw, c.zchunk};
  return std::max(c.T * TBGridFor(p0->dim, c).Blocks(), TBGridFor(p0->dim, one).Blocks());
}
int TBPartials(const __PSGrid *p0, const TBConfig &c) { return 2 * TBHalf(p0, c); }

// The pass schedule of a run: psk::TBSchedule.
using psk::TBSchedule;

// Partials of the launch before the current one, still to be summed.
struct Pending {
  const double *partials = nullptr;
  int levels = 0, blocks = 0;
  double *out = nullptr;
};

// One pass of T sweeps: u_{k0+T} = F^T(in) into out (interior only).  Its
// workgroup 0 sums the pending partials of the previous pass; this pass's
// own partials (in `partials`) become pending.
int TBPass(int storage, const TBConfig &c, int T, const void *in, void *out, __PSGrid *A, __PSGrid *B,
           __PSGrid *const *grids, float omega, int k0, double *gosa_dev, double *partials, Pending &pend,
           hipStream_t s) {
  const TBConfig cc{T, c.by, c.rpw, c.zchunk};
  const TBGrid g = TBGridFor(A->dim, cc);
  if (g.Blocks() == 0) return PS_HIP_OK;
  const void *k = TBKernelFor(storage, T, c.by, c.rpw, T == c.T ? c.pf : 0);
  if (!k) return PS_HIP_ERR_UNSUPPORTED;
  TBArgs a;
  a.in = in;
  a.out = out;
  a.shell[0] = A->p0;
  a.shell[1] = B->p0;
  for (int n = 0; n < NCOEF; ++n) a.c[n] = grids[n]->p0;
  for (int d = 0; d < 3; ++d) a.dim[d] = A->dim[d];
  a.k0 = k0;
  const int iz = A->dim[2] - 2;
  a.zchunk = c.zchunk <= 0 || c.zchunk > iz ? iz : c.zchunk;
  a.tiles_x = g.tiles_x;
  a.sc.omega = omega;
  a.partials = partials;
  a.prev_partials = pend.partials;
  a.prev_levels = pend.levels;
  a.prev_blocks = pend.blocks;
  a.prev_out = pend.out;
  void *args[] = {&a};
  if (hipLaunchKernel(k, dim3(g.tiles_x * g.tiles_y, g.chunks), dim3(64, c.by / c.rpw), args, 0, s) != hipSuccess)
    return PS_HIP_ERR_DEVICE;
  pend = Pending();
  if (gosa_dev) {
    pend.partials = partials;
    pend.levels = T;
    pend.blocks = g.Blocks();
    pend.out = gosa_dev + k0;
  }
  return PS_HIP_OK;
}

// `sweeps` sweeps starting from u_0 in A, leaving the grids as the
// one-sweep sequence leaves them (u_n in A for even n, else in B; u_{n-1} in
// the other).  Passes of c.T sweeps run while at least c.T + 2 sweeps remain
// (input and output alternate between the third buffer X and the buffer of
// the output's parity); the rest run one sweep per pass into the buffer of
// their parity, so the last two always land in A and B.  gosa_dev[i] (may
// be NULL) receives sweep i's FP64 sum of ss^2.  The schedule itself is
// TBSchedule (host code, exported for tests as PSHipHimenoTBSchedule).
int TBRun(const TBConfig &c, __PSGrid *A, __PSGrid *B, void *X, __PSGrid *const *grids, float omega, int sweeps,
          double *gosa_dev, double *partials, hipStream_t s) {
  const int storage = Rec(A)->storage;
  if (Rec(A)->bytes > 0xffffffffull) return PS_HIP_ERR_UNSUPPORTED;  // 32-bit byte offsets
  if (c.T > 1 && !X) return PS_HIP_ERR_INVALID;
  void *buf[3] = {A->p0, B->p0, X};
  const int half = TBHalf(A, c);
  Pending pend;
  int launches = 0;
  std::vector<int> plan(4 * (size_t)(sweeps + 1));
  const int npass = TBSchedule(sweeps, c.T, plan.data(), (int)(plan.size() / 4));
  if (npass < 0) return PS_HIP_ERR_INVALID;
  for (int i = 0; i < npass; ++i) {
    const int k = plan[4 * i], T = plan[4 * i + 1], src = plan[4 * i + 2], dst = plan[4 * i + 3];
    double *part = partials + (size_t)(launches++ & 1) * half;
    const int st = TBPass(storage, c, T, buf[src], buf[dst], A, B, grids, omega, k, gosa_dev, part, pend, s);
    if (st != PS_HIP_OK) return st;
  }
  if (pend.out) {
    hipLaunchKernelGGL(SumPartialsLevels, dim3(pend.levels), dim3(256), 0, s, pend.partials, pend.blocks, pend.out);
    if (hipGetLastError() != hipSuccess) return PS_HIP_ERR_DEVICE;
  }
  return PS_HIP_OK;
}

// One sweep per launch on the fused kernel's shape (T = 1, coefficient
// loads issued a plane ahead): variant 4 of PSHipHimenoRun.
const TBConfig kSweepShape{1, 8, 1, 16};

// The default one-sweep launch: a tile of 256 points whose width is the
// widest of 256, 128 and 64 that covers the interior row with at most 10%
// of its lanes idle (wider tiles read longer runs of each of the 14 rows
// per step), and the largest z-chunk from the storage's list that still
// gives kTileWorkgroups workgroups (the smallest chunk when none does).
// Timed on MI355X over 48^3 .. 513x257x257 against 35 tiles (64 x 2/4/8,
// 128 x 2/4, 256 x 1/2, chunks 2..32): FP32 is fastest at 2-4 planes,
// 16-bit storage at 8-16; the rule is within 5% of the best tile on every
// grid, 4-15% faster than 64 x 4 tiles on 256^3-513x257x257 and 1.05-4x
// faster than 32-plane chunks on grids below 257x129x129.
constexpr int kTileWorkgroups = 1024;

Variant AutoTile(const __PSGrid *p0) {
  static const int kChunksFP32[] = {4, 2};
  static const int kChunks16[] = {16, 8, 4, 2};
  const bool wide = Rec(p0)->storage == PS_HIP_STORAGE_NATIVE;
  const int *chunks = wide ? kChunksFP32 : kChunks16;
  const int n = wide ? 2 : 4;
  const psk::Box box = InteriorBox(p0);
  const int ex = std::max(box.hi[0] - box.lo[0], 1);
  int bx = 64;
  for (int w : {256, 128}) {
    const int covered = (ex + w - 1) / w * w;
    if (10 * ex >= 9 * covered) {
      bx = w;
      break;
    }
  }
  const int by = 256 / bx;
  for (int i = 0; i < n; ++i) {
    const Variant c{"skeleton_auto", bx, by, chunks[i]};
    if (Blocks(c, box) >= kTileWorkgroups) return c;
  }
  return Variant{"skeleton_auto", bx, by, chunks[n - 1]};
}

}  // namespace

int HimenoPartials(const __PSGrid *p0) { return 2 * Blocks(AutoTile(p0), InteriorBox(p0)); }

int HimenoEnqueue(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids, float omega, int sweeps,
                  double *gosa_dev, double *partials, hipStream_t s) {
  const Variant v = AutoTile(p0);
  switch (Rec(p0)->storage) {
    case PS_HIP_STORAGE_NATIVE:
      return EnqueueTyped<psk::Native<float>>(p0, p1, grids, omega, sweeps, gosa_dev, partials, v, s);
    case PS_HIP_STORAGE_BF16:
      return EnqueueTyped<psk::BF16>(p0, p1, grids, omega, sweeps, gosa_dev, partials, v, s);
    case PS_HIP_STORAGE_FP16:
      return EnqueueTyped<psk::FP16>(p0, p1, grids, omega, sweeps, gosa_dev, partials, v, s);
  }
  return PS_HIP_ERR_INVALID;
}

int HimenoTBPartials(const __PSGrid *p0, int degree, int by, int rpw, int zchunk) {
  return TBPartials(p0, TBConfig{degree, by, rpw, zchunk});
}

int HimenoEnqueueTB(__PSGrid *p0, __PSGrid *p1, void *aux, __PSGrid *const *grids, float omega, int sweeps,
                    int degree, int by, int rpw, int zchunk, double *gosa_dev, double *partials, hipStream_t s) {
  const TBConfig c{degree, by, rpw, zchunk};
  if (!TBSupported(c)) return PS_HIP_ERR_UNSUPPORTED;
  return TBRun(c, p0, p1, aux, grids, omega, sweeps, gosa_dev, partials, s);
}

}  // namespace physis_hip

using namespace physis_hip;

namespace {

bool SameGridSet(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids) {
  if (!p0 || !p1 || !grids) return false;
  const int storage = Rec(p0)->storage;
  auto same = [&](const __PSGrid *g) {
    if (!g || g->num_dims != 3 || g->elm_size != 4 || !g->p0 || Rec(g)->storage != storage) return false;
    for (int d = 0; d < 3; ++d)
      if (g->dim[d] != p0->dim[d]) return false;
    return true;
  };
  if (!same(p0) || !same(p1)) return false;
  for (int n = 0; n < NCOEF; ++n)
    if (!same(grids[n])) return false;
  return true;
}

}  // namespace

// The pass schedule PSHipHimenoRunTB follows for `sweeps` sweeps of degree
// `degree` (see TBSchedule); returns the number of passes or -1.
extern "C" int PSHipHimenoTBSchedule(int sweeps, int degree, int *passes, int max_passes) {
  if (sweeps < 0 || degree < 1 || !passes || max_passes < 0) return -1;
  return TBSchedule(sweeps, degree, passes, max_passes);
}

// Himeno sweeps with `degree` sweeps fused per HBM pass (1..4), tiles of
// 64 x `by` points (8 or 16), `rows_per_wave` tile rows per wave (1 or 2)
// and `zchunk` output planes per workgroup (0: all).  Same contract as PSHipHimenoRun: after the
// run p0 and p1 hold exactly what `sweeps` single sweeps leave in them, and
// gosa (may be NULL) receives the FP64 sum of ss^2 of every sweep.  Degree
// > 1 uses the runtime's third grid buffer.
extern "C" int PSHipHimenoRunTB(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids, float omega, int sweeps,
                                double *gosa, int degree, int by, int rows_per_wave, int zchunk,
                                int prefetch) {
  if (!SameGridSet(p0, p1, grids) || sweeps < 0) return PS_HIP_ERR_INVALID;
  const TBConfig c{degree, by, rows_per_wave, zchunk, prefetch};
  if (!TBSupported(c)) return PS_HIP_ERR_UNSUPPORTED;
  Runtime &rt = Runtime::Get();
  hipStream_t s = rt.stream();
  void *aux = nullptr;
  if (degree > 1) {
    aux = rt.Aux(Rec(p0)->bytes);
    if (!aux) return PS_HIP_ERR_NOMEM;
  }
  const size_t nres = sweeps > 0 ? (size_t)sweeps : 1;
  const size_t npart = (size_t)std::max(TBPartials(p0, c), 1);
  double *scratch = static_cast<double *>(rt.Scratch((nres + npart) * sizeof(double)));
  double *gosa_dev = scratch, *partials = scratch + nres;
  if (gosa && sweeps > 0 && TBGridFor(p0->dim, c).Blocks() == 0) {
    for (int i = 0; i < sweeps; ++i) gosa[i] = 0.0;
    gosa = nullptr;
  }
  int st = TBRun(c, p0, p1, aux, grids, omega, sweeps, gosa ? gosa_dev : nullptr, partials, s);
  if (st != PS_HIP_OK) return st;
  if (gosa && sweeps > 0) {
    if (hipMemcpyAsync(gosa, gosa_dev, sweeps * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
      return PS_HIP_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return PS_HIP_ERR_DEVICE;
  }
  return PS_HIP_OK;
}

// One sweep per launch on the generic skeleton with bx x by threads, each
// computing `rows` tile rows (tiles of bx x by*rows points): 64 x 2/4/8,
// 128 x 2/4, 256 x 1/2 with one row, 64 x 4, 128 x 2, 256 x 1 with two;
// `zchunk` planes per workgroup (0: all).
extern "C" int PSHipHimenoRunTile(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids, float omega, int sweeps,
                                  double *gosa, int bx, int by, int rows, int zchunk) {
  if (!SameGridSet(p0, p1, grids) || sweeps < 0 || zchunk < 0) return PS_HIP_ERR_INVALID;
  const bool one = (bx == 64 && (by == 2 || by == 4 || by == 8)) || (bx == 128 && (by == 2 || by == 4)) ||
                   (bx == 256 && (by == 1 || by == 2));
  const bool two = (bx == 64 && by == 4) || (bx == 128 && by == 2) || (bx == 256 && by == 1);
  if (!((rows == 1 && one) || (rows == 2 && two))) return PS_HIP_ERR_UNSUPPORTED;
  Variant v{"tile", bx, by, zchunk};
  v.rows = rows;
  switch (Rec(p0)->storage) {
    case PS_HIP_STORAGE_NATIVE: return RunTyped<psk::Native<float>>(p0, p1, grids, omega, sweeps, gosa, v);
    case PS_HIP_STORAGE_BF16: return RunTyped<psk::BF16>(p0, p1, grids, omega, sweeps, gosa, v);
    case PS_HIP_STORAGE_FP16: return RunTyped<psk::FP16>(p0, p1, grids, omega, sweeps, gosa, v);
  }
  return PS_HIP_ERR_INVALID;
}

extern "C" int PSHipHimenoRun(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids, float omega,
                              int sweeps, double *gosa, int variant) {
  if (!p0 || !p1 || !grids || sweeps < 0) return PS_HIP_ERR_INVALID;
  // Ids 1-3 are kVariants[0..2], id 4 the fused kernel's one-sweep shape,
  // ids 5.. the remaining kVariants; variant 0 is the tile of AutoTile.
  constexpr int kSweepId = 4;
  if (variant == 0) {
    if (!SameGridSet(p0, p1, grids) || sweeps < 0) return PS_HIP_ERR_INVALID;
    const Variant v = AutoTile(p0);
    switch (Rec(p0)->storage) {
      case PS_HIP_STORAGE_NATIVE: return RunTyped<psk::Native<float>>(p0, p1, grids, omega, sweeps, gosa, v);
      case PS_HIP_STORAGE_BF16: return RunTyped<psk::BF16>(p0, p1, grids, omega, sweeps, gosa, v);
      case PS_HIP_STORAGE_FP16: return RunTyped<psk::FP16>(p0, p1, grids, omega, sweeps, gosa, v);
    }
    return PS_HIP_ERR_INVALID;
  }
  if (variant == kSweepId)
    return PSHipHimenoRunTB(p0, p1, grids, omega, sweeps, gosa, kSweepShape.T, kSweepShape.by, kSweepShape.rpw,
                            kSweepShape.zchunk, 0);
  if (variant < 1 || variant > kNumVariants + 1) return PS_HIP_ERR_INVALID;
  const int storage = Rec(p0)->storage;
  auto same = [&](const __PSGrid *g) {
    if (!g || g->num_dims != 3 || g->elm_size != 4 || !g->p0 || Rec(g)->storage != storage) return false;
    for (int d = 0; d < 3; ++d)
      if (g->dim[d] != p0->dim[d]) return false;
    return true;
  };
  if (!same(p0) || !same(p1)) return PS_HIP_ERR_INVALID;
  for (int n = 0; n < NCOEF; ++n)
    if (!same(grids[n])) return PS_HIP_ERR_INVALID;
  const Variant &v = kVariants[variant < kSweepId ? variant - 1 : variant - 2];
  switch (storage) {
    case PS_HIP_STORAGE_NATIVE: return RunTyped<psk::Native<float>>(p0, p1, grids, omega, sweeps, gosa, v);
    case PS_HIP_STORAGE_BF16: return RunTyped<psk::BF16>(p0, p1, grids, omega, sweeps, gosa, v);
    case PS_HIP_STORAGE_FP16: return RunTyped<psk::FP16>(p0, p1, grids, omega, sweeps, gosa, v);
  }
  return PS_HIP_ERR_INVALID;
}
