// diffusion7.hip -- the 7-point diffusion program, as physisc-hip lowers
// reference examples/diffusion-benchmark/diffusion3d_physis.c:29-76
// (clamped form) and tests/system_tests/test_cases/test_7-pt-periodic.c
// (periodic form, PSGridGetPeriodic) with coefficients, plus hand-tuned
// kernels for the same update.  Every variant evaluates
//     cc*c + cw*w + ce*e + cs*s + cn*n + cb*b + ct*t
// left to right in FP32; on FP32 grids with no contraction into FMA, exactly
// as the sequential ref target does, so results are bitwise equal to it; on
// 16-bit grids with the products after the first fused into the running sum
// (Diffusion7Body's kFma), the same in every variant.
#include "runtime_hip.h"
#include "diffusion7_body.h"
#include "physis/hip/skeleton.h"
#include "physis/hip/skeleton_star.h"
#include "physis/hip/skeleton_pack.h"
#include "physis/hip/skeleton_star_tb.h"
#include "physis/hip/skeleton_stride.h"
#include "physis/hip/storage.h"

#include <cstdio>

namespace physis_hip {
namespace {

// ---- variant family 1: the generic skeleton with an emitted body ---------
// Plan for map(kernel, dom, g_in, g_out, ce, cw, cn, cs, ct, cb, cc).
template <class S, bool kPeriodic>
struct DiffusionPlan {
  using SV = float;
  static constexpr int NS = 1, RX = 1, RY = 1, RZ = 1;
  static constexpr bool kWrap = kPeriodic;
  static constexpr bool kReduce = false;
  struct Args {
    const typename S::stored_t *g1;
    typename S::stored_t *g2;
    int dim[3];
    Coef7 k;
  };
  __device__ static void stage_dims(const Args &a, int &nx, int &ny, int &nz) {
    nx = a.dim[0];
    ny = a.dim[1];
    nz = a.dim[2];
  }
  __device__ static SV stage_load(const Args &a, int, int64_t off) { return S::load(a.g1 + off); }
  // Star-skeleton interface: single input grid, one emitted value per point.
  using Storage = S;
  __device__ static const typename S::stored_t *stage_ptr(const Args &a) { return a.g1; }
  __device__ static typename S::stored_t *out_ptr(const Args &a) { return a.g2; }
  template <class C>
  __device__ static void point(const C &ctx, const Args &a, const int x, const int y, const int z) {
    const int nx = a.dim[0], ny = a.dim[1];
    S::store(a.g2 + (x + y * nx + (int64_t)z * nx * ny), compute(ctx, a, x, y, z));
  }
  template <class C>
  __device__ static float compute(const C &ctx, const Args &a, const int x, const int y, const int z) {
    return Diffusion7Body<kPeriodic, false, false, false, kFma16<S>>(ctx, a.k, x, y, z, a.dim[0], a.dim[1], a.dim[2]);
  }
  // compute() for a point whose x (kX), y (kY) and z (kZ) neighbours in ctx
  // are the elements at the clamped coordinates (skeleton_star_tb.h).
  static constexpr bool kExactNeighbours = true;
  template <bool kX, bool kY, bool kZ, class C>
  __device__ static float compute_exact(const C &ctx, const Args &a, const int x, const int y, const int z) {
    return Diffusion7Body<kPeriodic, kX, kY, kZ, kFma16<S>>(ctx, a.k, x, y, z, a.dim[0], a.dim[1], a.dim[2]);
  }
};

// ---- variant family 2: hand-written z-march with a register queue --------
// Thread per (x, y) column; z-1/z/z+1 of the column live in registers and
// the four in-plane neighbours are read from HBM/L2 directly.
template <bool kPeriodic>
__device__ __forceinline__ int Nb(int i, int d, int n) {
  int j = i + d;
  if (kPeriodic) return j < 0 ? j + n : (j >= n ? j - n : j);
  return j < 0 ? 0 : (j >= n ? n - 1 : j);
}

template <class S, bool kPeriodic, int BX, int BY>
__global__ void __launch_bounds__(BX * BY)
Diffusion7Direct(const typename S::stored_t *__restrict__ in, typename S::stored_t *__restrict__ out,
                 int nx, int ny, int nz, psk::Box box, int zchunk, Coef7 k) {
  const int x = box.lo[0] + (int)blockIdx.x * BX + (int)threadIdx.x;
  const int y = box.lo[1] + (int)blockIdx.y * BY + (int)threadIdx.y;
  if (x >= box.hi[0] || y >= box.hi[1]) return;
  const int zb = box.lo[2] + (int)blockIdx.z * zchunk;
  const int ze = min(zb + zchunk, box.hi[2]);
  const int64_t plane = (int64_t)nx * ny;
  const int xw = Nb<kPeriodic>(x, -1, nx), xe = Nb<kPeriodic>(x, 1, nx);
  const int yn = Nb<kPeriodic>(y, -1, ny), ys = Nb<kPeriodic>(y, 1, ny);
  const int64_t col = x + (int64_t)y * nx;
  const int64_t ow = xw + (int64_t)y * nx, oe = xe + (int64_t)y * nx;
  const int64_t on = x + (int64_t)yn * nx, os = x + (int64_t)ys * nx;
  float bv = S::load(in + col + (int64_t)Nb<kPeriodic>(zb, -1, nz) * plane);
  float cv = S::load(in + col + (int64_t)zb * plane);
  for (int z = zb; z < ze; ++z) {
    const int64_t zo = (int64_t)z * plane;
    const float tv = S::load(in + col + (int64_t)Nb<kPeriodic>(z, 1, nz) * plane);
    const float wv = S::load(in + ow + zo), ev = S::load(in + oe + zo);
    const float nv = S::load(in + on + zo), sv = S::load(in + os + zo);
    float v;
    if constexpr (kFma16<S>) {  // same evaluation as Diffusion7Body
      v = k.cc * cv;
      v = __builtin_fmaf(k.cw, wv, v);
      v = __builtin_fmaf(k.ce, ev, v);
      v = __builtin_fmaf(k.cs, sv, v);
      v = __builtin_fmaf(k.cn, nv, v);
      v = __builtin_fmaf(k.cb, bv, v);
      v = __builtin_fmaf(k.ct, tv, v);
    } else {
      v = k.cc * cv + k.cw * wv + k.ce * ev + k.cs * sv + k.cn * nv + k.cb * bv + k.ct * tv;
    }
    S::store(out + col + zo, v);
    bv = cv;
    cv = tv;
  }
}

// ---- variant table --------------------------------------------------------
struct Variant {
  const char *name;
  int family;  // 1 LDS-ring skeleton, 2 direct, 3 vectorised star skeleton
  int bx, by;
  int zchunk;  // 0 = whole column
  bool remap;
  int rows = 1;  // family 1: tile rows per thread
  bool renamed = false;  // family 4: the loop schedule without queue moves
  int sched = psk::kChunked;  // family 4: how workgroups split the z-march (psk::TB2Sched)
};

const Variant kVariants[] = {
    {"skeleton64x4_z64", 1, 64, 4, 64, true},
    {"skeleton64x8_z64", 1, 64, 8, 64, true},
    {"skeleton64x4_zall", 1, 64, 4, 0, true},
    {"skeleton64x4_z64_noremap", 1, 64, 4, 64, false},
    {"direct64x4_z64", 2, 64, 4, 64, false},
    {"direct64x4_zall", 2, 64, 4, 0, false},
    {"skeleton64x2_z32", 1, 64, 2, 32, true},
    {"skeleton64x8_z128", 1, 64, 8, 128, true},
    {"star4_by4_z64", 3, 256, 4, 64, true},
    {"star4_by8_z64", 3, 256, 8, 64, true},
    {"star4_by2_z64", 3, 256, 2, 64, true},
    {"star4_by4_z128", 3, 256, 4, 128, true},
    {"star4_by4_z32", 3, 256, 4, 32, true},
    {"star4_by8_z128", 3, 256, 8, 128, true},
    {"star4_by4_z64_noremap", 3, 256, 4, 64, false},
    {"star4_by1_z64", 3, 256, 1, 64, true},
    // Family 4: two fused steps per pass; bx = waves per row (rows <= 256*bx).
    {"tb2_w2_by8_z32", 4, 2, 8, 32, true},
    {"tb2_w2_by8_z64", 4, 2, 8, 64, true},
    {"tb2_w2_by6_z32", 4, 2, 6, 32, true},
    {"tb2_w2_by4_z32", 4, 2, 4, 32, true},
    {"tb2_w2_by8_z16", 4, 2, 8, 16, true},
    {"tb2_w1_by16_z32", 4, 1, 16, 32, true},
    {"tb2_w1_by8_z32", 4, 1, 8, 32, true},
    {"tb2_w4_by4_z32", 4, 4, 4, 32, true},
    {"skeleton64x4r2_z64", 1, 64, 4, 64, true, 2},
    {"skeleton64x4r4_z64", 1, 64, 4, 64, true, 4},
    {"skeleton64x8r2_z64", 1, 64, 8, 64, true, 2},
    {"skeleton64x2r4_z64", 1, 64, 2, 64, true, 4},
    {"tb2_w1_by8_z16", 4, 1, 8, 16, true},
    {"tb2_w1_by8_z8", 4, 1, 8, 8, true},
    {"tb2_w1_by16_z16", 4, 1, 16, 16, true},
    {"tb2_w1_by16_z8", 4, 1, 16, 8, true},
    {"tb2_w2_by8_z8", 4, 2, 8, 8, true},
    {"tb2_w2_by6_z16", 4, 2, 6, 16, true},
    {"tb2r_w2_by8_z32", 4, 2, 8, 32, true, 1, true},
    {"tb2r_w2_by8_z16", 4, 2, 8, 16, true, 1, true},
    {"tb2r_w1_by8_z32", 4, 1, 8, 32, true, 1, true},
    {"tb2r_w1_by16_z32", 4, 1, 16, 32, true, 1, true},
    {"tb2r_w2_by8_z8", 4, 2, 8, 8, true, 1, true},
    {"tb2r_w1_by8_z16", 4, 1, 8, 16, true, 1, true},
    {"tb2r_w1_by8_z8", 4, 1, 8, 8, true, 1, true},
    {"tb2g_w2_by8_m8", 4, 2, 8, 8, true, 1, true, psk::kGuided},
    {"tb2g_w2_by8_m16", 4, 2, 8, 16, true, 1, true, psk::kGuided},
    {"tb2g_w1_by8_m8", 4, 1, 8, 8, true, 1, true, psk::kGuided},
    {"tb2g2_w2_by8_m8", 4, 2, 8, 8, true, 1, true, psk::kGuided2},
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
// Measured on MI355X at 512^3 FP32 (bench.py --sweep): star4_by8_z128 ~710
// GLUPS one step per pass; two steps per pass ~1190 GLUPS with the shape of
// psk::ChooseTB2Shape, which variant 0 uses for every grid size.
// One-step variant used by family-4 runs for unfused steps.
constexpr int kStarVariant = 14;

// Launches one fused pair of steps src -> (on chip) -> dst with variant v
// (family 4); a non-null mid also receives the intermediate step (rotating-
// queue variants only).  Returns hipErrorNotSupported if the grid does not
// qualify.
template <class S, bool kPeriodic>
hipError_t LaunchPair(const Variant &v, const void *src, void *dst, void *mid, const int *dim, const Coef7 &k,
                      hipStream_t s) {
  using P = DiffusionPlan<S, kPeriodic>;
  using st = typename S::stored_t;
  typename P::Args a1, a2;
  a1.g1 = static_cast<const st *>(src);
  a1.g2 = static_cast<st *>(mid);
  a2.g1 = nullptr;
  a2.g2 = static_cast<st *>(dst);
  for (int d = 0; d < 3; ++d) a1.dim[d] = a2.dim[d] = dim[d];
  a1.k = a2.k = k;
  if (mid && !v.renamed) return hipErrorInvalidValue;
#define PSH_TBR(NWX, BY)                                                                                   \
  if (v.renamed && v.bx == NWX && v.by == BY) {                                                            \
    if (!psk::StarTB2Supported<NWX>(dim[0], dim[1], dim[2])) return hipErrorNotSupported;                \
    return mid ? psk::LaunchZMarchStarTB2<P, NWX, BY, true, true>(a1, a2, dim[1], dim[2], v.zchunk, v.remap, s,  \
                                                                  v.sched)                                  \
               : psk::LaunchZMarchStarTB2<P, NWX, BY, true, false>(a1, a2, dim[1], dim[2], v.zchunk, v.remap, s, \
                                                                   v.sched);                                \
  }
  PSH_TBR(2, 8)
  PSH_TBR(1, 8)
  PSH_TBR(1, 16)
  PSH_TBR(4, 4)
#undef PSH_TBR
  if (v.renamed) return hipErrorInvalidValue;
#define PSH_TB(NWX, BY)                                                                     \
  if (v.bx == NWX && v.by == BY) {                                                          \
    if (!psk::StarTB2Supported<NWX>(dim[0], dim[1], dim[2])) return hipErrorNotSupported; \
    return psk::LaunchZMarchStarTB2<P, NWX, BY>(a1, a2, dim[1], dim[2], v.zchunk, v.remap, s); \
  }
  PSH_TB(2, 8)
  PSH_TB(2, 6)
  PSH_TB(2, 4)
  PSH_TB(1, 16)
  PSH_TB(1, 8)
  PSH_TB(4, 4)
#undef PSH_TB
  return hipErrorInvalidValue;
}

// A fused pair for whole-grid runs whose rows do not fill whole 256-point
// star waves (rows of any length up to 512 points): the packed-row tile
// (rows of at least 4 points), else the stride tile.
template <class S, bool kPeriodic>
hipError_t LaunchStridePair(const void *src, void *dst, void *mid, const int *dim, const Coef7 &k, hipStream_t s) {
  using P = DiffusionPlan<S, kPeriodic>;
  using st = typename S::stored_t;
  typename P::Args a1, a2;
  a1.g1 = static_cast<const st *>(src);
  a1.g2 = static_cast<st *>(mid);
  a2.g1 = nullptr;
  a2.g2 = static_cast<st *>(dst);
  for (int d = 0; d < 3; ++d) a1.dim[d] = a2.dim[d] = dim[d];
  a1.k = a2.k = k;
  const hipError_t e = psk::LaunchPackTB2Grid<P>(a1, a2, dim[0], dim[1], dim[2], s, mid != nullptr);
  if (e != hipErrorNotSupported) return e;
  return psk::LaunchStrideTB2Grid<P>(a1, a2, dim[0], dim[1], dim[2], s, mid != nullptr);
}

hipError_t LaunchStridePairStorage(int storage, bool periodic, const void *src, void *dst, void *mid,
                                   const int *dim, const Coef7 &k, hipStream_t s) {
  switch (storage) {
    case PS_HIP_STORAGE_NATIVE:
      return periodic ? LaunchStridePair<psk::Native<float>, true>(src, dst, mid, dim, k, s)
                      : LaunchStridePair<psk::Native<float>, false>(src, dst, mid, dim, k, s);
    case PS_HIP_STORAGE_BF16:
      return periodic ? LaunchStridePair<psk::BF16, true>(src, dst, mid, dim, k, s)
                      : LaunchStridePair<psk::BF16, false>(src, dst, mid, dim, k, s);
    case PS_HIP_STORAGE_FP16:
      return periodic ? LaunchStridePair<psk::FP16, true>(src, dst, mid, dim, k, s)
                      : LaunchStridePair<psk::FP16, false>(src, dst, mid, dim, k, s);
  }
  return hipErrorInvalidValue;
}

hipError_t LaunchPairStorage(const Variant &v, int storage, bool periodic, const void *src, void *dst, void *mid,
                             const int *dim, const Coef7 &k, hipStream_t s) {
  switch (storage) {
    case PS_HIP_STORAGE_NATIVE:
      return periodic ? LaunchPair<psk::Native<float>, true>(v, src, dst, mid, dim, k, s)
                      : LaunchPair<psk::Native<float>, false>(v, src, dst, mid, dim, k, s);
    case PS_HIP_STORAGE_BF16:
      return periodic ? LaunchPair<psk::BF16, true>(v, src, dst, mid, dim, k, s)
                      : LaunchPair<psk::BF16, false>(v, src, dst, mid, dim, k, s);
    case PS_HIP_STORAGE_FP16:
      return periodic ? LaunchPair<psk::FP16, true>(v, src, dst, mid, dim, k, s)
                      : LaunchPair<psk::FP16, false>(v, src, dst, mid, dim, k, s);
  }
  return hipErrorInvalidValue;
}

// Whether family-4 variant v can fuse steps on this grid and domain.
bool PairApplies(const Variant &v, const int *dim, const psk::Box &box) {
  for (int d = 0; d < 3; ++d)
    if (box.lo[d] != 0 || box.hi[d] != dim[d]) return false;  // whole-grid domain only
  switch (v.bx) {
    case 1: return psk::StarTB2Supported<1>(dim[0], dim[1], dim[2]);
    case 2: return psk::StarTB2Supported<2>(dim[0], dim[1], dim[2]);
    case 4: return psk::StarTB2Supported<4>(dim[0], dim[1], dim[2]);
  }
  return false;
}

template <class S, bool kPeriodic>
hipError_t LaunchOne(const Variant &v, const void *in, void *out, const int *dim, const psk::Box &box,
                     const Coef7 &k, hipStream_t s) {
  using st = typename S::stored_t;
  if (v.family == 3) {
    // Whole-grid steps whose rows do not fill whole 256-point star waves:
    // the packed-row tile.
    bool whole = true;
    for (int d = 0; d < 3; ++d) whole = whole && box.lo[d] == 0 && box.hi[d] == dim[d];
    if (whole && dim[0] % 256 != 0) {
      typename DiffusionPlan<S, kPeriodic>::Args a;
      a.g1 = static_cast<const st *>(in);
      a.g2 = static_cast<st *>(out);
      for (int d = 0; d < 3; ++d) a.dim[d] = dim[d];
      a.k = k;
      const hipError_t e = psk::LaunchPackStepGrid<DiffusionPlan<S, kPeriodic>>(a, dim[0], dim[1], dim[2], s);
      if (e != hipErrorNotSupported) return e;
    }
  }
  if (v.family == 3 && !psk::StarSupported(dim[0])) {
    // Rows that are not whole 16-byte vectors: whole-grid steps with rows up
    // to 256 points run on the stride skeleton (1.3-5.6x the LDS ring on
    // 130^3-257^3), others on the LDS ring (two rows per thread when the grid
    // is large enough, as generated code; eight points per lane on the stride
    // skeleton cost occupancy and ran 14% slower at 510^3).
    typename DiffusionPlan<S, kPeriodic>::Args a;
    a.g1 = static_cast<const st *>(in);
    a.g2 = static_cast<st *>(out);
    for (int d = 0; d < 3; ++d) a.dim[d] = dim[d];
    a.k = k;
    bool whole = true;
    for (int d = 0; d < 3; ++d) whole = whole && box.lo[d] == 0 && box.hi[d] == dim[d];
    if (whole && dim[0] <= 256)
      return psk::LaunchStrideStepGrid<DiffusionPlan<S, kPeriodic>>(a, dim[0], dim[1], dim[2], s);
    return psk::LaunchZMarch3D<DiffusionPlan<S, kPeriodic>, true>(a, box, 64, true, s);
  }
  if (v.family == 1 || v.family == 3) {
    typename DiffusionPlan<S, kPeriodic>::Args a;
    a.g1 = static_cast<const st *>(in);
    a.g2 = static_cast<st *>(out);
    a.dim[0] = dim[0];
    a.dim[1] = dim[1];
    a.dim[2] = dim[2];
    a.k = k;
    using P = DiffusionPlan<S, kPeriodic>;
    if (v.family == 3) {
      switch (v.by) {
        case 1: return psk::LaunchZMarchStar<P, 1>(a, box, v.zchunk, v.remap, s);
        case 2: return psk::LaunchZMarchStar<P, 2>(a, box, v.zchunk, v.remap, s);
        case 4: return psk::LaunchZMarchStar<P, 4>(a, box, v.zchunk, v.remap, s);
        case 8: return psk::LaunchZMarchStar<P, 8>(a, box, v.zchunk, v.remap, s);
      }
      return hipErrorInvalidValue;
    }
#define PSH_RING(BX, BY, R) \
    if (v.bx == BX && v.by == BY && v.rows == R) return psk::LaunchZMarch<P, BX, BY, R>(a, box, v.zchunk, v.remap, s);
    PSH_RING(64, 4, 1) PSH_RING(64, 8, 1) PSH_RING(64, 2, 1) PSH_RING(64, 4, 2) PSH_RING(64, 4, 4) PSH_RING(64, 8, 2)
    PSH_RING(64, 2, 4)
#undef PSH_RING
    return hipErrorInvalidValue;
  }
  const int ex = box.hi[0] - box.lo[0], ey = box.hi[1] - box.lo[1], ez = box.hi[2] - box.lo[2];
  if (ex <= 0 || ey <= 0 || ez <= 0) return hipSuccess;
  int zc = v.zchunk <= 0 || v.zchunk > ez ? ez : v.zchunk;
  dim3 grid((ex + 63) / 64, (ey + 3) / 4, (ez + zc - 1) / zc);
  hipLaunchKernelGGL((Diffusion7Direct<S, kPeriodic, 64, 4>), grid, dim3(64, 4, 1), 0, s,
                     static_cast<const st *>(in), static_cast<st *>(out), dim[0], dim[1], dim[2], box,
                     zc, k);
  return hipGetLastError();
}

template <class S>
hipError_t LaunchStorage(const Variant &v, bool periodic, const void *in, void *out, const int *dim,
                         const psk::Box &box, const Coef7 &k, hipStream_t s) {
  return periodic ? LaunchOne<S, true>(v, in, out, dim, box, k, s)
                  : LaunchOne<S, false>(v, in, out, dim, box, k, s);
}

hipError_t LaunchStep(const Variant &v, int storage, bool periodic, const void *in, void *out,
                      const int *dim, const psk::Box &box, const Coef7 &k, hipStream_t s) {
  switch (storage) {
    case PS_HIP_STORAGE_NATIVE:
      return LaunchStorage<psk::Native<float>>(v, periodic, in, out, dim, box, k, s);
    case PS_HIP_STORAGE_BF16:
      return LaunchStorage<psk::BF16>(v, periodic, in, out, dim, box, k, s);
    case PS_HIP_STORAGE_FP16:
      return LaunchStorage<psk::FP16>(v, periodic, in, out, dim, box, k, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace
}  // namespace physis_hip

using namespace physis_hip;

extern "C" {

int PSHipDiffusion7Variants(void) { return kNumVariants; }

// The packed-row tiles' geometry for rows of nx points and workgroups of
// 64*nw lanes (psk::PackLaneOf, psk::PackHaloOf): lane_out[8*lane ..] =
// r, vi, xs, shadow, wj, wd, ej, ed for every lane; halo_out[4*x ..] = the
// two (slot, position) places of point x of a halo row whose vectors start
// at position 0.  Returns the rows per workgroup, or PS_HIP_ERR_INVALID.
int PSHipPackGeometry(int nx, int nw, int periodic, int *lane_out, int *halo_out) {
  if (nx < 4 || nw < 1 || nw > 16 || !lane_out || !halo_out) return PS_HIP_ERR_INVALID;
  const int R = psk::PackRows(nx, nw);
  if (R < 1) return PS_HIP_ERR_INVALID;
  for (int lane = 0; lane < 64 * nw; ++lane) {
    const psk::PackLane p = psk::PackLaneOf(lane, nx, nw, periodic != 0);
    const int v[8] = {p.r, p.vi, p.xs, p.shadow, p.wj, p.wd, p.ej, p.ed};
    for (int k = 0; k < 8; ++k) lane_out[8 * lane + k] = v[k];
  }
  for (int x = 0; x < nx; ++x) {
    const psk::PackHalo h = psk::PackHaloOf(x, nx, 0);
    halo_out[4 * x] = h.ja;
    halo_out[4 * x + 1] = h.pa;
    halo_out[4 * x + 2] = h.jb;
    halo_out[4 * x + 3] = h.pb;
  }
  return R;
}

const char *PSHipDiffusion7VariantName(int variant) {
  if (variant < 1 || variant > kNumVariants) return nullptr;
  return kVariants[variant - 1].name;
}

int PSHipDiffusion7Run(__PSGrid *g1, __PSGrid *g2, const __PSDomain *dom, const float *coef,
                       int boundary, int steps, int variant) {
  if (!g1 || !g2 || !dom || !coef || steps < 0) return PS_HIP_ERR_INVALID;
  if (boundary != PS_HIP_CLAMPED && boundary != PS_HIP_PERIODIC) return PS_HIP_ERR_INVALID;
  if (g1->num_dims != 3 || g2->num_dims != 3 || g1->elm_size != 4 || g2->elm_size != 4)
    return PS_HIP_ERR_INVALID;
  for (int d = 0; d < 3; ++d) {
    if (g1->dim[d] != g2->dim[d]) return PS_HIP_ERR_INVALID;
    if (dom->local_min[d] < 0 || dom->local_max[d] > g1->dim[d]) return PS_HIP_ERR_INVALID;
  }
  if (Rec(g1)->storage != Rec(g2)->storage) return PS_HIP_ERR_INVALID;
  if (variant < 0 || variant > kNumVariants) return PS_HIP_ERR_INVALID;
  // Variant 0: two fused steps per pass with the shape psk::ChooseTB2Shape
  // picks for this grid.
  const psk::TB2Shape sh = psk::ChooseTB2Shape(g1->dim[0], g1->dim[1], g1->dim[2]);
  const Variant autov{"tb2_auto", 4, sh.nwx, sh.by, sh.zchunk, true, 1, true, sh.sched};
  const Variant &v = variant == 0 ? autov : kVariants[variant - 1];
  Coef7 k{coef[0], coef[1], coef[2], coef[3], coef[4], coef[5], coef[6]};
  psk::Box box = psk::BoxFromDomain(*dom, 3);
  hipStream_t s = Runtime::Get().stream();
  const int storage = Rec(g1)->storage;
  const bool periodic = boundary == PS_HIP_PERIODIC;
  if (v.family == 4) {
    const Variant &one = kVariants[kStarVariant - 1];
    // The default variant fuses whole-grid runs on the packed-row tile
    // unless the rows fill whole 256-point star waves (rows of 4 to 512
    // points; the star tile leaves lanes idle at other widths).
    bool whole = true;
    for (int d = 0; d < 3; ++d) whole = whole && box.lo[d] == 0 && box.hi[d] == g1->dim[d];
    const bool pack_first = variant == 0 && whole && g1->dim[0] % 256 != 0 && g1->dim[0] >= 4 && g1->dim[0] <= 512;
    const bool star_pair = !pack_first && PairApplies(v, g1->dim, box);
    const bool stride_pair = !star_pair && variant == 0 && whole && g1->dim[0] <= 512;
    auto pair = [&](const void *src, void *dst, void *mid) {
      return star_pair ? LaunchPairStorage(v, storage, periodic, src, dst, mid, g1->dim, k, s)
                       : LaunchStridePairStorage(storage, periodic, src, dst, mid, g1->dim, k, s);
    };
    auto single = [&](const void *in, void *out) { return LaunchStep(one, storage, periodic, in, out, g1->dim, box, k, s); };
    const int pairs = steps / 2;
    if (pairs >= 2 && (star_pair || stride_pair)) {
      void *aux = Runtime::Get().Aux(Rec(g1)->bytes);
      if (!aux) return PS_HIP_ERR_NOMEM;
      // Map i reads g1 and writes g2 for even i, the reverse for odd i, so
      // after `steps` maps the grid the last map wrote holds u_steps and the
      // other u_(steps-1).  An odd count first runs map 0 (g1 -> g2) alone;
      // the pairs then start from A and leave u_steps in A, u_(steps-1) in B.
      void *A = g1->p0, *B = g2->p0;
      if (steps & 1) {
        if (single(g1->p0, g2->p0) != hipSuccess) return PS_HIP_ERR_DEVICE;
        A = g2->p0;
        B = g1->p0;
      }
      if (v.renamed) {
        // Every pair fused; the last pass also stores its intermediate step
        // into B (psk::FusedPathDst).
        const void *src = A;
        for (int j = 0; j < pairs; ++j) {
          void *dst = psk::FusedPathDst(j, pairs, A, B, aux);
          if (pair(src, dst, j == pairs - 1 ? B : nullptr) != hipSuccess) return PS_HIP_ERR_DEVICE;
          src = dst;
        }
        return PS_HIP_OK;
      }
      // Moving-queue variants: pairs - 1 fused passes between A and the third
      // buffer, then the last pair unfused (src -> B -> A).
      const void *src = A;
      for (int j = 0; j < pairs - 1; ++j) {
        void *dst = (j & 1) ? A : aux;
        if (pair(src, dst, nullptr) != hipSuccess) return PS_HIP_ERR_DEVICE;
        src = dst;
      }
      if (single(src, B) != hipSuccess || single(B, A) != hipSuccess) return PS_HIP_ERR_DEVICE;
      return PS_HIP_OK;
    }
    for (int i = 0; i < steps; ++i) {
      if (single((i & 1) ? g2->p0 : g1->p0, (i & 1) ? g1->p0 : g2->p0) != hipSuccess) return PS_HIP_ERR_DEVICE;
    }
    return PS_HIP_OK;
  }
  for (int i = 0; i < steps; ++i) {
    const __PSGrid *in = (i & 1) ? g2 : g1;
    __PSGrid *out = (i & 1) ? g1 : g2;
    if (LaunchStep(v, storage, periodic, in->p0, out->p0, g1->dim, box, k, s) != hipSuccess)
      return PS_HIP_ERR_DEVICE;
  }
  return PS_HIP_OK;
}

}  // extern "C"
