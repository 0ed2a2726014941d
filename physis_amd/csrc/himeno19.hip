// himeno19.hip -- the Himeno 19-point Jacobi program, as physisc-hip lowers
// reference examples/himeno/himenobmtxpa_physis.c:331-361 (jacobi_kernel)
// over the interior domain of :372-374, with a fused FP64 epilogue that sums
// ss*ss per sweep (the reference port leaves gosa at 0, :387-392; the
// original accumulates it in float, himenobmtxpa_original.c:334).
//
// p0 is the only grid read at offsets: it is staged in the skeleton's LDS
// ring (radius 1 in x, y and z, including the xy/yz/xz diagonals).  The 12
// coefficient grids are read at the point itself, straight from HBM.
#include "runtime_hip.h"
#include "physis/hip/skeleton.h"
#include "physis/hip/storage.h"

namespace physis_hip {
namespace {

enum { A0, A1, A2, A3, B0, B1, B2, C0, C1, C2, BND, WRK1, NCOEF };

template <class S>
struct HimenoPlan {
  using SV = float;
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
  };
  __device__ static void stage_dims(const Args &a, int &nx, int &ny, int &nz) {
    nx = a.dim[0];
    ny = a.dim[1];
    nz = a.dim[2];
  }
  __device__ static SV stage_load(const Args &a, int, int64_t off) { return S::load(a.p + off); }
  template <class C>
  __device__ static double point(const C &P, const Args &a, const int i, const int j, const int k) {
    const int64_t o = i + (int64_t)j * a.dim[0] + (int64_t)k * a.dim[0] * a.dim[1];
    auto g = [&](int n) { return S::load(a.c[n] + o); };
    float s0, ss;
    s0 = g(A0) * P.at(0, 0, 0, 1) + g(A1) * P.at(0, 0, 1, 0) + g(A2) * P.at(0, 1, 0, 0) +
         g(B0) * (P.at(0, 0, 1, 1) - P.at(0, 0, -1, 1) - P.at(0, 0, 1, -1) + P.at(0, 0, -1, -1)) +
         g(B1) * (P.at(0, 1, 1, 0) - P.at(0, 1, -1, 0) - P.at(0, -1, 1, 0) + P.at(0, -1, -1, 0)) +
         g(B2) * (P.at(0, 1, 0, 1) - P.at(0, 1, 0, -1) - P.at(0, -1, 0, 1) + P.at(0, -1, 0, -1)) +
         g(C0) * P.at(0, 0, 0, -1) + g(C1) * P.at(0, 0, -1, 0) + g(C2) * P.at(0, -1, 0, 0) + g(WRK1);
    ss = (s0 * g(A3) - P.at(0, 0, 0, 0)) * g(BND);
    const float v = P.at(0, 0, 0, 0) + a.omega * ss;
    S::store(a.q + o, v);
    return (double)ss * (double)ss;
  }
};

// Sums the per-workgroup partials of one sweep in index order.
__global__ void __launch_bounds__(256) SumPartials(const double *partials, int n, double *out) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += partials[i];
  const double r = psk::block_reduce<256>(acc, psk::OpSum(), red, threadIdx.x);
  if (threadIdx.x == 0) *out = r;
}

struct Variant {
  const char *name;
  int bx, by, zchunk;
};
const Variant kVariants[] = {
    {"skeleton64x4_z32", 64, 4, 32},
    {"skeleton64x8_z32", 64, 8, 32},
    {"skeleton64x4_zall", 64, 4, 0},
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
  return v.by == 4 ? psk::ZMarchBlocks<64, 4>(box, v.zchunk) : psk::ZMarchBlocks<64, 8>(box, v.zchunk);
}

// Queues `sweeps` sweeps p0 -> p1 -> p0 ... on stream s.  gosa_dev (may be
// NULL) receives the FP64 sum of ss^2 of each sweep; partials needs
// Blocks(v, box) doubles.
template <class S>
int EnqueueTyped(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids, float omega, int sweeps,
                 double *gosa_dev, double *partials, const Variant &v, hipStream_t s) {
  using P = HimenoPlan<S>;
  using st = typename S::stored_t;
  const psk::Box box = InteriorBox(p0);
  const int nblocks = Blocks(v, box);
  typename P::Args a;
  for (int n = 0; n < NCOEF; ++n) a.c[n] = static_cast<const st *>(grids[n]->p0);
  for (int d = 0; d < 3; ++d) a.dim[d] = p0->dim[d];
  a.omega = omega;
  a.partials = partials;
  for (int i = 0; i < sweeps; ++i) {
    a.p = static_cast<const st *>((i & 1) ? p1->p0 : p0->p0);
    a.q = static_cast<st *>((i & 1) ? p0->p0 : p1->p0);
    hipError_t e = v.by == 4 ? psk::LaunchZMarch<P, 64, 4>(a, box, v.zchunk, true, s)
                             : psk::LaunchZMarch<P, 64, 8>(a, box, v.zchunk, true, s);
    if (e != hipSuccess) return PS_HIP_ERR_DEVICE;
    if (gosa_dev && nblocks > 0) {
      hipLaunchKernelGGL(SumPartials, dim3(1), dim3(256), 0, s, partials, nblocks, gosa_dev + i);
      if (hipGetLastError() != hipSuccess) return PS_HIP_ERR_DEVICE;
    }
  }
  return PS_HIP_OK;
}

template <class S>
int RunTyped(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids, float omega, int sweeps,
             double *gosa, const Variant &v) {
  Runtime &rt = Runtime::Get();
  hipStream_t s = rt.stream();
  const int nblocks = Blocks(v, InteriorBox(p0));
  const size_t nres = sweeps > 0 ? (size_t)sweeps : 1;
  double *scratch = static_cast<double *>(rt.Scratch((nres + (size_t)nblocks + 1) * sizeof(double)));
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

}  // namespace

int HimenoPartials(const __PSGrid *p0) { return Blocks(kVariants[0], InteriorBox(p0)); }

int HimenoEnqueue(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids, float omega, int sweeps,
                  double *gosa_dev, double *partials, hipStream_t s) {
  const Variant &v = kVariants[0];
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

}  // namespace physis_hip

using namespace physis_hip;

extern "C" int PSHipHimenoRun(__PSGrid *p0, __PSGrid *p1, __PSGrid *const *grids, float omega,
                              int sweeps, double *gosa, int variant) {
  if (!p0 || !p1 || !grids || sweeps < 0) return PS_HIP_ERR_INVALID;
  if (variant == 0) variant = 1;
  if (variant < 1 || variant > kNumVariants) return PS_HIP_ERR_INVALID;
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
  const Variant &v = kVariants[variant - 1];
  switch (storage) {
    case PS_HIP_STORAGE_NATIVE: return RunTyped<psk::Native<float>>(p0, p1, grids, omega, sweeps, gosa, v);
    case PS_HIP_STORAGE_BF16: return RunTyped<psk::BF16>(p0, p1, grids, omega, sweeps, gosa, v);
    case PS_HIP_STORAGE_FP16: return RunTyped<psk::FP16>(p0, p1, grids, omega, sweeps, gosa, v);
  }
  return PS_HIP_ERR_INVALID;
}
