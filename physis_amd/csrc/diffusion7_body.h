// diffusion7_body.h -- the 7-point diffusion body shared by the single-grid
// and batched built-in plans.
//
// The DSL body of the reference benchmark (examples/diffusion-benchmark/
// diffusion3d_physis.c:29-58; periodic form test_7-pt-periodic.c:12-24):
//   out = cc*c + cw*w + ce*e + cs*s + cn*n + cb*b + ct*t
// where a clamped face reads the point itself (`if (x == 0) w = c; else
// w = g(x-1)`).  kX/kY/kZ: the skeleton supplies, along that axis, the
// neighbour at the clamped coordinate (the point itself at a face), so the
// face test is dropped; the value, and so the result, is the same.
//
// kFma: the same terms in the same order, each product after the first
// fused into the running sum (one rounding per term instead of two: 7
// instructions instead of 13).  Used for grids stored in 16 bits, whose
// results are judged against FP64 within the storage's rounding, not bit
// for bit against the FP32 reference target; FP32 grids keep the
// expression as written, rounded as the sequential reference rounds it.
#pragma once

#include <hip/hip_runtime.h>

namespace physis_hip {

struct Coef7 {
  float ce, cw, cn, cs, ct, cb, cc;
};

// Whether grids of storage S evaluate the body with fused multiply-adds.
template <class S>
constexpr bool kFma16 = sizeof(typename S::stored_t) == 2;

template <bool kPeriodic, bool kX, bool kY, bool kZ, bool kFma = false, class C>
__device__ __forceinline__ float Diffusion7Body(const C &ctx, const Coef7 &k, const int x, const int y, const int z,
                                                const int nx, const int ny, const int nz) {
  float c, w, e, n, s, b, t;
  c = ctx.at(0, 0, 0, 0);
  if (kPeriodic || kX) {
    w = ctx.at(0, -1, 0, 0);
    e = ctx.at(0, 1, 0, 0);
  } else {
    if (x == 0) w = c; else w = ctx.at(0, -1, 0, 0);
    if (x == nx - 1) e = c; else e = ctx.at(0, 1, 0, 0);
  }
  if (kPeriodic || kY) {
    n = ctx.at(0, 0, -1, 0);
    s = ctx.at(0, 0, 1, 0);
  } else {
    if (y == 0) n = c; else n = ctx.at(0, 0, -1, 0);
    if (y == ny - 1) s = c; else s = ctx.at(0, 0, 1, 0);
  }
  if (kPeriodic || kZ) {
    b = ctx.at(0, 0, 0, -1);
    t = ctx.at(0, 0, 0, 1);
  } else {
    if (z == 0) b = c; else b = ctx.at(0, 0, 0, -1);
    if (z == nz - 1) t = c; else t = ctx.at(0, 0, 0, 1);
  }
  if constexpr (kFma) {
    float r = k.cc * c;
    r = __builtin_fmaf(k.cw, w, r);
    r = __builtin_fmaf(k.ce, e, r);
    r = __builtin_fmaf(k.cs, s, r);
    r = __builtin_fmaf(k.cn, n, r);
    r = __builtin_fmaf(k.cb, b, r);
    return __builtin_fmaf(k.ct, t, r);
  } else {
    return k.cc * c + k.cw * w + k.ce * e + k.cs * s + k.cn * n + k.cb * b + k.ct * t;
  }
}

}  // namespace physis_hip
