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
#pragma once

#include <hip/hip_runtime.h>

namespace physis_hip {

struct Coef7 {
  float ce, cw, cn, cs, ct, cb, cc;
};

template <bool kPeriodic, bool kX, bool kY, bool kZ, class C>
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
  return k.cc * c + k.cw * w + k.ce * e + k.cs * s + k.cn * n + k.cb * b + k.ct * t;
}

}  // namespace physis_hip
