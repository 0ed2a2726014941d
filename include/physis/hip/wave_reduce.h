// wave_reduce.h -- wave64 reductions for gfx950.
//
// Within each 16-lane row the tree uses DPP lane permutations
// (quad_perm xor-1, quad_perm xor-2, row_half_mirror, row_mirror); the two
// cross-row steps use lane shuffles.  Every lane ends with the same value and
// the pairing order is fixed, so results are bitwise reproducible.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace psk {

// DPP control words (GFX9 encoding).
enum : int {
  kDppQuadXor1 = 0xB1,       // quad_perm [1,0,3,2]
  kDppQuadXor2 = 0x4E,       // quad_perm [2,3,0,1]
  kDppRowHalfMirror = 0x141, // lane i <-> 7-i within each half row
  kDppRowMirror = 0x140      // lane i <-> 15-i within each row
};

template <int kCtrl>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xF, 0xF, false));
}

template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double v) {
  int64_t bits = __double_as_longlong(v);
  int lo = (int)(bits & 0xffffffff), hi = (int)(bits >> 32);
  lo = __builtin_amdgcn_update_dpp(0, lo, kCtrl, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(0, hi, kCtrl, 0xF, 0xF, false);
  return __longlong_as_double(((int64_t)(uint32_t)lo) | ((int64_t)hi << 32));
}

template <class T, int kCtrl>
__device__ __forceinline__ T dpp_move(T v) {
  if constexpr (sizeof(T) == 8) return dpp_f64<kCtrl>(v);
  else return dpp_f32<kCtrl>(v);
}

// Reduces v over the 64 lanes of the wave with the binary operator op
// (which must be commutative); every lane receives the result.
template <class T, class Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
  v = op(v, dpp_move<T, kDppQuadXor1>(v));
  v = op(v, dpp_move<T, kDppQuadXor2>(v));
  v = op(v, dpp_move<T, kDppRowHalfMirror>(v));
  v = op(v, dpp_move<T, kDppRowMirror>(v));
  v = op(v, __shfl_xor(v, 16, 64));
  v = op(v, __shfl_xor(v, 32, 64));
  return v;
}

struct OpSum {
  template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return a + b; }
};
struct OpProd {
  template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return a * b; }
};
struct OpMax {
  template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return a > b ? a : b; }
};
struct OpMin {
  template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return a < b ? a : b; }
};

// Reduces v over a workgroup of kThreads (a multiple of 64) threads; the
// result is valid in every thread.  `lds` needs kThreads/64 elements; tid is
// the flat thread index (waves are consecutive runs of 64 tids).
template <int kThreads, class T, class Op>
__device__ __forceinline__ T block_reduce(T v, Op op, T *lds, int tid) {
  static_assert(kThreads % 64 == 0, "workgroup must be whole waves");
  constexpr int kWaves = kThreads / 64;
  v = wave_reduce(v, op);
  const int lane = tid & 63, wave = tid >> 6;
  if (kWaves == 1) return v;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  T r = lds[0];
#pragma unroll
  for (int w = 1; w < kWaves; ++w) r = op(r, lds[w]);
  __syncthreads();
  return r;
}

// Sum of p[0 .. n) over a workgroup of kThreads threads: thread i adds
// p[i], p[i + kThreads], ... in order, then the workgroup tree-reduces.  The
// grouping is fixed, so the result is bitwise reproducible.
template <int kThreads>
__device__ __forceinline__ double block_sum(const double *p, int n, double *lds, int tid) {
  double acc = 0.0;
  for (int i = tid; i < n; i += kThreads) acc += p[i];
  return block_reduce<kThreads>(acc, OpSum(), lds, tid);
}

}  // namespace psk
