// storage.h -- how grid elements are held in HBM on the MI355X target.
//
// A grid of a primitive type T is stored as T.  FP32 grids may instead be
// stored as bfloat16 or IEEE half: loads widen to FP32 and stores round to
// nearest-even, so the stencil arithmetic is FP32 either way and only the
// bytes moved per update change.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace psk {

// Grid elements live in HBM.  Every grid access goes through these two
// helpers, which address global memory explicitly: a grid pointer a kernel
// loads from memory (a batch's argument table) is generic to the compiler,
// and an access through a generic pointer becomes a flat instruction, which
// also counts against the LDS wait counter, so that every wait for an LDS
// read would also wait for the grid loads in flight.
template <class T>
__device__ __forceinline__ T GlobalLoad(const T *p) {
  return *(const __attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ __forceinline__ void GlobalStoreNT(T *p, T v) {
  __builtin_nontemporal_store(v, (__attribute__((address_space(1))) T *)p);
}
template <class T>
__device__ __forceinline__ void GlobalStore(T *p, T v) {
  *(__attribute__((address_space(1))) T *)p = v;
}

// Native storage: the stored type is the value type.
template <class T>
struct Native {
  using value_t = T;
  using stored_t = T;
  static constexpr int kBytes = sizeof(T);
  __device__ __forceinline__ static T load(const stored_t *p) { return GlobalLoad(p); }
  __device__ __forceinline__ static void store(stored_t *p, T v) { GlobalStoreNT(p, v); }
  __device__ __forceinline__ static void store_cached(stored_t *p, T v) { GlobalStore(p, v); }
  __device__ __forceinline__ static T widen(stored_t s) { return s; }
  __device__ __forceinline__ static stored_t narrow(T v) { return v; }
};

typedef float v2f_t __attribute__((ext_vector_type(2)));

// bfloat16 storage of an FP32 value: the upper 16 bits, rounded to nearest
// even (gfx950's v_cvt_pk_bf16_f32, one instruction for two values); NaN
// stays NaN.
struct BF16 {
  using value_t = float;
  using stored_t = uint16_t;
  static constexpr int kBytes = 2;
  __device__ __forceinline__ static float widen(uint16_t s) {
    return __uint_as_float(((uint32_t)s) << 16);
  }
  __device__ __forceinline__ static uint16_t narrow(float v) {
    const __bf16 h = static_cast<__bf16>(v);
    uint16_t s;
    __builtin_memcpy(&s, &h, 2);
    return s;
  }
  // a and b rounded as narrow() rounds them, packed (a in the low half) by
  // one instruction; the halves of such a pair widened.
  __device__ __forceinline__ static uint32_t narrow2(float a, float b) {
    typedef __bf16 v2bf_t __attribute__((ext_vector_type(2)));
    const v2bf_t h = __builtin_convertvector((v2f_t){a, b}, v2bf_t);
    uint32_t r;
    __builtin_memcpy(&r, &h, 4);
    return r;
  }
  __device__ __forceinline__ static float widen_lo(uint32_t r) { return __uint_as_float(r << 16); }
  __device__ __forceinline__ static float widen_hi(uint32_t r) { return __uint_as_float(r & 0xffff0000u); }
  __device__ __forceinline__ static float load(const stored_t *p) { return widen(GlobalLoad(p)); }
  __device__ __forceinline__ static void store(stored_t *p, float v) { GlobalStoreNT(p, narrow(v)); }
  __device__ __forceinline__ static void store_cached(stored_t *p, float v) { GlobalStore(p, narrow(v)); }
};

// IEEE binary16 storage of an FP32 value (hardware conversion, nearest even).
struct FP16 {
  using value_t = float;
  using stored_t = uint16_t;
  static constexpr int kBytes = 2;
  __device__ __forceinline__ static float widen(uint16_t s) {
    _Float16 h;
    __builtin_memcpy(&h, &s, 2);
    return (float)h;
  }
  __device__ __forceinline__ static uint16_t narrow(float v) {
    // v is an FP32 result: the empty asm keeps the compiler from fusing the
    // FMA that produced it with this conversion into one FP16-rounding
    // v_fma_mixlo_f16 (one rounding instead of FP32 then FP16), which it
    // would do in some kernels and not in others.
    asm("" : "+v"(v));
    _Float16 h = (_Float16)v;
    uint16_t s;
    __builtin_memcpy(&s, &h, 2);
    return s;
  }
  // a and b rounded as narrow() rounds them, packed (a in the low half) by
  // one instruction (v_cvt_pk_f16_f32; the empty asm as in narrow()); the
  // halves of such a pair widened.
  __device__ __forceinline__ static uint32_t narrow2(float a, float b) {
    typedef _Float16 v2h_t __attribute__((ext_vector_type(2)));
    asm("" : "+v"(a));
    asm("" : "+v"(b));
    const v2h_t h = __builtin_convertvector((v2f_t){a, b}, v2h_t);
    uint32_t r;
    __builtin_memcpy(&r, &h, 4);
    return r;
  }
  __device__ __forceinline__ static float widen_lo(uint32_t r) { return widen((uint16_t)(r & 0xffffu)); }
  __device__ __forceinline__ static float widen_hi(uint32_t r) { return widen((uint16_t)(r >> 16)); }
  __device__ __forceinline__ static float load(const stored_t *p) { return widen(GlobalLoad(p)); }
  __device__ __forceinline__ static void store(stored_t *p, float v) { GlobalStoreNT(p, narrow(v)); }
  __device__ __forceinline__ static void store_cached(stored_t *p, float v) { GlobalStore(p, narrow(v)); }
};

}  // namespace psk
