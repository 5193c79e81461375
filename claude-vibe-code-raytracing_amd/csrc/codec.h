// Population storage codecs of the fused coupled step.
//
// ST_F32 stores populations as the reference does (float32) and computes in
// FP64, so results match src/sim.js.  ST_F16 and ST_BF16 store each population
// as a 16-bit deviation from a per-direction baseline and compute in FP32:
//   f_k is stored as f_k - w_k rho0 (rho0 = 1: the pressure proxy sum f
//     stays near 1 in both phases, so the deviations are small);
//   h_k is stored as h_k - w7_k phi_c, with phi_c the phi of the cell at the
//     time of the store (phi(t) for the post-collision h(t), phi(t+1) for the
//     post-stream h); the reader knows that phi: it is the phi array of the
//     same parity, at the source cell of the pull.
// In the bulk of either phase the deviations are near zero, where a 16-bit
// float has its finest absolute resolution.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

enum : int { ST_F32 = 0, ST_F16 = 1, ST_BF16 = 2 };

// Round a float to bfloat16, nearest even; NaN stays a quiet NaN.
// float -> bfloat16, round to nearest even (NaN: quiet, upper payload kept):
// the bits of (u + 0x7fff + lsb) >> 16, or (u >> 16) | 0x40 for a NaN.  gfx950
// computes exactly that in one v_cvt_pk_bf16_f32 (checked on MI355X against
// the integer form for all 2^32 float bit patterns, f32 denormals enabled as
// every kernel here runs); both halves of the result hold it.
__device__ __forceinline__ uint32_t f32_to_bf16_pk(float v) {
  uint32_t r;
  asm("v_cvt_pk_bf16_f32 %0, %1, %1" : "=v"(r) : "v"(v));
  return r;
}
__device__ __forceinline__ uint16_t f32_to_bf16(float v) {
  return (uint16_t)(f32_to_bf16_pk(v) & 0xffffu);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Element of an array at a per-lane byte offset from a wave-uniform base (an
// array, or a plane shifted by a uniform pull offset).  Byte offsets stay below
// 4 GiB: grids hold fewer than 2^30 cells.
template <typename T>
__device__ __forceinline__ T ldb(const T *base, uint32_t bo) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + bo);
}
template <typename T>
__device__ __forceinline__ void stb(T *base, uint32_t bo, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + bo) = v;
}
// A population store marked non-temporal: no step reads the value back before
// the whole grid has streamed through the caches, so the line should not take
// L2 room from the phi stencils and the pulls' halo lines (measured on MI355X,
// 256^3: -6% on the float32 stream kernel, -3% on the FP16 one, and a faster
// phase stream after them).
template <typename T>
__device__ __forceinline__ void stn(T *base, uint32_t bo, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<T *>(reinterpret_cast<char *>(base) + bo));
}

template <int KIND>
struct Codec;

// T: stored element; R: arithmetic of the collide and macroscopic update.
template <>
struct Codec<ST_F32> {
  using T = float;
  using R = double;
  static constexpr int bytes = 4;
  __device__ __forceinline__ static float ld(const T *p, int64_t a, float) { return p[a]; }
  __device__ __forceinline__ static void st(T *p, int64_t a, float v, float) { p[a] = v; }
  // ld / st at a byte offset from a uniform base (ldb / stb).
  __device__ __forceinline__ static float ldo(const T *p, uint32_t bo, float) { return ldb(p, bo); }
  __device__ __forceinline__ static void sto(T *p, uint32_t bo, float v, float) { stn(p, bo, v); }
  // The value st then ld give back.
  __device__ __forceinline__ static float rt(float v, float) { return v; }
};

// float -> binary16, round to nearest even, subnormals kept, as one
// v_cvt_f16_f32 of a float held in a register.  Written out so that every
// rounding to 16 bits is this instruction: the compiler would otherwise fold
// some conversions into mixed-precision FMAs (v_fma_mixlo_f16), which round the
// product-sum to 16 bits once instead of first to float, and whether it does
// depends on the surrounding code -- the same arithmetic in two kernels (the
// fused step, and a call split between stream and collide) would then store
// different bits where the float result falls on a 16-bit rounding tie (mostly
// subnormal deviations, in the gas).  The result's upper 16 bits are masked.
__device__ __forceinline__ uint16_t f32_to_f16_bits(float v) {
  uint32_t r;
  asm("v_cvt_f16_f32 %0, %1" : "=v"(r) : "v"(v));
  return (uint16_t)(r & 0xffffu);
}

template <>
struct Codec<ST_F16> {
  using T = _Float16;
  using R = float;
  static constexpr int bytes = 2;
  __device__ __forceinline__ static T h(float v) { return __builtin_bit_cast(T, f32_to_f16_bits(v)); }
  __device__ __forceinline__ static float ld(const T *p, int64_t a, float base) {
    return (float)p[a] + base;
  }
  __device__ __forceinline__ static void st(T *p, int64_t a, float v, float base) {
    p[a] = h(v - base);
  }
  __device__ __forceinline__ static float ldo(const T *p, uint32_t bo, float base) {
    return (float)ldb(p, bo) + base;
  }
  __device__ __forceinline__ static void sto(T *p, uint32_t bo, float v, float base) {
    stn(reinterpret_cast<uint16_t *>(p), bo, f32_to_f16_bits(v - base));
  }
  __device__ __forceinline__ static float rt(float v, float base) {
    return (float)h(v - base) + base;
  }
  // The stored bits of an element: dec(enc(v, b), b) == ld after st.
  __device__ __forceinline__ static uint16_t enc(float v, float base) {
    return f32_to_f16_bits(v - base);
  }
  __device__ __forceinline__ static float dec(uint16_t bits, float base) {
    return (float)__builtin_bit_cast(T, bits) + base;
  }
  // Deviations themselves (the stored value, no baseline added): dv of the
  // stored bits, ev the bits of a deviation, rtv a deviation as stored and
  // reloaded; ldd / std load and store a deviation at a byte offset.
  __device__ __forceinline__ static float dv(uint16_t bits) {
    return (float)__builtin_bit_cast(T, bits);
  }
  __device__ __forceinline__ static uint16_t ev(float v) { return f32_to_f16_bits(v); }
  __device__ __forceinline__ static float rtv(float v) { return (float)h(v); }
  __device__ __forceinline__ static float ldd(const T *p, uint32_t bo) { return (float)ldb(p, bo); }
  __device__ __forceinline__ static void std_(T *p, uint32_t bo, float v) {
    stn(reinterpret_cast<uint16_t *>(p), bo, f32_to_f16_bits(v));
  }
  // ev(v) written into the low (HI false) or high half of a packed pair, the
  // other half kept: the same v_cvt_f16_f32, with its result routed to the
  // half by the instruction's destination select (one instruction instead of
  // a conversion, a mask and a merge).
  template <bool HI>
  __device__ __forceinline__ static void put(uint32_t &w, float v) {
    if constexpr (HI)
      asm("v_cvt_f16_f32_sdwa %0, %1 dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE src0_sel:DWORD"
          : "+v"(w) : "v"(v));
    else
      asm("v_cvt_f16_f32_sdwa %0, %1 dst_sel:WORD_0 dst_unused:UNUSED_PRESERVE src0_sel:DWORD"
          : "+v"(w) : "v"(v));
  }
};

template <>
struct Codec<ST_BF16> {
  using T = uint16_t;
  using R = float;
  static constexpr int bytes = 2;
  __device__ __forceinline__ static float ld(const T *p, int64_t a, float base) {
    return bf16_to_f32(p[a]) + base;
  }
  __device__ __forceinline__ static void st(T *p, int64_t a, float v, float base) {
    p[a] = f32_to_bf16(v - base);
  }
  __device__ __forceinline__ static float ldo(const T *p, uint32_t bo, float base) {
    return bf16_to_f32(ldb(p, bo)) + base;
  }
  __device__ __forceinline__ static void sto(T *p, uint32_t bo, float v, float base) {
    stb(p, bo, f32_to_bf16(v - base));
  }
  __device__ __forceinline__ static float rt(float v, float base) {
    return bf16_to_f32(f32_to_bf16(v - base)) + base;
  }
  __device__ __forceinline__ static uint16_t enc(float v, float base) { return f32_to_bf16(v - base); }
  __device__ __forceinline__ static float dec(uint16_t bits, float base) {
    return bf16_to_f32(bits) + base;
  }
  __device__ __forceinline__ static float dv(uint16_t bits) { return bf16_to_f32(bits); }
  __device__ __forceinline__ static uint16_t ev(float v) { return f32_to_bf16(v); }
  __device__ __forceinline__ static float rtv(float v) { return bf16_to_f32(f32_to_bf16(v)); }
  __device__ __forceinline__ static float ldd(const T *p, uint32_t bo) { return bf16_to_f32(ldb(p, bo)); }
  __device__ __forceinline__ static void std_(T *p, uint32_t bo, float v) { stn(p, bo, f32_to_bf16(v)); }
  // ev(v) written into the low (HI false) or high half of a packed pair.
  template <bool HI>
  __device__ __forceinline__ static void put(uint32_t &w, float v) {
    const uint32_t b = f32_to_bf16_pk(v);
    w = HI ? (w & 0x0000ffffu) | (b & 0xffff0000u) : (w & 0xffff0000u) | (b & 0x0000ffffu);
  }
};

}  // namespace lbm
