// gc_device.hpp -- device utilities for the gfx950 kernels: counter-based
// RNG, uniform/normal transforms, wavefront reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GC_DEV __device__ __forceinline__

namespace gc {

constexpr float kPiF = 3.14159265358979323846f;

// Device global memory, said so: loads through such a pointer compile to
// global_load (a plain pointer read from a descriptor in memory gives
// flat_load, which also waits on the LDS counter).
using gfloat = __attribute__((address_space(1))) const float;
GC_DEV gfloat* as_global(const float* p) { return (gfloat*)p; }

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11), from the paper's round
// function: two 32x32->64 multiplies per round, Weyl key schedule.
GC_DEV void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                          uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
    const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
    const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
    c1 = uint32_t(p1);
    c3 = uint32_t(p0);
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Native stream: counter = {iteration, chain_lo, chain_hi, purpose}, key = seed.
enum : uint32_t { PURPOSE_NORMAL0 = 0, PURPOSE_ZZ = 16, PURPOSE_INIT = 32 };

GC_DEV void draw4(uint64_t seed, uint32_t it, uint64_t chain, uint32_t purpose, uint32_t out[4]) {
  philox4x32_10(it, uint32_t(chain), uint32_t(chain >> 32), purpose, uint32_t(seed), uint32_t(seed >> 32), out);
}

// uint32 -> (0,1): ((x >> 8) + 0.5) * 2^-24, formed as one FMA
// (x >> 8) * 2^-24 + 2^-25: the product is exact, so the FMA rounds the same
// exact value once, as the add-then-scale form does.
GC_DEV float u01(uint32_t x) { return fmaf(float(x >> 8), 1.0f / 16777216.0f, 1.0f / 33554432.0f); }

// Box-Muller pair from two uniforms: r = sqrt(-2 ln u1), angle 2 pi u2.
// v_sin_f32 / v_cos_f32 take their argument in revolutions, so the angle
// needs no range reduction.
GC_DEV float sin_rev(float x) { return __builtin_amdgcn_sinf(x); }
GC_DEV float cos_rev(float x) { return __builtin_amdgcn_cosf(x); }
// sin and cos of 2 pi x for any x (in revolutions): x - rint(x) is exact
// and lies in [-1/2, 1/2], the hardware instructions' best range
GC_DEV void sincos_rev(float x, float& s, float& c) {
  const float f = x - __builtin_rintf(x);
  s = sin_rev(f);
  c = cos_rev(f);
}
constexpr float kInv2PiF = 0.15915494309189535f;
// ln(x) = ln(2) log2(x) with the hardware v_log_f32 (x normal and positive
// here: uniforms are at least 2^-25)
constexpr float kLn2F = 0.6931471805599453f;
GC_DEV float log2v(float x) { return __builtin_amdgcn_logf(x); }
GC_DEV void box_muller(float u1, float u2, float& n0, float& n1) {
  // sqrt(-2 ln u1) = sqrt(-2 ln(2) log2(u1))
  const float r = __builtin_amdgcn_sqrtf((-2.0f * kLn2F) * log2v(u1));
  n0 = r * cos_rev(u2);
  n1 = r * sin_rev(u2);
}

// Standard normal quantile sqrt(2) erfinv(2u - 1) for u in (0, 1), with the
// single-precision erfinv of M. Giles, "Approximating the erfinv function"
// (GPU Computing Gems Jade Edition, 2011); (1 - x)(1 + x) is formed as
// 4 u (1 - u), exact near the tails.
GC_DEV float norm_inv(float u) {
  const float x = 2.f * u - 1.f;
  float w = -kLn2F * log2v(4.f * u * (1.f - u));
  float p;
  if (w < 5.f) {
    w -= 2.5f;
    p = 2.81022636e-08f;
    p = 3.43273939e-07f + p * w;
    p = -3.5233877e-06f + p * w;
    p = -4.39150654e-06f + p * w;
    p = 0.00021858087f + p * w;
    p = -0.00125372503f + p * w;
    p = -0.00417768164f + p * w;
    p = 0.246640727f + p * w;
    p = 1.50140941f + p * w;
  } else {
    w = __builtin_amdgcn_sqrtf(w) - 3.f;
    p = -0.000200214257f;
    p = 0.000100950558f + p * w;
    p = 0.00134934322f + p * w;
    p = -0.00367342844f + p * w;
    p = 0.00573950773f + p * w;
    p = -0.0076224613f + p * w;
    p = 0.00943887047f + p * w;
    p = 1.00167406f + p * w;
    p = 2.83297682f + p * w;
  }
  return 1.41421356237f * p * x;
}

// ---------------------------------------------------------------------------
// Wavefront (64-lane) reductions.  Exchanges inside a row of 16 lanes use DPP
// (quad_perm xor 1 / xor 2, row_ror 4 / 8), which fold into the VALU add;
// the two cross-row steps (lane ^ 16, lane ^ 32) go through ds_bpermute.
// (The gfx950 permlane16/32 swap builtins are not used: this compiler
// returns the first register of the swapped pair for both results.)
template <int CTRL>
GC_DEV float dpp(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
constexpr int kDppXor1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int kDppRor4 = 0x124;  // row_ror:4
constexpr int kDppRor8 = 0x128;  // row_ror:8

// Sum over the 16 lanes of each row that share lane % 4, then over rows.
GC_DEV float sum_class4(float k) {
  k += dpp<kDppRor4>(k);
  k += dpp<kDppRor8>(k);
  k += __shfl_xor(k, 16, 64);
  return k + __shfl_xor(k, 32, 64);
}

GC_DEV float wave_sum(float v) {
  v += dpp<kDppXor1>(v);
  v += dpp<kDppXor2>(v);
  return sum_class4(v);
}

// Reduce four values across the wavefront with a butterfly transpose: after
// the two in-quad steps lane L carries value index
// ((L & 1) << 1) | ((L >> 1) & 1) and the rest is a sum over that class.
// On return every lane holds the total of its index; lanes 0..3 carry
// indices 0, 2, 1, 3.
GC_DEV float wave_sum4(float v0, float v1, float v2, float v3, int lane) {
  const bool b0 = lane & 1;
  const float k0 = b0 ? v2 : v0, k1 = b0 ? v3 : v1;
  const float s0 = b0 ? v0 : v2, s1 = b0 ? v1 : v3;
  const float a0 = k0 + dpp<kDppXor1>(s0);
  const float a1 = k1 + dpp<kDppXor1>(s1);
  const bool b1 = lane & 2;
  float k = b1 ? a1 : a0;
  const float s = b1 ? a0 : a1;
  k += dpp<kDppXor2>(s);
  return sum_class4(k);
}
GC_DEV int wave_sum4_index(int lane) { return ((lane & 1) << 1) | ((lane >> 1) & 1); }

// Byte addresses of the two cross-row partners (lane ^ 16, lane ^ 32) for
// ds_bpermute, computed once per thread outside the loops that reduce.
struct XRow {
  int a16, a32;
};
GC_DEV XRow xrow_addr(int lane) { return XRow{(lane ^ 16) << 2, (lane ^ 32) << 2}; }
GC_DEV float bperm(int addr, float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, v)));
}
GC_DEV float sum_class4(float k, XRow x) {
  k += dpp<kDppRor4>(k);
  k += dpp<kDppRor8>(k);
  k += bperm(x.a16, k);
  return k + bperm(x.a32, k);
}
GC_DEV float wave_sum4(float v0, float v1, float v2, float v3, int lane, XRow x) {
  const bool b0 = lane & 1;
  const float k0 = b0 ? v2 : v0, k1 = b0 ? v3 : v1;
  const float s0 = b0 ? v0 : v2, s1 = b0 ? v1 : v3;
  const float a0 = k0 + dpp<kDppXor1>(s0);
  const float a1 = k1 + dpp<kDppXor1>(s1);
  const bool b1 = lane & 2;
  float k = b1 ? a1 : a0;
  const float s = b1 ? a0 : a1;
  k += dpp<kDppXor2>(s);
  return sum_class4(k, x);
}

// Sum over the 16 lanes of each row: every lane holds its row's total.
GC_DEV float row_sum16(float v) {
  v += dpp<kDppXor1>(v);
  v += dpp<kDppXor2>(v);
  v += dpp<kDppRor4>(v);
  return v + dpp<kDppRor8>(v);
}

GC_DEV double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

GC_DEV float sgnf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }

// reciprocal with the hardware v_rcp_f32 (1 ulp)
GC_DEV float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
// square root with the hardware v_sqrt_f32 (1 ulp), not the correctly
// rounded expansion
GC_DEV float sqrtv(float x) { return __builtin_amdgcn_sqrtf(x); }

// r[i] = 1 / d[i] for N values with one v_rcp_f32 (the transcendental unit
// bounds the fitting loops): the prefix products P_i = d_0 .. d_i, R = 1 / P,
// then r_i = P_{i-1} R_i and R_{i-1} = R_i d_i downwards (3 (N-1) products,
// a few ulp).  A product outside [2^-100, 2^100] (or not a number) takes one
// reciprocal per value instead, in a branch that the wavefront skips when
// no lane needs it.
template <int N>
GC_DEV void rcp_batch(const float* d, float* r) {
  if constexpr (N == 1) {
    r[0] = rcpf(d[0]);
  } else {
    float pre[N];
    pre[0] = d[0];
#pragma unroll
    for (int i = 1; i < N; ++i) pre[i] = pre[i - 1] * d[i];
    float R = rcpf(pre[N - 1]);
#pragma unroll
    for (int i = N - 1; i > 0; --i) {
      r[i] = pre[i - 1] * R;
      R *= d[i];
    }
    r[0] = R;
    const bool ok = pre[N - 1] >= 0x1p-100f && pre[N - 1] <= 0x1p100f;
    if (__builtin_amdgcn_ballot_w64(!ok) != 0) {
      // (the empty volatile asm keeps the reciprocals inside the branch:
      // speculated, they would run on every iteration)
      asm volatile("" ::: "memory");
#pragma unroll
      for (int i = 0; i < N; ++i)
        if (!ok) r[i] = rcpf(d[i]);
    }
  }
}

}  // namespace gc
