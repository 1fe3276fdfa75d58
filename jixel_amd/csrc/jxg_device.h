// jxg_device.h - gfx950 device helpers of the VarDCT front end: the
// reference's scalar arithmetic (fmla, ties-away rounding, saturating
// casts, the bit-trick cube root, fast_exp2, dirty_log2f) and the scaled
// DCT-II butterflies, all in the reference's operation order.
//
// The whole library is compiled with -ffp-contract=off; FUSED selects
// whether the reference's fmla (src/dct.rs:33-53) is a fused multiply-add.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jxg_internal.h"

namespace jxg {

template <bool F>
__device__ __forceinline__ float fmla(float a, float b, float c) {
  if constexpr (F)
    return __builtin_fmaf(a, b, c);
  else
    return a * b + c;
}

__device__ __forceinline__ float bits_f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t f_bits(float f) { return __float_as_uint(f); }

// Rust `as i32` / `as i16`: saturating, NaN -> 0.  Branch-free: the value is
// clamped into the target range (2147483520 is the largest float below 2^31;
// a NaN clamps to the low bound and is then replaced by 0) and truncated.
__device__ __forceinline__ int32_t sat_i32(float v) {
  const int32_t r = (int32_t)fminf(fmaxf(v, -2147483648.0f), 2147483520.0f);
  return v != v ? 0 : (v >= 2147483648.0f ? INT32_MAX : r);
}
__device__ __forceinline__ int16_t sat_i16(float v) {
  const int32_t r = (int32_t)fminf(fmaxf(v, -32768.0f), 32767.0f);
  return (int16_t)(v != v ? 0 : r);
}

// f32::round: ties away from zero.
__device__ __forceinline__ float round_away(float x) { return __builtin_roundf(x); }

// ---- cube roots (src/xyb.rs:146-170; src/avx/xyb.rs:128-205) ----
template <bool F>
__device__ __forceinline__ float halley(float x, float a) {
  float tx = x * x * x;
  return x * fmla<F>(2.0f, a, tx) / fmla<F>(2.0f, tx, a);
}

template <bool F>
__device__ __forceinline__ float cbrt_scalar(float x) {
  if (x == 0.0f) return x;
  uint32_t ui = f_bits(x);
  uint32_t hx = (ui & 0x7fffffffu) / 3u + 709958130u;
  float t = bits_f((ui & 0x80000000u) | hx);
  t = halley<F>(t, x);
  return halley<F>(t, x);
}

__device__ __forceinline__ float cbrt_avx2_positive(float a) {
  uint32_t hx = f_bits(a) & 0x7fffffffu;
  float x = bits_f((uint32_t)(((uint64_t)hx * 341u) >> 10) + 709958130u);
#pragma unroll
  for (int it = 0; it < 2; it++) {
    float tx = (x * x) * x;
    float num = __builtin_fmaf(a, 2.0f, tx);
    float den = __builtin_fmaf(tx, 2.0f, a);
    x = x * (num / den);
  }
  return a == 0.0f ? 0.0f : x;
}

// ---- fast_exp2 / dirty_log2f / fast_exp (src/adaptive_quant.rs:739-784,
//      src/util.rs:342-357) ----
template <bool F>
__device__ __forceinline__ float fast_exp2(float d, const float *table) {
  const float redux = 196608.0f;  // f32::from_bits(0x4b400000) / 64
  uint32_t ui = f_bits(d + redux);
  uint32_t i0 = ui + 32u;
  uint32_t k = i0 / 64u;
  i0 &= 63u;
  float uf = bits_f(ui) - redux;
  float z0 = table[i0];
  float f = d - uf;
  float u = 0.24022650695908768f;
  u = fmla<F>(u, f, 0.69314718055994973f);
  u *= f;
  float i2 = bits_f(((uint32_t)((int32_t)k + 0x7f)) << 23);
  return fmla<F>(u, z0, z0) * i2;
}

template <bool F>
__device__ __forceinline__ float dirty_log2f(float d) {
  uint32_t ix = f_bits(d) + (0x3f800000u - 0x3f3504f3u);
  int32_t n = (int32_t)(ix >> 23) - 0x7f;
  ix = (ix & 0x007fffffu) + 0x3f3504f3u;
  float a = bits_f(ix);
  float x = (a - 1.0f) / (a + 1.0f);
  float x2 = x * x;
  float u = 0.4121985850084821691e+0f;
  u = fmla<F>(u, x2, 0.5770780163490337802e+0f);
  u = fmla<F>(u, x2, 0.9617966939259845749e+0f);
  return fmla<F>(x2 * x, u, fmla<F>(x, 0.2885390081777926802e+1f, (float)n));
}

__device__ __forceinline__ int32_t round_i32_portable(float v) {
  float m = fabsf(v);
  float b = (m >= 0.5f && m < 8388608.0f) ? v + copysignf(0.5f, v) : v;
  return sat_i32(b);
}

template <bool F>
__device__ __forceinline__ float fast_exp(float x) {
  int32_t e = round_i32_portable(x * 1.44269504088896340736f);
  float r = fmla<F>((float)e, -0.69314575f, x);
  r = fmla<F>((float)e, -1.4286068e-6f, r);
  float sq = r * r;
  float p = fmla<F>(-0.002775669f, sq, 0.16666707f);
  p = fmla<F>(p, sq, 2.0f);
  float scale = bits_f(((uint32_t)(e + 127)) << 23);
  return (1.0f + 2.0f * r / (p - r)) * scale;
}

// ---- scaled DCT-II butterflies (src/dct.rs:68-127,687-720,1477-1500) ----
#define JXG_SQRT2 1.41421356237309504880f

template <bool F>
__device__ __forceinline__ void dct4(float *b) {
  float t0 = b[0] + b[3], t1 = b[1] + b[2];
  float e0 = t0 + t1, e1 = t0 - t1;
  float t2 = (b[0] - b[3]) * 0.5411961f;
  float t3 = (b[1] - b[2]) * 1.306563f;
  float o0 = t2 + t3, o1 = t2 - t3;
  o0 = fmla<F>(o0, JXG_SQRT2, o1);
  b[0] = e0;
  b[2] = e1;
  b[1] = o0;
  b[3] = o1;
}

template <bool F>
__device__ __forceinline__ void dct8(float *b) {
  const float wc[4] = {0.5097956f, 0.6013449f, 0.8999762f, 2.5629156f};
  float t[8];
#pragma unroll
  for (int i = 0; i < 4; i++) t[i] = b[i] + b[7 - i];
  dct4<F>(t);
#pragma unroll
  for (int i = 0; i < 4; i++) t[4 + i] = (b[i] - b[7 - i]) * wc[i];
  dct4<F>(t + 4);
  t[4] = fmla<F>(t[4], JXG_SQRT2, t[5]);
  t[5] += t[6];
  t[6] += t[7];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    b[2 * i] = t[i];
    b[2 * i + 1] = t[4 + i];
  }
}

template <bool F>
__device__ __forceinline__ void dct16(float *b) {
  const float wc[8] = {0.5024193f, 0.5224986f, 0.566944f,  0.6468218f,
                       0.7881546f, 1.0606777f, 1.7224471f, 5.1011486f};
  float t[16];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    t[i] = b[i] + b[15 - i];
    t[8 + i] = b[i] - b[15 - i];
  }
  dct8<F>(t);
#pragma unroll
  for (int i = 0; i < 8; i++) t[8 + i] *= wc[i];
  dct8<F>(t + 8);
  t[8] = fmla<F>(t[8], JXG_SQRT2, t[9]);
#pragma unroll
  for (int i = 9; i < 15; i++) t[i] += t[i + 1];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    b[2 * i] = t[i];
    b[2 * i + 1] = t[8 + i];
  }
}

template <bool F>
__device__ __forceinline__ void dct32(float *b) {
  const float wc[16] = {0.50060298f, 0.50547096f, 0.5154473f, 0.5310426f,
                        0.5531039f,  0.582935f,   0.6225041f, 0.67480834f,
                        0.7445363f,  0.83934965f, 0.97256825f, 1.16944f,
                        1.4841646f,  2.057781f,   3.4076084f, 10.190008f};
  float t[32];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    t[i] = b[i] + b[31 - i];
    t[16 + i] = b[i] - b[31 - i];
  }
  dct16<F>(t);
#pragma unroll
  for (int i = 0; i < 16; i++) t[16 + i] *= wc[i];
  dct16<F>(t + 16);
  t[16] = fmla<F>(t[16], JXG_SQRT2, t[17]);
#pragma unroll
  for (int i = 17; i < 31; i++) t[i] += t[i + 1];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    b[2 * i] = t[i];
    b[2 * i + 1] = t[16 + i];
  }
}

// WC64 (src/dct.rs:2404-2437), the reference's literal constants.
static __constant__ float WC64_DEV[32] = {
    0.500150621f, 0.501358449f, 0.50378871f,  0.507471144f, 0.51245147f,  0.518792689f,
    0.526577294f, 0.535909832f, 0.546920419f, 0.559769809f, 0.574655175f, 0.591818511f,
    0.611557364f, 0.634238958f, 0.660319805f, 0.690372109f, 0.725120544f, 0.765494168f,
    0.812702119f, 0.868344724f, 0.934583604f, 1.01440823f,  1.11207163f,  1.23383272f,
    1.38929391f,  1.59397233f,  1.87467599f,  2.28204989f,  2.9246285f,   4.08461094f,
    6.79675055f,  20.3738785f};

// dct1d_64 (src/dct.rs:2529-2546) on a line held in LDS (element stride
// `step`), computed in place so that only 32 values are in registers at a
// time.  The reference's final even / odd interleave is not performed:
// position i < 32 holds frequency 2i and position 32 + i frequency 2i + 1
// (dct64_pos gives the position of a frequency).
template <bool F>
__device__ __forceinline__ void dct64_line(float *p, int step, float scale, bool do_scale) {
  // tmp[i] = buf[i] + buf[63-i], tmp[32+i] = (buf[i] - buf[63-i]) * WC64[i];
  // positions {i, 31-i, 32+i, 63-i} are read and written by one group.
  for (int i = 0; i < 16; i++) {
    const float a = p[i * step], b = p[(63 - i) * step];
    const float c = p[(31 - i) * step], d = p[(32 + i) * step];
    p[i * step] = a + b;
    p[(32 + i) * step] = (a - b) * WC64_DEV[i];
    p[(31 - i) * step] = c + d;
    p[(63 - i) * step] = (c - d) * WC64_DEV[31 - i];
  }
#pragma unroll
  for (int half = 0; half < 2; half++) {
    float b[32];
    float *q = p + half * 32 * step;
#pragma unroll
    for (int i = 0; i < 32; i++) b[i] = q[i * step];
    dct32<F>(b);
#pragma unroll
    for (int i = 0; i < 32; i++) q[i * step] = b[i];
  }
  p[32 * step] = fmla<F>(p[32 * step], JXG_SQRT2, p[33 * step]);
  for (int i = 33; i < 63; i++) p[i * step] += p[(i + 1) * step];
  if (do_scale)
    for (int i = 0; i < 64; i++) p[i * step] = p[i * step] * scale;
}

// LDS position of frequency f along an axis transformed by dct64_line.
__host__ __device__ __forceinline__ int dct64_pos(int f) { return (f & 1) ? 32 + (f >> 1) : f >> 1; }

// inv_dct1d_2 / inv_dct1d_4 / inv_dct1d_8 (src/dct.rs:1500-1543).
__device__ __forceinline__ void inv_dct2(float *b) {
  const float s = b[0], d = b[1];
  b[0] = (s + d) * 0.5f;
  b[1] = (s - d) * 0.5f;
}
__device__ __forceinline__ void inv_dct4(float *b) {
  const float IS2 = 0.70710678118654752440f;
  float t[4] = {b[0], b[2], b[1], b[3]};
  t[2] = (t[2] - t[3]) * IS2;
  inv_dct2(t + 2);
  t[2] *= 1.0f / 0.5411961f;
  t[3] *= 1.0f / 1.306563f;
  inv_dct2(t);
  b[0] = (t[0] + t[2]) * 0.5f;
  b[3] = (t[0] - t[2]) * 0.5f;
  b[1] = (t[1] + t[3]) * 0.5f;
  b[2] = (t[1] - t[3]) * 0.5f;
}
__device__ __forceinline__ void inv_dct8(float *b) {
  const float IS2 = 0.70710678118654752440f;
  const float iwc[4] = {1.0f / 0.5097956f, 1.0f / 0.6013449f, 1.0f / 0.8999762f,
                        1.0f / 2.5629156f};
  float t[8];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    t[i] = b[2 * i];
    t[4 + i] = b[2 * i + 1];
  }
  t[6] -= t[7];
  t[5] -= t[6];
  t[4] = (t[4] - t[5]) * IS2;
  inv_dct4(t + 4);
#pragma unroll
  for (int i = 0; i < 4; i++) t[4 + i] *= iwc[i];
  inv_dct4(t);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    b[i] = (t[i] + t[4 + i]) * 0.5f;
    b[7 - i] = (t[i] - t[4 + i]) * 0.5f;
  }
}

// inv_dct1d_16 / inv_dct1d_32 (src/dct.rs:1547-1627).
__device__ __forceinline__ void inv_dct16(float *b) {
  const float IS2 = 0.70710678118654752440f;
  const float iwc[8] = {1.0f / 0.5024193f, 1.0f / 0.5224986f, 1.0f / 0.566944f, 1.0f / 0.6468218f,
                        1.0f / 0.7881546f, 1.0f / 1.0606777f, 1.0f / 1.7224471f, 1.0f / 5.1011486f};
  float t[16];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    t[i] = b[2 * i];
    t[8 + i] = b[2 * i + 1];
  }
#pragma unroll
  for (int i = 14; i >= 9; i--) t[i] -= t[i + 1];
  t[8] = (t[8] - t[9]) * IS2;
  inv_dct8(t + 8);
#pragma unroll
  for (int i = 0; i < 8; i++) t[8 + i] *= iwc[i];
  inv_dct8(t);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    b[i] = (t[i] + t[8 + i]) * 0.5f;
    b[15 - i] = (t[i] - t[8 + i]) * 0.5f;
  }
}
__device__ __forceinline__ void inv_dct32(float *b) {
  const float IS2 = 0.70710678118654752440f;
  const float iwc[16] = {1.0f / 0.50060298f, 1.0f / 0.50547096f, 1.0f / 0.5154473f,
                         1.0f / 0.5310426f,  1.0f / 0.5531039f,  1.0f / 0.582935f,
                         1.0f / 0.6225041f,  1.0f / 0.67480834f, 1.0f / 0.7445363f,
                         1.0f / 0.83934965f, 1.0f / 0.97256825f, 1.0f / 1.16944f,
                         1.0f / 1.4841646f,  1.0f / 2.057781f,   1.0f / 3.4076084f,
                         1.0f / 10.190008f};
  float t[32];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    t[i] = b[2 * i];
    t[16 + i] = b[2 * i + 1];
  }
#pragma unroll
  for (int i = 30; i >= 17; i--) t[i] -= t[i + 1];
  t[16] = (t[16] - t[17]) * IS2;
  inv_dct16(t + 16);
#pragma unroll
  for (int i = 0; i < 16; i++) t[16 + i] *= iwc[i];
  inv_dct16(t);
#pragma unroll
  for (int i = 0; i < 16; i++) {
    b[i] = (t[i] + t[16 + i]) * 0.5f;
    b[31 - i] = (t[i] - t[16 + i]) * 0.5f;
  }
}

// RESAMPLE_SCALE_64_TO_8 (src/dct.rs:2005-2014)
static __constant__ float RS64_8_DEV[8] = {
    1.0f, 0.9936866f, 0.97488683f, 0.94401807f, 0.9017642f, 0.8490575f, 0.7870549f, 0.71710813f};

// One line transform of length N held in LDS with element stride `step`,
// optionally scaled afterwards (the last pass of a 2-D assembly multiplies
// by 1/(W*H)).
template <bool F, int N>
__device__ __forceinline__ void dct_line(float *p, int step, float scale, bool do_scale) {
  float b[N];
#pragma unroll
  for (int i = 0; i < N; i++) b[i] = p[i * step];
  if constexpr (N == 8) dct8<F>(b);
  if constexpr (N == 16) dct16<F>(b);
  if constexpr (N == 32) dct32<F>(b);
  if (do_scale) {
#pragma unroll
    for (int i = 0; i < N; i++) b[i] = b[i] * scale;
  }
#pragma unroll
  for (int i = 0; i < N; i++) p[i * step] = b[i];
}

// Line of length n (8, 16, 32, or 64 when BIG).  The 64-point line is only
// compiled into kernels that take DCT64-family transforms, so it does not
// raise the register budget of the others.
template <bool F, bool BIG = false>
__device__ __forceinline__ void dct_line_n(int n, float *p, int step, float scale, bool do_scale) {
  if (n == 8)
    dct_line<F, 8>(p, step, scale, do_scale);
  else if (n == 16)
    dct_line<F, 16>(p, step, scale, do_scale);
  else if (!BIG || n == 32)
    dct_line<F, 32>(p, step, scale, do_scale);
  else if constexpr (BIG)
    dct64_line<F>(p, step, scale, do_scale);
}

// idct1d_4 (src/dct.rs:2144-2168)
template <bool F>
__device__ __forceinline__ void idct4(const float v[4], float out[4]) {
  float t0 = v[0], t1 = v[2], t2 = v[1], t3 = v[3];
  float e0 = t0 + t1, e1 = t0 - t1;
  t3 += t2;
  t2 *= JXG_SQRT2;
  float o0 = t2 + t3, o1 = t2 - t3;
  out[0] = fmla<F>(0.5411961f, o0, e0);
  out[3] = fmla<F>(-0.5411961f, o0, e0);
  out[1] = fmla<F>(1.306563f, o1, e1);
  out[2] = fmla<F>(-1.306563f, o1, e1);
}

// ---- strategy geometry (src/dc_group_data.rs:140-170) ----
// Per-strategy values live in 4-bit fields of two 64-bit constants (codes
// 0..15, then 16..23), so a lookup by a per-lane strategy code is a shift in
// registers; an array indexed that way would be a memory load per lane.
namespace jxg_geom {
constexpr uint8_t kCovX[24] = {1, 1, 2, 2, 4, 1, 1, 1, 2, 4, 1, 1, 1, 1, 8, 4, 8, 1, 1, 16, 8, 16, 1, 4};
constexpr uint8_t kCovY[24] = {1, 2, 1, 2, 4, 1, 1, 1, 4, 2, 1, 1, 1, 1, 8, 8, 4, 1, 1, 16, 16, 8, 4, 1};
constexpr int log2i(int v) { return v <= 1 ? 0 : 1 + log2i(v >> 1); }
constexpr uint64_t pack_log2(const uint8_t *v, int lo) {
  uint64_t r = 0;
  for (int i = 0; i < 16 && lo + i < 24; i++) r |= (uint64_t)log2i(v[lo + i]) << (4 * i);
  return r;
}
constexpr uint64_t pack_kind(int lo) {
  uint64_t r = 0;
  for (int i = 0; i < 16 && lo + i < 24; i++) r |= (uint64_t)jxg_qt_kind_dev(lo + i) << (4 * i);
  return r;
}
constexpr uint64_t kLogCovX[2] = {pack_log2(kCovX, 0), pack_log2(kCovX, 16)};
constexpr uint64_t kLogCovY[2] = {pack_log2(kCovY, 0), pack_log2(kCovY, 16)};
constexpr uint64_t kKind[2] = {pack_kind(0), pack_kind(16)};
constexpr int field(const uint64_t (&w)[2], int s) {
  return (int)((s < 16 ? w[0] >> (4 * s) : w[1] >> (4 * (s - 16))) & 15u);
}
constexpr bool packed_ok() {
  for (int s = 0; s < 24; s++)
    if ((1 << field(kLogCovX, s)) != kCovX[s] || (1 << field(kLogCovY, s)) != kCovY[s] ||
        field(kKind, s) != jxg_qt_kind_dev(s))
      return false;
  return true;
}
static_assert(packed_ok(), "packed strategy geometry");
}  // namespace jxg_geom

__device__ __forceinline__ int nibble24(uint64_t lo, uint64_t hi, int s) {
  return (int)((s < 16 ? lo >> (4 * s) : hi >> (4 * (s - 16))) & 15u);
}
__device__ __forceinline__ int cov_x(int s) {
  return 1 << nibble24(jxg_geom::kLogCovX[0], jxg_geom::kLogCovX[1], s);
}
__device__ __forceinline__ int cov_y(int s) {
  return 1 << nibble24(jxg_geom::kLogCovY[0], jxg_geom::kLogCovY[1], s);
}
// jxg_qt_kind_dev as a register lookup.
__device__ __forceinline__ int qt_kind(int s) {
  return nibble24(jxg_geom::kKind[0], jxg_geom::kKind[1], s);
}

// Whether the 2-D assembly transforms rows first (src/dct.rs: 8x8, 8x16,
// 16x32, 8x32, 32x64) or columns first (16x8, 16x16, 32x32, 32x16, 32x8,
// 64x64, 64x32).
__device__ __forceinline__ bool rows_first(int s) {
  return s == JXG_DCT || s == JXG_DCT8X16 || s == JXG_DCT16X32 || s == JXG_DCT8X32 ||
         s == JXG_DCT32X64;
}

// The DCT64 family (64X64, 64X32, 32X64).
__host__ __device__ __forceinline__ bool is_dct64(int s) {
  return s == JXG_DCT64X64 || s == JXG_DCT64X32 || s == JXG_DCT32X64;
}

// Storage index of the coefficient with vertical frequency v and horizontal
// frequency u of a w x h footprint (see the 2-D assemblies of src/dct.rs):
// the square 8x8 stores u-major, wide shapes v-major, column-first shapes
// u-major.
__device__ __forceinline__ int storage_index(int s, int w, int h, int v, int u) {
  if (s == JXG_DCT) return u * 8 + v;
  if (rows_first(s)) return v * w + u;
  return u * h + v;
}

// Spatial (v, u) of storage index k (inverse of storage_index).
__device__ __forceinline__ void storage_to_spatial(int s, int w, int h, int k, int *v, int *u) {
  if (s == JXG_DCT) {
    *u = k >> 3;
    *v = k & 7;
  } else if (rows_first(s)) {
    *v = k / w;
    *u = k % w;
  } else {
    *u = k / h;
    *v = k % h;
  }
}

// DequantBias::level / level_f32 (src/group.rs:748-790)
__device__ __forceinline__ float dequant_level(float one, float num, int32_t q) {
  uint32_t a = q < 0 ? (uint32_t)(-(int64_t)q) : (uint32_t)q;
  if (a <= 1) return q == 0 ? 0.0f : (q > 0 ? one : -one);
  float qf = (float)q;
  if (a < 64) {
    float t = num / (float)a;
    return qf - (q < 0 ? -t : t);
  }
  return qf - num / qf;
}

__device__ __forceinline__ float level_f32(float one, float num, float q) {
  float aq = fabsf(q);
  if (aq < 1.125f) return q == 0.0f ? 0.0f : copysignf(one, q);
  uint32_t a = (uint32_t)aq;
  if (a < 64 && (float)a == aq) {
    float t = num / (float)a;
    return q - (q < 0.0f ? -t : t);
  }
  return q - num / q;
}

// aggregate_quant (src/ac_strategy/selection.rs:484-498): the quantizer of
// a merged footprint from its blocks' maximum and sum.
__device__ __forceinline__ uint8_t jxg_aggregate_quant(uint8_t max_q, uint32_t sum, int covered,
                                                       float d) {
  if (covered < 4) return max_q;
  const float mixer = fmaxf(1.0f - fmaxf(d - 1.54138f, 0.0f) * 0.56391f, 0.0f);
  const float mean = (float)sum / (float)covered;
  const float mixed = (float)max_q * mixer + mean * (1.0f - mixer);
  return (uint8_t)fminf(fmaxf(mixed + 0.5f, 1.0f), 255.0f);
}

// rate_log2 (src/inflated_cost.rs:52-60) of a level magnitude aq (an
// integer-valued float): the small table, the context's large table, then
// the device log2f beyond both.
__device__ __forceinline__ float jxg_rate_log2(const jxg_derived *D, float aq) {
  const uint32_t ki = aq >= 4294967295.0f ? 0xffffffffu : (uint32_t)aq;
  if (ki < JXG_RATE_LUT_N) return D->rate_lut[ki];
  if (ki < JXG_RATE_BIG_N && D->rate_big) return D->rate_big[ki];
  return log2f(1.0f + aq);
}

}  // namespace jxg
