/*
 * jxo_pixel.c - oracle: scalar helpers, frame parameters, sRGB -> XYB and the
 * adaptive quant field.  TEST INFRASTRUCTURE ONLY (see jxo.h).
 *
 * Constants are written as f32 literals and combined in the reference's
 * order, so compile-time folding gives the same bits as the reference's
 * f32 const evaluation.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "jxo.h"
#include "jxo_internal.h"

int jxo_fused_flag = 1;

void jxo_set_fused(int fused) { jxo_fused_flag = fused ? 1 : 0; }
int jxo_get_fused(void) { return jxo_fused_flag; }

/* ------------------------------------------------------------------ */
/* src/util.rs:285-337 ties-away rounding and saturating conversion.   */

float jxo_round(float x) { return roundf(x); }

int32_t jxo_sat_i32(float v) {
  if (v != v) return 0;
  if (v >= 2147483648.0f) return INT32_MAX;
  if (v <= -2147483648.0f) return INT32_MIN;
  return (int32_t)v;
}

int16_t jxo_sat_i16(float v) {
  if (v != v) return 0;
  if (v >= 32767.0f) return INT16_MAX;
  if (v <= -32768.0f) return INT16_MIN;
  return (int16_t)v;
}

/* round_i32_portable (src/util.rs:325-335): the x86 form of round_i32. */
int32_t jxo_round_i32(float value) {
  float magnitude = fabsf(value);
  float biased = value;
  if (magnitude >= 0.5f && magnitude < 8388608.0f) biased = value + copysignf(0.5f, value);
  return jxo_sat_i32(biased);
}

static float bits_f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint32_t f_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

/* fast_exp (src/util.rs:342-357). */
float jxo_fast_exp(float x) {
  const float LN2_HI = 0.69314575f;
  const float LN2_LO = 1.4286068e-6f;
  int32_t exponent = jxo_round_i32(x * 1.44269504088896340736f);
  float r = jxo_fmla((float)exponent, -LN2_HI, x);
  r = jxo_fmla((float)exponent, -LN2_LO, r);
  float square = r * r;
  float p = jxo_fmla(-0.002775669f, square, 0.16666707f);
  p = jxo_fmla(p, square, 2.0f);
  float scale = bits_f(((uint32_t)(exponent + 127)) << 23);
  return (1.0f + 2.0f * r / (p - r)) * scale;
}

/* ------------------------------------------------------------------ */
/* fast_exp2 (src/adaptive_quant.rs:735-766).  The 64-entry table holds  */
/* 2^((i-32)/64) rounded to f32; it is derived here from the closed form */
/* and checked against the reference's EXP2FT constants in the tests.    */

static float g_exp2_table[64];
static int g_exp2_ready = 0;

static void exp2_table_init(void) {
  if (g_exp2_ready) return;
  for (int i = 0; i < 64; i++) g_exp2_table[i] = (float)exp2((double)(i - 32) / 64.0);
  g_exp2_ready = 1;
}

const float *jxo_exp2_table(void) {
  exp2_table_init();
  return g_exp2_table;
}

float jxo_fast_exp2(float d) {
  exp2_table_init();
  const float redux = bits_f(0x4b400000u) / 64.0f;
  uint32_t ui = f_bits(d + redux);
  uint32_t i0 = ui + 32u;
  uint32_t k = i0 / 64u;
  i0 &= 63u;
  float uf = bits_f(ui);
  uf -= redux;
  float z0 = g_exp2_table[i0];
  float f = d - uf;
  float u = 0.24022650695908768f;
  u = jxo_fmla(u, f, 0.69314718055994973f);
  u *= f;
  float i2 = bits_f(((uint32_t)((int32_t)k + 0x7f)) << 23);
  return jxo_fmla(u, z0, z0) * i2;
}

/* dirty_log2f (src/adaptive_quant.rs:768-784). */
float jxo_dirty_log2f(float d) {
  uint32_t ix = f_bits(d);
  ix += 0x3f800000u - 0x3f3504f3u;
  int32_t n = (int32_t)(ix >> 23) - 0x7f;
  ix = (ix & 0x007fffffu) + 0x3f3504f3u;
  float a = bits_f(ix);
  float x = (a - 1.0f) / (a + 1.0f);
  float x2 = x * x;
  float u = 0.4121985850084821691e+0f;
  u = jxo_fmla(u, x2, 0.5770780163490337802e+0f);
  u = jxo_fmla(u, x2, 0.9617966939259845749e+0f);
  return jxo_fmla(x2 * x, u, jxo_fmla(x, 0.2885390081777926802e+1f, (float)n));
}

/* ------------------------------------------------------------------ */
/* distance_from_quality (src/encode_image.rs:1321-1337).              */
float jxo_distance_from_quality(float quality) {
  float q = quality < 0.0f ? 0.0f : (quality > 100.0f ? 100.0f : quality);
  float d;
  if (q >= 99.0f)
    d = 0.05f * powf(2.0f, 100.0f - q);
  else if (q >= 90.0f)
    d = powf(10.0f, (99.0f - q) / 9.0f - 1.0f);
  else if (q >= 30.0f)
    d = 0.1f + (100.0f - q) * 0.09f;
  else
    d = 6.24f + powf(2.5f, (30.0f - q) / 5.0f) / 6.25f;
  return d < 25.0f ? d : 25.0f;
}

/* ------------------------------------------------------------------ */
/* compute_distance_params and helpers (src/frame/quantization.rs:52-149) */

static float clampf_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

static float dc_refinement(float distance) {
  const float PEAK = 1.35f, HOLD = 3.0f, RELEASE = 5.0f;
  if (distance <= HOLD) return PEAK;
  if (distance >= RELEASE) return 1.0f;
  const float RECIP = 1.0f / (RELEASE - HOLD);
  float t = (distance - HOLD) * RECIP;
  return jxo_fmla(1.0f - PEAK, t, PEAK);
}

static float dcepf_vlq_t(float distance) {
  return clampf_((distance - 8.0f) / (15.0f - 8.0f), 0.0f, 1.0f);
}

static float quant_dc(float distance) {
  float refine = dc_refinement(distance);
  float vlq = jxo_fmla(dcepf_vlq_t(distance), 0.7f - 1.0f, 1.0f);
  float effective = 2.9f * powf(distance / 2.9f, 0.57f);
  effective = clampf_(effective, 0.5f * distance, distance);
  float v = 1.12f / effective;
  v = v < 50.0f ? v : 50.0f;
  return v * refine * vlq;
}

static float luma_dc_coarsening(float distance) {
  const float a = 1.25f, b = 1.75f, c = 4.25f, e = 5.0f;
  float t;
  if (distance <= a || distance >= e)
    t = 0.0f;
  else if (distance < b)
    t = (distance - a) / (b - a);
  else if (distance <= c)
    t = 1.0f;
  else
    t = (e - distance) / (e - c);
  return jxo_fmla(t, 1.15f - 1.0f, 1.0f);
}

/* Round-to-nearest-even f32 -> binary16 bits and back (the F16C
 * conversions src/util.rs:427-518 dispatches to on x86-64). */
uint16_t jxo_f32_to_f16(float v) {
  uint32_t x = f_bits(v);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t absx = x & 0x7fffffffu;
  if (absx >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (absx > 0x7f800000u ? 0x200u : 0));
  if (absx >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* overflow */
  if (absx < 0x38800000u) {                                    /* subnormal half */
    if (absx < 0x33000000u) return (uint16_t)sign;
    uint32_t e = absx >> 23;
    uint32_t m = (absx & 0x7fffffu) | 0x800000u;
    uint32_t shift = 126 - e; /* 14 - (e - 112) + 10 ... */
    uint32_t half = m >> (shift);
    uint32_t rem = m & ((1u << shift) - 1);
    uint32_t mid = 1u << (shift - 1);
    if (rem > mid || (rem == mid && (half & 1))) half++;
    return (uint16_t)(sign | half);
  }
  uint32_t e = (absx >> 23) - 112;
  uint32_t m = absx & 0x7fffffu;
  uint32_t h = (e << 10) | (m >> 13);
  uint32_t rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1))) h++;
  return (uint16_t)(sign | h);
}

float jxo_f16_to_f32(uint16_t b) {
  uint32_t sign = (uint32_t)(b >> 15);
  uint32_t e = (b >> 10) & 0x1f;
  uint32_t m = b & 0x3ff;
  if (e == 0) {
    float v = (1.0f / 16384.0f) * ((float)m * (1.0f / 1024.0f));
    return sign ? -v : v;
  }
  if (e == 31) return bits_f((sign << 31) | 0x7f800000u | (m << 13));
  return bits_f((sign << 31) | ((e + 112) << 23) | (m << 13));
}

static const float DC_QUANT[3] = {1.0f / 4096.0f, 1.0f / 512.0f, 1.0f / 256.0f};

/* signaled_dc_step (src/color_correlation.rs:1835-1838). */
static float signaled_dc_step(int c, float multiplier) {
  float wire = DC_QUANT[c] * 128.0f * multiplier;
  return jxo_f16_to_f32(jxo_f32_to_f16(wire)) / (DC_QUANT[c] * 128.0f);
}

void jxo_distance_params(float distance, int32_t *global_scale, float *scale,
                         int32_t *quant_dc_out, float *scale_dc, uint32_t *x_qm_scale,
                         uint32_t *epf_iters, float dc_step[3]) {
  float qdc_fine = quant_dc(distance);
  float coarsen = luma_dc_coarsening(distance);
  float qdc = qdc_fine / coarsen;
  float s = (float)(1 << 16) * 0.8f / (distance * 10.0f);
  int32_t gs = jxo_sat_i32(roundf(s));
  gs = gs < 1 ? 1 : (gs > (1 << 15) ? (1 << 15) : gs);
  float scale_f = (float)gs / (float)(1 << 16);
  int32_t qd = jxo_sat_i32((qdc / scale_f) + 0.5f);
  qd = qd < 1 ? 1 : (qd > (1 << 16) ? (1 << 16) : qd);
  *global_scale = gs;
  *scale = scale_f;
  *quant_dc_out = qd;
  *scale_dc = (float)qd * scale_f;
  uint32_t xq = 2;
  if (distance >= 1.25f) xq += 1;
  if (distance > 9.0f) xq += 1;
  if (distance < 0.299f) xq += 1;
  *x_qm_scale = xq;
  uint32_t it = 0;
  if (distance >= 0.7f) it++;
  if (distance >= 1.5f) it++;
  /* epf_pass0_scale is Some for 2.5 <= d with vlq_t < 0.5 */
  float t = dcepf_vlq_t(distance);
  int pass0 = distance >= 2.5f && t < 0.5f;
  if (pass0 || t >= 0.5f) it = 3;
  *epf_iters = it;
  /* chroma_dc_steps (src/frame/quantization.rs:210-217) */
  float growth = quant_dc(3.5f) / qdc_fine;
  if (growth < 1.0f) growth = 1.0f;
  float steps[3] = {2.0f * powf(growth, 0.75f - 1.0f), 1.0f, powf(growth, 0.5f - 1.0f)};
  steps[0] = steps[0] / coarsen;
  steps[2] = steps[2] / coarsen;
  for (int c = 0; c < 3; c++) dc_step[c] = signaled_dc_step(c, steps[c]);
}

/* ------------------------------------------------------------------ */
/* sRGB transfer LUT (src/color.rs:35-45): f64 math, one f32 rounding.  */
void jxo_srgb_lut(float out[256]) {
  for (int i = 0; i < 256; i++) {
    double v = (double)i * (1.0 / 255.0);
    double lin = v <= 0.04045 ? v / 12.92 : pow((v + 0.055) / 1.055, 2.4);
    out[i] = (float)lin;
  }
}

/* ------------------------------------------------------------------ */
/* XYB (src/xyb.rs:35-47,141-185; src/avx/xyb.rs:117-262).             */

#define OPSIN_BIAS 0.0037930734f
#define NEG_BIAS_CBRT (-0.1559542f)

static float halley_refine(float x, float a) {
  float tx = x * x * x;
  return x * jxo_fmla(2.0f, a, tx) / jxo_fmla(2.0f, tx, a);
}

float jxo_cbrtf(float x) {
  if (x == 0.0f) return x;
  const uint32_t B1 = 709958130u;
  uint32_t ui = f_bits(x);
  uint32_t hx = ui & 0x7fffffffu;
  hx = hx / 3u + B1;
  ui &= 0x80000000u;
  ui |= hx;
  float t = bits_f(ui);
  t = halley_refine(t, x);
  return halley_refine(t, x);
}

/* AVX2 production cube root of a non-negative value: seed (hx*341)>>10,
 * Halley x*(num/den) with fused num/den, zero mapped to zero at the end. */
static float cbrt_avx2_positive(float a) {
  uint32_t hx = f_bits(a) & 0x7fffffffu;
  uint32_t seed = (uint32_t)(((uint64_t)hx * 341u) >> 10) + 709958130u;
  float x = bits_f(seed);
  for (int it = 0; it < 2; it++) {
    float tx = (x * x) * x;
    float num = fmaf(a, 2.0f, tx);
    float den = fmaf(tx, 2.0f, a);
    x = x * (num / den);
  }
  return a == 0.0f ? 0.0f : x;
}

void jxo_xyb_pixel(const float m[9], int order, float r, float g, float b, float out[3]) {
  float tm[3];
  for (int row = 0; row < 3; row++) {
    const float *mr = m + 3 * row;
    if (order == JXG_XYB_AVX2) {
      float mixed = fmaf(b, mr[2], OPSIN_BIAS);
      mixed = fmaf(g, mr[1], mixed);
      mixed = fmaf(r, mr[0], mixed);
      mixed = fmaxf(mixed, 0.0f);
      tm[row] = cbrt_avx2_positive(mixed) + NEG_BIAS_CBRT;
    } else {
      float mixed = jxo_fmla(mr[0], r, jxo_fmla(mr[1], g, jxo_fmla(mr[2], b, OPSIN_BIAS)));
      tm[row] = jxo_cbrtf(fmaxf(mixed, 0.0f)) + NEG_BIAS_CBRT;
    }
  }
  if (order == JXG_XYB_AVX2) {
    out[0] = (tm[0] - tm[1]) * 0.5f;
    out[1] = (tm[0] + tm[1]) * 0.5f;
  } else {
    out[0] = 0.5f * (tm[0] - tm[1]);
    out[1] = 0.5f * (tm[0] + tm[1]);
  }
  out[2] = tm[2];
}

void jxo_xyb_image(const uint8_t *rgb, size_t row_bytes, int channels, int width, int height,
                   const float m[9], int order, float *xyb) {
  float lut[256];
  jxo_srgb_lut(lut);
  size_t plane = (size_t)width * height;
  for (int y = 0; y < height; y++) {
    const uint8_t *row = rgb + (size_t)y * row_bytes;
    for (int x = 0; x < width; x++) {
      const uint8_t *px = row + (size_t)x * channels;
      float o[3];
      if (channels <= 2) {
        /* linearize_gray (src/encode_image.rs:1571-1599), then
         * snap_achromatic_xyb (src/frame/encoder.rs:347-357) */
        jxo_xyb_pixel(m, order, lut[px[0]], lut[px[0]], lut[px[0]], o);
        o[0] = 0.0f;
        o[2] = o[1];
      } else {
        jxo_xyb_pixel(m, order, lut[px[0]], lut[px[1]], lut[px[2]], o);
      }
      size_t i = (size_t)y * width + x;
      xyb[i] = o[0];
      xyb[plane + i] = o[1];
      xyb[2 * plane + i] = o[2];
    }
  }
}

/* ------------------------------------------------------------------ */
/* Adaptive quant field (src/adaptive_quant.rs:41-206,410-470,547-729). */

#define K_SG_MUL 226.77216153508914f
#define K_SG_MUL2 (1.0f / 73.377132366608819f)
#define K_LOG2 0.693147181f
#define K_SG_RET_MUL (K_SG_MUL2 * 18.6580932135f * K_LOG2)
#define K_SG_V_OFFSET 7.7825991679894591f
#define MATCH_GAMMA_OFFSET 0.019f

static float ratio_cubic_to_simple_gamma(float v, int invert) {
  const float k_epsilon = 1e-2f;
  if (v < 0.0f) v = 0.0f;
  const float k_num_mul = K_SG_RET_MUL * 3.0f * K_SG_MUL;
  const float k_v_offset = K_SG_V_OFFSET * K_LOG2 + k_epsilon;
  const float k_den_mul = K_LOG2 * K_SG_MUL;
  float v2 = v * v;
  float num = k_num_mul * v2 + k_epsilon;
  float den = (k_den_mul * v) * v2 + k_v_offset;
  return invert ? num / den : den / num;
}

static float compute_mask(float out_val) {
  const float k_base = -0.699765f;
  const float k_mul4 = 9.4708735624378946f;
  const float k_mul2 = 17.35036561631863f;
  const float k_offset2 = 302.59587815579727f;
  const float k_mul3 = 6.7943250517376494f;
  const float k_offset3 = 3.7179635626140772f;
  const float k_offset4 = 0.25f * k_offset3;
  const float k_mul0 = 1.07267f;
  float v1 = fmaxf(out_val * k_mul0, 1e-3f);
  float v2 = 1.0f / (v1 + k_offset2);
  float v3 = 1.0f / (v1 * v1 + k_offset3);
  float v4 = 1.0f / (v1 * v1 + k_offset4);
  return jxo_fmla(k_mul4, v4, jxo_fmla(k_mul2, v2, jxo_fmla(k_mul3, v3, k_base)));
}

static float masking_sqrt(float v) {
  const float k_log_offset = 27.505837037000106f;
  const float k_mul = 211.66567973503678f;
  const float mul_v = k_mul * 1e8f;
  return 0.25f * sqrtf(jxo_fmla(v, sqrtf(mul_v), k_log_offset));
}

static float hf_modulation_strength(float distance) {
  return clampf_(1.06836f * distance, 0.1f, 0.890398f);
}

static const float AQ_KNOTS[7] = {0.5f, 1.0f, 1.4f, 1.75f, 3.0f, 5.0f, 8.0f};
static const float AQ_STRENGTH[7] = {0.26565f, 0.275459f, 0.058287f, 0.0f, 0.0f, 0.0f, 0.0f};
static const float AQ_BASE_LEVEL[7] = {0.636493f, 0.648351f, 0.606759f, 0.48f,
                                       0.48f,     0.48f,     0.48f};

static float aq_schedule(const float values[7], float distance) {
  if (distance <= AQ_KNOTS[0]) return values[0];
  for (int i = 1; i < 7; i++) {
    if (distance <= AQ_KNOTS[i]) {
      float f = (distance - AQ_KNOTS[i - 1]) / (AQ_KNOTS[i] - AQ_KNOTS[i - 1]);
      return values[i - 1] + f * (values[i] - values[i - 1]);
    }
  }
  return values[6];
}

float jxo_aq_dampen(float distance) { return clampf_(aq_schedule(AQ_STRENGTH, distance), 0.0f, 1.0f); }

static inline float px_at(const float *plane, int width, int x, int y) {
  return plane[(size_t)y * width + x];
}

static int mini(int a, int b) { return a < b ? a : b; }
static int maxi(int a, int b) { return a > b ? a : b; }

static float hf_modulation(int x, int y, const float *py_, int xs, int ys, float out_val,
                           float strength) {
  const float valmin_y = 0.0206f;
  float sum_y = 0.0f;
  for (int dy = 0; dy < 8; dy++) {
    int ry = mini(y + dy, ys - 1);
    int ry_next = dy == 7 ? ry : mini(y + dy + 1, ys - 1);
    for (int dx = 0; dx < 8; dx++) {
      int c0 = mini(x + dx, xs - 1);
      float p = px_at(py_, xs, c0, ry);
      if (dx < 7) sum_y += fminf(valmin_y, fabsf(p - px_at(py_, xs, mini(x + dx + 1, xs - 1), ry)));
      sum_y += fminf(valmin_y, fabsf(p - px_at(py_, xs, c0, ry_next)));
    }
  }
  float k_mul_y = -0.506626f * strength;
  const float k_offset = 0.655866f;
  return jxo_fmla(sum_y, k_mul_y, k_offset) + out_val;
}

static float blue_modulation(int x, int y, const float *xyb, int xs, int ys, float out_val) {
  size_t plane = (size_t)xs * ys;
  const float k_limit = 0.010474084867598155f;
  const float k_offset = 0.0031994768654636393f;
  float sum = 0.0f;
  for (int dy = 0; dy < 8; dy++) {
    int ry = mini(y + dy, ys - 1);
    for (int dx = 0; dx < 8; dx++) {
      int cx = mini(x + dx, xs - 1);
      size_t i = (size_t)ry * xs + cx;
      float p_x = xyb[i];
      float p_b = xyb[2 * plane + i];
      float p_y_eff = xyb[plane + i] + k_offset + fabsf(p_x);
      if (p_b > p_y_eff) sum += fminf(p_b - p_y_eff, k_limit);
    }
  }
  const float k_mul = 0.127191f;
  float s = sum;
  if (s >= 32.0f * k_limit) s = 64.0f * k_limit - s;
  const float k_max_limit = 15.463398341612438f;
  if (s >= k_max_limit * k_limit) s = k_max_limit * k_limit;
  s *= k_mul;
  return s + out_val;
}

static float gamma_modulation(int x, int y, const float *xyb, int xs, int ys, float out_val) {
  size_t plane = (size_t)xs * ys;
  const float k_bias = 0.16f;
  float overall = 0.0f;
  for (int dy = 0; dy < 8; dy++) {
    int ry = mini(y + dy, ys - 1);
    for (int dx = 0; dx < 8; dx++) {
      int cx = mini(x + dx, xs - 1);
      size_t i = (size_t)ry * xs + cx;
      float iny = xyb[plane + i] + k_bias;
      float inx = xyb[i];
      float r = iny - inx;
      float g = iny + inx;
      float rr = ratio_cubic_to_simple_gamma(r, 1);
      float rg = ratio_cubic_to_simple_gamma(g, 1);
      overall = jxo_fmla(0.5f, rr + rg, overall);
    }
  }
  overall *= 1.0f / 64.0f;
  const float k_gam = 0.119146f;
  return jxo_fmla(k_gam, jxo_dirty_log2f(overall), out_val);
}

static void sort4(float a[4]) {
#define CSWAP(i, j)    \
  if (a[i] > a[j]) {   \
    float t = a[i];    \
    a[i] = a[j];       \
    a[j] = t;          \
  }
  CSWAP(0, 1);
  CSWAP(2, 3);
  CSWAP(0, 2);
  CSWAP(1, 3);
  CSWAP(1, 2);
#undef CSWAP
}

static void store_min4(float v, float m[4]) {
  if (v < m[3]) {
    if (v < m[0]) {
      m[3] = m[2];
      m[2] = m[1];
      m[1] = m[0];
      m[0] = v;
    } else if (v < m[1]) {
      m[3] = m[2];
      m[2] = m[1];
      m[1] = v;
    } else if (v < m[2]) {
      m[3] = m[2];
      m[2] = v;
    } else {
      m[3] = v;
    }
  }
}

/* One DC group; `fq`, when not NULL, receives the field before it is scaled
 * and rounded (the value `q` below), for tests of the float chain. */
static void fill_quant_field_group(const float *xyb, int img_xsize, int img_ysize, int x0, int y0,
                                   int xblocks, int yblocks, float distance, float inv_scale,
                                   uint8_t *qf, float *fq) {
  const float *py_ = xyb + (size_t)img_xsize * img_ysize;
  float scale = 0.8294f / distance;
  int region_w = xblocks * 8, region_h = yblocks * 8;
  int pre_w = region_w / 4, pre_h = region_h / 4;
  float *row_acc = (float *)calloc((size_t)region_w, sizeof(float));
  float *pre = (float *)calloc((size_t)pre_w * pre_h, sizeof(float));
  float *aq_map = (float *)calloc((size_t)xblocks * yblocks, sizeof(float));
  for (int ry = 0; ry < region_h; ry++) {
    int gy = y0 + ry;
    int gy_c = mini(maxi(gy, 0), img_ysize - 1);
    int gy1 = mini(maxi(gy - 1, 0), img_ysize - 1);
    int gy2 = mini(maxi(gy + 1, 0), img_ysize - 1);
    for (int rx = 0; rx < region_w; rx++) {
      int gx = x0 + rx;
      int gx_c = mini(maxi(gx, 0), img_xsize - 1);
      int gx1 = mini(maxi(gx - 1, 0), img_xsize - 1);
      int gx2 = mini(maxi(gx + 1, 0), img_xsize - 1);
      float in_y = px_at(py_, img_xsize, gx_c, gy_c);
      float base = 0.25f * (px_at(py_, img_xsize, gx_c, gy2) + px_at(py_, img_xsize, gx_c, gy1) +
                            px_at(py_, img_xsize, gx1, gy_c) + px_at(py_, img_xsize, gx2, gy_c));
      float gammac = ratio_cubic_to_simple_gamma(in_y + MATCH_GAMMA_OFFSET, 0);
      float diff = gammac * (in_y - base);
      diff *= diff;
      if (diff >= 0.2f) diff = 0.2f;
      diff = masking_sqrt(diff);
      if ((ry & 3) != 0)
        row_acc[rx] += diff;
      else
        row_acc[rx] = diff;
    }
    if (ry % 4 == 3) {
      float *prow = pre + (size_t)(ry / 4) * pre_w;
      for (int p = 0; p < pre_w; p++)
        prow[p] = (row_acc[p * 4] + row_acc[p * 4 + 1] + row_acc[p * 4 + 2] + row_acc[p * 4 + 3]) *
                  0.25f;
    }
  }
  float fe_mul = distance < 2.0f ? (2.0f - distance) * 0.5f : 0.0f;
  const float fe_base[4] = {0.125f, 0.1f, 0.09f, 0.06f};
  const float fe_add[4] = {0.0f, -0.1f, -0.09f, -0.06f};
  float k_mul[4];
  float norm_sum = 0.0f;
  for (int i = 0; i < 4; i++) {
    k_mul[i] = fe_base[i] + fe_mul * fe_add[i];
    norm_sum += k_mul[i];
  }
  const float k_total = 0.29959705784054957f;
  for (int i = 0; i < 4; i++) k_mul[i] *= k_total / norm_sum;
  for (int fy = 0; fy < pre_h; fy++) {
    int ym1 = fy >= 1 ? fy - 1 : fy;
    int yp1 = fy + 1 < pre_h ? fy + 1 : fy;
    const float *rowt = pre + (size_t)ym1 * pre_w;
    const float *row = pre + (size_t)fy * pre_w;
    const float *rowb = pre + (size_t)yp1 * pre_w;
    for (int fx = 0; fx < pre_w; fx++) {
      int xm1 = fx >= 1 ? fx - 1 : fx;
      int xp1 = fx + 1 < pre_w ? fx + 1 : fx;
      float mins[4] = {row[fx], row[xm1], row[xp1], rowt[xm1]};
      sort4(mins);
      store_min4(rowt[fx], mins);
      store_min4(rowt[xp1], mins);
      store_min4(rowb[xm1], mins);
      store_min4(rowb[fx], mins);
      store_min4(rowb[xp1], mins);
      float v = k_mul[0] * mins[0] + k_mul[1] * mins[1] + k_mul[2] * mins[2] + k_mul[3] * mins[3];
      size_t idx = (size_t)(fy / 2) * xblocks + fx / 2;
      if (fx % 2 == 0 && fy % 2 == 0)
        aq_map[idx] = v;
      else
        aq_map[idx] += v;
    }
  }
  float base_level = aq_schedule(AQ_BASE_LEVEL, distance) * scale;
  float dampen = jxo_aq_dampen(distance);
  float mul = scale * dampen;
  float add = (1.0f - dampen) * base_level;
  float strength = hf_modulation_strength(distance);
  for (int by = 0; by < yblocks; by++) {
    int pyy = y0 + by * 8;
    for (int bx = 0; bx < xblocks; bx++) {
      int pxx = x0 + bx * 8;
      uint8_t *out = qf + (size_t)by * xblocks + bx;
      if (pxx >= img_xsize || pyy >= img_ysize) {
        *out = 1;
        if (fq) fq[(size_t)by * xblocks + bx] = 0.0f;
        continue;
      }
      int bx_px = mini(pxx, maxi(img_xsize - 8, 0));
      int by_px = mini(pyy, maxi(img_ysize - 8, 0));
      float aq = aq_map[(size_t)by * xblocks + bx];
      float mask_val = compute_mask(aq);
      mask_val = gamma_modulation(bx_px, by_px, xyb, img_xsize, img_ysize, mask_val);
      float out_val = hf_modulation(bx_px, by_px, py_, img_xsize, img_ysize, mask_val, strength);
      float blue = blue_modulation(bx_px, by_px, xyb, img_xsize, img_ysize, mask_val);
      out_val = fminf(out_val, blue);
      float q = jxo_fast_exp2(out_val * 1.442695041f) * mul + add;
      if (fq) fq[(size_t)by * xblocks + bx] = q;
      int32_t qi = jxo_sat_i32(jxo_fmla(q, inv_scale, 0.5f));
      qi = qi < 1 ? 1 : (qi > 255 ? 255 : qi);
      *out = (uint8_t)qi;
    }
  }
  free(row_acc);
  free(pre);
  free(aq_map);
}

void jxo_fill_quant_field_group(const float *xyb, int img_xsize, int img_ysize, int x0, int y0,
                                int xblocks, int yblocks, float distance, float inv_scale,
                                uint8_t *qf) {
  fill_quant_field_group(xyb, img_xsize, img_ysize, x0, y0, xblocks, yblocks, distance, inv_scale,
                         qf, NULL);
}

void jxo_quant_field_float_image(const float *xyb, int width, int height, float distance,
                                 float *fq) {
  int xb_total = (width + 7) / 8;
  for (int gy = 0; gy * 2048 < height; gy++) {
    for (int gx = 0; gx * 2048 < width; gx++) {
      int x0 = gx * 2048, y0 = gy * 2048;
      int xs = mini(2048, width - x0), ys = mini(2048, height - y0);
      int xb = (xs + 7) / 8, yb = (ys + 7) / 8;
      uint8_t *tmp = (uint8_t *)malloc((size_t)xb * yb);
      float *ftmp = (float *)malloc((size_t)xb * yb * sizeof(float));
      fill_quant_field_group(xyb, width, height, x0, y0, xb, yb, distance, 1.0f, tmp, ftmp);
      for (int by = 0; by < yb; by++)
        memcpy(fq + (size_t)(y0 / 8 + by) * xb_total + x0 / 8, ftmp + (size_t)by * xb,
               xb * sizeof(float));
      free(tmp);
      free(ftmp);
    }
  }
}

void jxo_fill_quant_field_image(const float *xyb, int width, int height, float distance,
                                float inv_scale, uint8_t *qf) {
  int xb_total = (width + 7) / 8;
  for (int gy = 0; gy * 2048 < height; gy++) {
    for (int gx = 0; gx * 2048 < width; gx++) {
      int x0 = gx * 2048, y0 = gy * 2048;
      int xs = mini(2048, width - x0), ys = mini(2048, height - y0);
      int xb = (xs + 7) / 8, yb = (ys + 7) / 8;
      uint8_t *tmp = (uint8_t *)malloc((size_t)xb * yb);
      jxo_fill_quant_field_group(xyb, width, height, x0, y0, xb, yb, distance, inv_scale, tmp);
      for (int by = 0; by < yb; by++)
        memcpy(qf + (size_t)(y0 / 8 + by) * xb_total + x0 / 8, tmp + (size_t)by * xb, xb);
      free(tmp);
    }
  }
}

/* ------------------------------------------------------------------ */
/* Chroma-HF protection (apply_chroma_hf_protection_with_stats,         */
/* src/adaptive_quant.rs:344-408, stats chroma_hf_stats_scalar :222-279) */

void jxo_chroma_hf_stats(const float *xyb, int width, int height, int px, int py, int w, int h,
                         float out[3], int *edges) {
  const size_t plane = (size_t)width * height;
  const float *X = xyb, *Y = xyb + plane, *B = xyb + 2 * plane;
  float chroma_sum = 0.0f, chroma_grad = 0.0f, luma_grad = 0.0f;
  int e = 0;
  for (int y = py; y < py + h; y++)
    for (int x = px; x < px + w; x++) {
      const size_t i = (size_t)y * width + x;
      chroma_sum += fabsf(X[i]) + fabsf(B[i] - Y[i]);
    }
  if (w > 1) {
    for (int y = py; y < py + h; y++)
      for (int x = px; x + 1 < px + w; x++) {
        const size_t i = (size_t)y * width + x, j = i + 1;
        chroma_grad += fabsf(X[j] - X[i]) + fabsf((B[j] - Y[j]) - (B[i] - Y[i]));
        luma_grad += fabsf(Y[j] - Y[i]);
      }
    e += (w - 1) * h;
  }
  if (h > 1) {
    for (int y = py; y + 1 < py + h; y++)
      for (int x = px; x < px + w; x++) {
        const size_t i = (size_t)y * width + x, j = i + (size_t)width;
        chroma_grad += fabsf(X[j] - X[i]) + fabsf((B[j] - Y[j]) - (B[i] - Y[i]));
        luma_grad += fabsf(Y[j] - Y[i]);
      }
    e += w * (h - 1);
  }
  out[0] = chroma_sum;
  out[1] = chroma_grad;
  out[2] = luma_grad;
  *edges = e;
}

static float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

float jxo_chroma_hf_strength(float distance) {
  const float fade_in_start = 0.35f, fade_in_width = 0.40f, fade_out_start = 1.0f,
              fade_out_width = 1.0f, strength = 0.40f;
  if (distance < fade_in_start) return 0.0f;
  const float fade_in = clamp01((distance - fade_in_start) / fade_in_width);
  const float fade_out = 1.0f - clamp01((distance - fade_out_start) / fade_out_width);
  return strength * fade_in * fade_out;
}

void jxo_chroma_hf_protection(const float *xyb, int width, int height, uint8_t *qf,
                              float distance) {
  const float ceil_base = 0.35f, ceil_share = 1.4f;
  const float ds = jxo_chroma_hf_strength(distance);
  if (ds <= 0.0f) return;
  const int xb = (width + 7) / 8, yb = (height + 7) / 8;
  for (int by = 0; by < yb; by++)
    for (int bx = 0; bx < xb; bx++) {
      const int px = bx * 8, py = by * 8;
      const int w = mini(width - px, 8), h = mini(height - py, 8);
      float s[3];
      int edges;
      jxo_chroma_hf_stats(xyb, width, height, px, py, w, h, s, &edges);
      if (edges == 0) continue;
      const float mean_chroma = s[0] / (float)(w * h);
      const float mean_gradient = s[1] / (float)edges;
      const float chroma_share = s[1] / (s[1] + s[2] + 1e-4f);
      const float chroma_w = clamp01((mean_chroma - 0.025f) / 0.075f);
      const float gradient_w = clamp01((mean_gradient - 0.006f) / 0.030f);
      const float share_w = clamp01((chroma_share - 0.25f) / 0.55f);
      const float risk = chroma_w * gradient_w * share_w;
      if (risk > 0.0f) {
        const float ceiling = jxo_fmla(ceil_share, share_w * share_w, ceil_base);
        const float gain = jxo_fmla(ceiling * ds, risk, 1.0f);
        uint8_t *q = qf + (size_t)by * xb + bx;
        float v = roundf((float)*q * gain);
        v = v < 1.0f ? 1.0f : (v > 255.0f ? 255.0f : v);
        *q = (uint8_t)v;
      }
    }
}
