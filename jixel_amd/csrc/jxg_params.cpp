// jxg_params.cpp - host-side parameters of the front end: quality mapping,
// distance parameters, transfer LUT, quant-matrix blob layout and defaults,
// and the per-frame derived constants the kernels read.  The default quant
// tables are built in jxg_tables.cpp.
//
// Compiled with -ffp-contract=off: every a*b+c stays unfused unless written
// with fma(), matching the reference's Rust float semantics; `fused`
// selects what the reference's fmla (src/dct.rs:33-53) is.
#include <math.h>
#include <string.h>

#include <cmath>
#include <initializer_list>

#include "jxg_internal.h"

namespace {

inline float fmla(bool fused, float a, float b, float c) {
  return fused ? std::fma(a, b, c) : a * b + c;
}

inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

inline float bits_to_f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
inline uint32_t f_to_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

const uint32_t kQtSizes[JXG_QT_NUM_KINDS] = {64, 128, 256, 1024, 512, 256, 4096,
                                             2048, 64, 64, 64, 64, 64};

// ---- f16 round trip (the F16C conversions of src/util.rs:427-518) ----
uint16_t f32_to_f16(float v) {
  uint32_t x = f_to_bits(v);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t a = x & 0x7fffffffu;
  if (a >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (a > 0x7f800000u ? 0x200u : 0));
  if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
  if (a < 0x38800000u) {
    if (a < 0x33000000u) return (uint16_t)sign;
    uint32_t e = a >> 23, m = (a & 0x7fffffu) | 0x800000u;
    uint32_t shift = 126 - e;
    uint32_t h = m >> shift, rem = m & ((1u << shift) - 1), mid = 1u << (shift - 1);
    if (rem > mid || (rem == mid && (h & 1))) h++;
    return (uint16_t)(sign | h);
  }
  uint32_t h = (((a >> 23) - 112) << 10) | ((a & 0x7fffffu) >> 13);
  uint32_t rem = a & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1))) h++;
  return (uint16_t)(sign | h);
}

float f16_to_f32(uint16_t b) {
  uint32_t sign = b >> 15, e = (b >> 10) & 0x1f, m = b & 0x3ff;
  if (e == 0) {
    float v = (1.0f / 16384.0f) * ((float)m * (1.0f / 1024.0f));
    return sign ? -v : v;
  }
  if (e == 31) return bits_to_f((sign << 31) | 0x7f800000u | (m << 13));
  return bits_to_f((sign << 31) | ((e + 112) << 23) | (m << 13));
}

// ---- compute_distance_params pieces (src/frame/quantization.rs:52-149) ----
float dc_refinement(bool fused, float d) {
  const float peak = 1.35f, hold = 3.0f, release = 5.0f;
  if (d <= hold) return peak;
  if (d >= release) return 1.0f;
  const float recip = 1.0f / (release - hold);
  return fmla(fused, 1.0f - peak, (d - hold) * recip, peak);
}

float dcepf_vlq_t(float d) { return clampf((d - 8.0f) / (15.0f - 8.0f), 0.0f, 1.0f); }

float quant_dc(bool fused, float d) {
  float refine = dc_refinement(fused, d);
  float vlq = fmla(fused, dcepf_vlq_t(d), 0.7f - 1.0f, 1.0f);
  float eff = 2.9f * powf(d / 2.9f, 0.57f);
  eff = clampf(eff, 0.5f * d, d);
  float v = fminf(1.12f / eff, 50.0f);
  return v * refine * vlq;
}

float luma_dc_coarsening(bool fused, float d) {
  float t;
  if (d <= 1.25f || d >= 5.0f)
    t = 0.0f;
  else if (d < 1.75f)
    t = (d - 1.25f) / (1.75f - 1.25f);
  else if (d <= 4.25f)
    t = 1.0f;
  else
    t = (5.0f - d) / (5.0f - 4.25f);
  return fmla(fused, t, 1.15f - 1.0f, 1.0f);
}

const float kDcQuant[3] = {1.0f / 4096.0f, 1.0f / 512.0f, 1.0f / 256.0f};
const float kInvDcQuant[3] = {4096.0f, 512.0f, 256.0f};

float signaled_dc_step(int c, float mult) {
  float wire = kDcQuant[c] * 128.0f * mult;
  return f16_to_f32(f32_to_f16(wire)) / (kDcQuant[c] * 128.0f);
}

int32_t sat_i32(float v) {
  if (v != v) return 0;
  if (v >= 2147483648.0f) return INT32_MAX;
  if (v <= -2147483648.0f) return INT32_MIN;
  return (int32_t)v;
}

// ---- AQ schedule (src/adaptive_quant.rs:100-125) ----
const float kAqKnots[7] = {0.5f, 1.0f, 1.4f, 1.75f, 3.0f, 5.0f, 8.0f};
const float kAqStrength[7] = {0.26565f, 0.275459f, 0.058287f, 0.0f, 0.0f, 0.0f, 0.0f};

float aq_schedule(const float v[7], float d) {
  if (d <= kAqKnots[0]) return v[0];
  for (int i = 1; i < 7; i++)
    if (d <= kAqKnots[i]) {
      float f = (d - kAqKnots[i - 1]) / (kAqKnots[i] - kAqKnots[i - 1]);
      return v[i - 1] + f * (v[i] - v[i - 1]);
    }
  return v[6];
}

// ---- quantize_ac_thresholds_scaled (src/group.rs:404-457) ----
void ac_thresholds(int c, int area, float d, float qm_mult, float thr[4]) {
  float normal[4] = {0.58f, 0.635f, 0.66f, 0.7f};
  if (c == 0) {
    float up = clampf((d - 0.5f) / (0.75f - 0.5f), 0.0f, 1.0f);
    float down = clampf((2.0f - d) / (2.0f - 1.5f), 0.0f, 1.0f);
    float w = fminf(up, down);
    normal[0] += w * (0.5f - normal[0]);
    for (int i = 1; i < 4; i++) normal[i] += (1.0f - w) * 0.08f;
  }
  if (c == 2) {
    float bt = qm_mult > 1.0f ? 0.65f : 0.75f;
    for (int i = 1; i < 4; i++) normal[i] = bt;
  }
  if (area > 1) {
    // 0.003 * xsize * ysize: any factorisation of the area gives the same
    // product because both factors are powers of two.
    int xs = 1, ys = 1;
    while (xs * ys < area) {
      if (xs <= ys)
        xs *= 2;
      else
        ys *= 2;
    }
    float delta = clampf(0.003f * (float)xs * (float)ys, 0.0f, c > 0 ? 0.08f : 0.12f);
    for (int i = 0; i < 4; i++) normal[i] -= delta;
  }
  float hq[4];
  if (c == 1) {
    hq[0] = 0.50f;
    hq[1] = 0.51f;
    hq[2] = 0.52f;
    hq[3] = 0.54f;
  } else {
    memcpy(hq, normal, sizeof(hq));
  }
  float t = clampf((d - 0.1f) / 0.9f, 0.0f, 1.0f);
  for (int i = 0; i < 4; i++) thr[i] = hq[i] + t * (normal[i] - hq[i]);
}

}  // namespace

extern "C" {

uint16_t jxg_f32_to_f16(float v) { return f32_to_f16(v); }
float jxg_f16_to_f32(uint16_t b) { return f16_to_f32(b); }

uint32_t jxg_qt_size(int kind) {
  return (kind >= 0 && kind < JXG_QT_NUM_KINDS) ? kQtSizes[kind] : 0;
}

uint32_t jxg_qt_offset(int kind) {
  uint32_t off = 0;
  for (int k = 0; k < kind && k < JXG_QT_NUM_KINDS; k++) off += 6 * kQtSizes[k];
  return off;
}

int jxg_qt_kind_of(int s) {
  switch (s) {
    case JXG_DCT: return JXG_QT_DCT8;
    case JXG_DCT16X8:
    case JXG_DCT8X16: return JXG_QT_DCT16X8;
    case JXG_DCT16X16: return JXG_QT_DCT16X16;
    case JXG_DCT32X32: return JXG_QT_DCT32X32;
    case JXG_DCT32X16:
    case JXG_DCT16X32: return JXG_QT_DCT32X16;
    case JXG_DCT32X8:
    case JXG_DCT8X32: return JXG_QT_DCT8X32;
    case JXG_DCT64X64: return JXG_QT_DCT64X64;
    case JXG_DCT64X32:
    case JXG_DCT32X64: return JXG_QT_DCT64X32;
    case JXG_IDENTITY: return JXG_QT_IDENTITY;
    case JXG_DCT2X2: return JXG_QT_DCT2X2;
    case JXG_DCT4X4: return JXG_QT_DCT4X4;
    case JXG_DCT4X8:
    case JXG_DCT8X4: return JXG_QT_DCT4X8;
    case JXG_AFV0:
    case JXG_AFV1:
    case JXG_AFV2:
    case JXG_AFV3: return JXG_QT_AFV;
    default: return -1;
  }
}

float jxg_distance_from_quality(float quality) {
  float q = quality != quality ? 0.0f : clampf(quality, 0.0f, 100.0f);
  float d;
  if (q >= 99.0f)
    d = 0.05f * powf(2.0f, 100.0f - q);
  else if (q >= 90.0f)
    d = powf(10.0f, (99.0f - q) / 9.0f - 1.0f);
  else if (q >= 30.0f)
    d = 0.1f + (100.0f - q) * 0.09f;
  else
    d = 6.24f + powf(2.5f, (30.0f - q) / 5.0f) / 6.25f;
  return fminf(d, 25.0f);
}

int jxg_frame_params_init(float distance, int speed, int fused, jxg_frame_params *out) {
  if (!out) return JXG_ERR_INVALID_ARG;
  if (!(distance > 0.0f) || !std::isfinite(distance)) return JXG_ERR_INVALID_DISTANCE;
  if (speed < JXG_SPEED_ULTRAFAST || speed > JXG_SPEED_EXTRASLOW) return JXG_ERR_INVALID_ARG;
  const bool f = fused != 0;
  memset(out, 0, sizeof(*out));
  float qdc_fine = quant_dc(f, distance);
  float coarsen = luma_dc_coarsening(f, distance);
  float qdc = qdc_fine / coarsen;
  float s = (float)(1 << 16) * 0.8f / (distance * 10.0f);
  int32_t gs = sat_i32(roundf(s));
  gs = gs < 1 ? 1 : (gs > (1 << 15) ? (1 << 15) : gs);
  float scale = (float)gs / (float)(1 << 16);
  int32_t qd = sat_i32((qdc / scale) + 0.5f);
  qd = qd < 1 ? 1 : (qd > (1 << 16) ? (1 << 16) : qd);
  out->distance = distance;
  out->global_scale = gs;
  out->scale = scale;
  out->quant_dc = qd;
  out->scale_dc = (float)qd * scale;
  uint32_t xq = 2;
  if (distance >= 1.25f) xq++;
  if (distance > 9.0f) xq++;
  if (distance < 0.299f) xq++;
  out->x_qm_scale = xq;
  int32_t it = 0;
  if (distance >= 0.7f) it++;
  if (distance >= 1.5f) it++;
  float t = dcepf_vlq_t(distance);
  if ((distance >= 2.5f && t < 0.5f) || t >= 0.5f) it = 3;
  out->epf_iters = it;
  // chroma_dc_steps (src/frame/quantization.rs:210-217) then
  // signaled_dc_steps (src/color_correlation.rs:1842)
  float growth = fmaxf(quant_dc(f, 3.5f) / qdc_fine, 1.0f);
  float steps[3] = {2.0f * powf(growth, 0.75f - 1.0f), 1.0f, powf(growth, 0.5f - 1.0f)};
  steps[0] = steps[0] / coarsen;
  steps[2] = steps[2] / coarsen;
  for (int c = 0; c < 3; c++) out->dc_step[c] = signaled_dc_step(c, steps[c]);
  out->ytob_dc = 0;
  // EncodingContext defaults (src/encoding_context.rs:609-700)
  out->cfl_base_x = 0.0f;
  out->cfl_base_b = 1.0f;
  out->cfl_color_factor = 84.0f;
  out->b_qm_scale = 2;
  // DequantBias::DEFAULT (src/group.rs:626-632,705-711)
  out->dequant_one[0] = 1.0f - 0.05465007330715401f;
  out->dequant_one[1] = 1.0f - 0.07005449891748593f;
  out->dequant_one[2] = 1.0f - 0.049935103337343655f;
  out->dequant_numerator = 0.145f;
  out->fused = f ? 1 : 0;
  out->speed = speed;
  // EncodeConfig::default: fill_invisible on, alpha_distance unset (lossy alpha)
  out->flags = JXG_FLAG_FILL_INVISIBLE | JXG_FLAG_LOSSY_ALPHA;
  // channel_weights_for_bias at the spec matrix (src/encoding_context.rs:39-51)
  out->channel_weight[0] = 0.30f;
  out->channel_weight[1] = 1.0f;
  out->channel_weight[2] = 0.28f;
  out->arbiter = JXG_ARBITER_SSIMULACRA2;  // EncodeConfig::default (src/encode_image.rs:717)
  // XybMatrix::SPEC (src/xyb.rs:35-47,81-96), scalar kernel order
  out->xyb_order = JXG_XYB_SCALAR;
  jxg_spec_xyb_matrix(out->xyb_fwd);
  return JXG_OK;
}

void jxg_spec_xyb_matrix(float fwd[9]) {
  const float M00 = 0.30f, M02 = 0.078f, M01 = 1.0f - M02 - M00;
  const float M10 = 0.23f, M12 = 0.078f, M11 = 1.0f - M12 - M10;
  const float B_BIAS = 0.55180986f;
  const float B_R_RATIO = 0.24342269f / (0.24342269f + 0.20476745f);
  const float M20 = B_R_RATIO * (1.0f - B_BIAS);
  const float M21 = (1.0f - B_R_RATIO) * (1.0f - B_BIAS);
  const float M22 = 1.0f - M20 - M21;
  const float m[9] = {M00, M01, M02, M10, M11, M12, M20, M21, M22};
  memcpy(fwd, m, sizeof(m));
}

void jxg_srgb_lut(float out[256]) {
  for (int i = 0; i < 256; i++) {
    double v = (double)i * (1.0 / 255.0);
    out[i] = (float)(v <= 0.04045 ? v / 12.92 : pow((v + 0.055) / 1.055, 2.4));
  }
}

void jxg_plane_geometry(uint32_t width, uint32_t height, uint32_t *stride, uint32_t *padded_h) {
  if (stride) *stride = (width + 63u) & ~63u;
  if (padded_h) *padded_h = (height + 63u) & ~63u;
}

int jxg_derive(const jxg_frame_params *p, jxg_derived *o) {
  if (!p || !o) return JXG_ERR_INVALID_ARG;
  if (!(p->distance > 0.0f) || !std::isfinite(p->distance)) return JXG_ERR_INVALID_DISTANCE;
  if (!(p->scale > 0.0f) || !(p->cfl_color_factor != 0.0f)) return JXG_ERR_INVALID_ARG;
  for (int c = 0; c < 3; c++)
    if (!(p->dc_step[c] > 0.0f)) return JXG_ERR_INVALID_ARG;
  const bool f = p->fused != 0;
  memset(o, 0, sizeof(*o));
  float d = p->distance;
  o->distance = d;
  o->scale = p->scale;
  o->inv_scale = 1.0f / p->scale;
  for (int c = 0; c < 3; c++) o->inv_factor[c] = kInvDcQuant[c] / p->dc_step[c] * p->scale_dc;
  o->cfl_factor_b = kInvDcQuant[2] / p->dc_step[2] * (kDcQuant[1] * p->dc_step[1]) *
                    (p->cfl_base_b + (float)p->ytob_dc / p->cfl_color_factor);
  o->cfl_factor_x = kInvDcQuant[0] / p->dc_step[0] * (kDcQuant[1] * p->dc_step[1]) * p->cfl_base_x;
  o->cfl_base_x = p->cfl_base_x;
  o->cfl_base_b = p->cfl_base_b;
  o->cfl_inv_color_factor = 1.0f / p->cfl_color_factor;
  static const float kBQm[6] = {1.0f, 1.25f, 1.5625f, 1.953125f, 2.4414062f, 3.0517578f};
  uint32_t bq = p->b_qm_scale < 2 ? 2 : (p->b_qm_scale > 7 ? 7 : p->b_qm_scale);
  o->qm_mult[0] = powf(1.25f, (float)p->x_qm_scale - 2.0f);
  o->qm_mult[1] = 1.0f;
  o->qm_mult[2] = kBQm[bq - 2];
  for (int c = 0; c < 3; c++) {
    o->dequant_one[c] = p->dequant_one[c];
    o->channel_weight[c] = p->channel_weight[c];
    for (int a = 0; a < JXG_AREA_CLASSES; a++)
      ac_thresholds(c, 1 << a, d, o->qm_mult[c], o->thr[c][a]);
  }
  o->dequant_numerator = p->dequant_numerator;
  // MergeTuning::new (src/ac_strategy.rs:288-350) at the Fast tiers.
  auto banded = [&](float hq, float base) {
    float t = clampf((d - 0.5f) / (1.5f - 0.5f), 0.0f, 1.0f);
    return fmla(f, t, base - hq, hq);
  };
  auto margin = [&](float hq, float base) {
    float fade = clampf((d - 1.0f) / (4.0f - 1.0f), 0.0f, 1.0f);
    return banded(hq, base) * (1.0f - fade * (1.0f - 0.82f));
  };
  bool fast = p->speed == JXG_SPEED_FAST || p->speed == JXG_SPEED_FASTEST;
  o->site_k = fast ? 0.2f : 0.0f;
  o->bias_16x16 = banded(1.2059275f, 1.05f);
  o->bias_32x32 = banded(1.2059275f, 1.06f);
  o->accept_16 = 1.0f - margin(0.19243888f, 0.048f);
  o->accept_32 = 1.0f - margin(0.21398508f, 0.234f);
  o->risk_k = clampf(aq_schedule(kAqStrength, d), 0.0f, 1.0f) > 0.0f ? 1.44269504f : 0.0f;
  float vlq_t = clampf((d - 4.0f) / (7.5f - 4.0f), 0.0f, 1.0f);
  if (vlq_t > 0.0f) {
    o->bias_16x16 = fmla(f, vlq_t, 0.880f - o->bias_16x16, o->bias_16x16);
    o->bias_32x32 = fmla(f, vlq_t, 0.867f - o->bias_32x32, o->bias_32x32);
    o->accept_16 = fmla(f, vlq_t, (1.0f - 0.067f) - o->accept_16, o->accept_16);
    o->accept_32 = fmla(f, vlq_t, (1.0f - 0.02925f) - o->accept_32, o->accept_32);
  }
  if (d >= 1.5f) o->bias_32x32 *= 0.94f;
  // selection_quantizer::enabled (src/ac_strategy/selection_quantizer.rs:46-66)
  o->selection_quantizer =
      (p->arbiter == JXG_ARBITER_SSIMULACRA2 && d >= 1.0f && d <= 4.0f && fast) ? 1 : 0;
  o->run_selection = (d > 0.056713393f && p->speed >= JXG_SPEED_FASTEST) ? 1 : 0;
  // layout_dp_coefficient at Fastest and Fast; layout_dp_unaligned and
  // layout_crossing at Fast (src/effort.rs:311-349)
  o->retile = !o->run_selection ? 0 : (p->speed == JXG_SPEED_FAST ? 2 : (p->speed == JXG_SPEED_FASTEST ? 1 : 0));
  // quant_refine at Fast, quant_refinement_steps (src/ac_strategy/selection.rs:534-542)
  o->refine_steps = (o->run_selection && p->speed == JXG_SPEED_FAST) ? (d < 1.5f ? 0 : (d >= 2.0f ? 2 : 1)) : 0;
  // SearchScope::rerank without full_rerank (src/ac_strategy.rs:71-73,410-412)
  // and MergeTuning::rerank_margin (:320,346-348; src/arbiter.rs:50-62)
  o->rerank = (o->run_selection && p->speed == JXG_SPEED_FAST && d <= 1.7090991223128462f) ? 1 : 0;
  o->rerank_margin = p->arbiter == JXG_ARBITER_SSIMULACRA2 ? 1.0f : 0.0f;
  {
    // reconstruction_lambda (src/ac_strategy.rs:1667-1686)
    const float recip = 1.0f / (4.0f - 1.0f);
    const float ramp = clampf((d - 1.0f) * recip, 0.0f, 1.0f);
    float mult = fmla(f, ramp, 10.0f - 0.25f, 0.25f);
    if (d > 1.0f && d < 2.0f) {
      const float knee = (2.0f - 1.0f) * recip;
      const float t = (d - 1.0f) / (2.0f - 1.0f);
      mult = fmla(f, powf(t, 1.33f), knee * (10.0f - 0.25f), 0.25f);
    }
    o->recon_lambda = 0.08086717f * mult;
  }
  o->fused = f ? 1 : 0;
  for (int k = 0; k < JXG_QT_NUM_KINDS; k++) {
    o->qt_off[k] = jxg_qt_offset(k);
    o->qt_n[k] = kQtSizes[k];
  }
  for (int i = 0; i < 64; i++) o->exp2_table[i] = (float)exp2((double)(i - 32) / 64.0);
  for (int i = 0; i < JXG_RATE_LUT_N; i++) o->rate_lut[i] = log2f(1.0f + (float)i);
  return JXG_OK;
}

}  // extern "C"
