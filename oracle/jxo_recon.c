/*
 * jxo_recon.c - oracle: the reconstruction-domain cost of one transform at
 * one quantizer (reconstruction_dist_and_rate without learned prices and
 * without the hue term, src/ac_strategy.rs:1538-1650 ->
 * src/inflated_cost.rs:1136-1500), the scalar kernels it runs
 * (recon_quantize_scalar::<true, true> :1725-1802, prepare_reconstruction
 * :1686-1723, ssim_deficit_scalar_kernel :1912-1965, ssim_deficit_visible
 * :1844-1910, combine_error_scalar :922-934) and rd_cost with the
 * reconstruction lambda (src/ac_strategy.rs:1667-1700).
 * TEST INFRASTRUCTURE ONLY (see jxo.h).
 */
#include <stdlib.h>
#include <string.h>

#include "jxo.h"
#include "jxo_internal.h"

float jxo_b_qm_mul(uint32_t b_qm_scale);
float jxo_x_qm_mul(uint32_t x_qm_scale);
int jxo_inverse_transform(int s, const float *coef, float *out);

static float level_f32(float one, float numerator, float quant) {
  float aq = fabsf(quant);
  if (aq < 1.125f) return quant == 0.0f ? 0.0f : copysignf(one, quant);
  uint32_t a = (uint32_t)aq;
  if (a < 64 && (float)a == aq) {
    float t = numerator / (float)a;
    return quant - (quant < 0.0f ? -t : t);
  }
  return quant - numerator / quant;
}

static float rate_log2(float qabs) {
  uint32_t k = qabs >= 4294967295.0f ? 0xffffffffu : (uint32_t)qabs;
  return log2f(1.0f + (k < 1024 ? (float)k : qabs));
}

/* recon_quantize_scalar::<true, true>: the biased quantizer's error in the
 * coefficient domain, (scaled - level) * matrix / quant_scale, and the
 * coefficient rate model. */
static float recon_quantize(const float *coef, const float *inv, const float *mat, float quant_scale,
                            float one, float numerator, const float thr[4], int width, int height,
                            int cx, int cy, const uint32_t *scan_pos, float *err) {
  const float inv_scale = 1.0f / quant_scale;
  const int half = width / 2;
  uint32_t nonzero = 0, max_scan = 0;
  float mag = 0.0f;
  for (int y = 0; y < height; y++) {
    const int yfix = y >= height / 2 ? 2 : 0;
    for (int x = 0; x < width; x++) {
      const int i = y * width + x;
      if (x < cx && y < cy) {
        err[i] = 0.0f;
        continue;
      }
      const float threshold = x >= half ? thr[yfix + 1] : thr[yfix];
      const float denominator = inv[i] * quant_scale;
      const float scaled = denominator * coef[i];
      const float quantized = fabsf(scaled) >= threshold ? roundf(scaled) : 0.0f;
      const float to_coefficient = mat[i] * inv_scale;
      err[i] = (scaled - level_f32(one, numerator, quantized)) * to_coefficient;
      if (quantized != 0.0f) {
        nonzero++;
        mag += rate_log2(fabsf(quantized));
        if (scan_pos[i] > max_scan) max_scan = scan_pos[i];
      }
    }
  }
  float visited = 0.0f;
  if (nonzero != 0) {
    const int64_t span = (int64_t)max_scan + 1 - (int64_t)(cx * cy + nonzero);
    visited = span > 0 ? (float)span : 0.0f;
  }
  return (float)nonzero * 1.6f + 1.0f * mag + 0.4f * rate_log2((float)nonzero) + 0.5f * visited;
}

/* ssim_deficit_scalar_kernel: 8x8 windows of a w x h footprint. */
static float ssim_full(const float *o, const float *r, int w, int h) {
  const float C1 = 1e-4f, C2 = 9e-4f, K = 1.0f / 64.0f;
  float deficit = 0.0f;
  for (int wy = 0; wy < h / 8; wy++)
    for (int wx = 0; wx < w / 8; wx++) {
      const int x0 = wx * 8, y0 = wy * 8;
      const float bo = o[y0 * w + x0], br = r[y0 * w + x0];
      float so = 0.0f, sr = 0.0f;
      for (int y = y0; y < y0 + 8; y++)
        for (int x = x0; x < x0 + 8; x++) {
          so += o[y * w + x] - bo;
          sr += r[y * w + x] - br;
        }
      const float mo = bo + so * K, mr = br + sr * K;
      float vo = 0.0f, vr = 0.0f, cv = 0.0f;
      for (int y = y0; y < y0 + 8; y++)
        for (int x = x0; x < x0 + 8; x++) {
          const float co = o[y * w + x] - mo, cr = r[y * w + x] - mr;
          vo = jxo_fmla(co, co, vo);
          vr = jxo_fmla(cr, cr, vr);
          cv = jxo_fmla(co, cr, cv);
        }
      vo *= K;
      vr *= K;
      cv *= K;
      const float lum = (2.0f * mo * mr + C1) / (mo * mo + mr * mr + C1);
      const float st = (2.0f * cv + C2) / (vo + vr + C2);
      deficit += (1.0f - lum * st) * 64.0f;
    }
  return deficit;
}

/* ssim_deficit_visible: the leading vw x vh window of stride-wide planes,
 * clipped windows weighted by their pixel count. */
static float ssim_visible(const float *o, const float *r, int stride, int vw, int vh) {
  const float C1 = 1e-4f, C2 = 9e-4f;
  if (vw == 0 || vh == 0) return 0.0f;
  float deficit = 0.0f;
  for (int y0 = 0; y0 < vh; y0 += 8) {
    const int wh = vh - y0 < 8 ? vh - y0 : 8;
    for (int x0 = 0; x0 < vw; x0 += 8) {
      const int ww = vw - x0 < 8 ? vw - x0 : 8;
      const float count = (float)(ww * wh), inv_count = 1.0f / count;
      const float bo = o[y0 * stride + x0], br = r[y0 * stride + x0];
      float so = 0.0f, sr = 0.0f;
      for (int y = y0; y < y0 + wh; y++)
        for (int x = x0; x < x0 + ww; x++) {
          so += o[y * stride + x] - bo;
          sr += r[y * stride + x] - br;
        }
      const float mo = bo + so * inv_count, mr = br + sr * inv_count;
      float vo = 0.0f, vr = 0.0f, cv = 0.0f;
      for (int y = y0; y < y0 + wh; y++)
        for (int x = x0; x < x0 + ww; x++) {
          const float co = o[y * stride + x] - mo, cr = r[y * stride + x] - mr;
          vo = jxo_fmla(co, co, vo);
          vr = jxo_fmla(cr, cr, vr);
          cv = jxo_fmla(co, cr, cv);
        }
      vo *= inv_count;
      vr *= inv_count;
      cv *= inv_count;
      const float lum = (2.0f * mo * mr + C1) / (mo * mo + mr * mr + C1);
      const float st = (2.0f * cv + C2) / (vo + vr + C2);
      deficit += (1.0f - lum * st) * count;
    }
  }
  return deficit;
}

/* prepare_reconstruction_scalar + the channel's weighted SSIM deficit: the
 * footprint's source pixels (edge-replicated past the image), the
 * reconstruction source - error. */
static float channel_ssim(const float *plane, int width, int height, int px, int py, int pw,
                          int ph, const float *error, float weight) {
  static __thread float orig[64 * 64], rec[64 * 64];
  const int sx = px < width - 1 ? px : width - 1;
  const int copied = pw < width - sx ? pw : width - sx;
  for (int y = 0; y < ph; y++) {
    const int yy = py + y < height - 1 ? py + y : height - 1;
    const float *src = plane + (size_t)yy * width;
    for (int x = 0; x < pw; x++) {
      orig[y * pw + x] = x < copied ? src[sx + x] : src[width - 1];
      rec[y * pw + x] = orig[y * pw + x] - error[y * pw + x];
    }
  }
  const int full = px + pw <= width && py + ph <= height;
  if (full) return weight * ssim_full(orig, rec, pw, ph);
  const int vw = px < width ? (width - px < pw ? width - px : pw) : 0;
  const int vh = py < height ? (height - py < ph ? height - py : ph) : 0;
  return weight * ssim_visible(orig, rec, pw, vw, vh);
}

/* reconstruction_lambda (src/ac_strategy.rs:1667-1686) */
float jxo_reconstruction_lambda(float d) {
  const float RD_LAMBDA = 0.08086717f;
  const float RECIP = 1.0f / (4.0f - 1.0f);
  float ramp = (d - 1.0f) * RECIP;
  ramp = ramp < 0.0f ? 0.0f : (ramp > 1.0f ? 1.0f : ramp);
  float multiplier = jxo_fmla(ramp, 10.0f - 0.25f, 0.25f);
  if (d > 1.0f && d < 2.0f) {
    const float knee_ramp = (2.0f - 1.0f) * RECIP;
    const float t = (d - 1.0f) / (2.0f - 1.0f);
    multiplier = jxo_fmla(powf(t, 1.33f), knee_ramp * (10.0f - 0.25f), 0.25f);
  }
  return RD_LAMBDA * multiplier;
}

/* The coefficients of strategy s at (px, py) with the tile's CfL applied
 * (prepare_strategy_coeffs, src/ac_strategy.rs:1382-1410).  Returns size. */
int jxo_recon_coeffs(const float *xyb, int width, int height, int s, int px, int py,
                     const float cmap[3], float *co /* 3 x 4096 */) {
  const size_t plane = (size_t)width * height;
  int size = 0;
  for (int c = 0; c < 3; c++)
    size = jxo_forward_transform(s, xyb + c * plane, width, height, px, py, co + c * 4096);
  const float ncx = -cmap[0], ncb = -cmap[2];
  for (int i = 0; i < size; i++) {
    co[i] = jxo_fmla(ncx, co[4096 + i], co[i]);
    co[2 * 4096 + i] = jxo_fmla(ncb, co[4096 + i], co[2 * 4096 + i]);
  }
  return size;
}

/* reconstruction_dist_and_rate + rd_cost(Reconstruction, d, 0, ...) for the
 * coefficients `co` of strategy s at (px, py), quantizer qac. */
float jxo_recon_cost(const float *xyb, int width, int height, int s, int px, int py, float qac,
                     const float cmap[3], const float *co, const float *tables,
                     const jxg_frame_params *p, float *dist_out, float *rate_out) {
  const size_t plane = (size_t)width * height;
  int pw, ph;
  jxo_strategy_pixel_dims(s, &pw, &ph);
  const int cxs = jxo_covered_x(s), cys = jxo_covered_y(s);
  const int cx = cxs > cys ? cxs : cys, cy = cxs > cys ? cys : cxs;
  const int width_c = cx * 8, height_c = cy * 8;
  const int kind = jxo_qt_kind(s);
  const uint32_t n = jxo_qt_size(kind), off = jxo_qt_offset(kind);
  const float qm[3] = {jxo_x_qm_mul(p->x_qm_scale), 1.0f, jxo_b_qm_mul(p->b_qm_scale)};
  static __thread uint32_t scan[4096];
  jxo_scan_pos_lut(width_c, height_c, scan);
  float thr[3][4];
  for (int c = 0; c < 3; c++) jxo_ac_thresholds(c, cx, cy, p->distance, qm[c], thr[c]);
  static __thread float err[3][4096], spatial[4096], combined[4096];
  /* luma */
  float rate = 0.0f;
  rate += recon_quantize(co + 4096, tables + off + 4 * n, tables + off + 1 * n, qac * qm[1],
                         p->dequant_one[1], p->dequant_numerator, thr[1], width_c, height_c, cx,
                         cy, scan, err[1]);
  jxo_inverse_transform(s, err[1], spatial);
  const float dist_y = channel_ssim(xyb + plane, width, height, px, py, pw, ph, spatial,
                                    p->channel_weight[1]);
  /* chroma against the luma error */
  for (int c = 0; c <= 2; c += 2) {
    const float factor = cmap[c];
    for (int i = 0; i < width_c * height_c; i++)
      combined[i] = jxo_fmla(factor, err[1][i], co[c * 4096 + i]);
    rate += recon_quantize(combined, tables + off + (3 + c) * n, tables + off + c * n,
                           qac * qm[c], p->dequant_one[c], p->dequant_numerator, thr[c], width_c,
                           height_c, cx, cy, scan, err[c]);
  }
  float distortion = 0.0f;
  for (int c = 0; c < 3; c++) {
    float ch;
    if (c == 1) {
      ch = dist_y;
    } else {
      jxo_inverse_transform(s, err[c], spatial);
      ch = channel_ssim(xyb + c * plane, width, height, px, py, pw, ph, spatial,
                        p->channel_weight[c]);
    }
    distortion += ch;
  }
  if (dist_out) *dist_out = distortion;
  if (rate_out) *rate_out = rate;
  return jxo_fmla(jxo_reconstruction_lambda(p->distance), rate + 0.0f, distortion);
}

/* Test entry points. */
float jxo_ssim_deficit(const float *orig, const float *rec, int stride, int w, int h, int visible) {
  return visible ? ssim_visible(orig, rec, stride, w, h) : ssim_full(orig, rec, w, h);
}
