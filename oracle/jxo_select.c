/*
 * jxo_select.c - oracle: coefficient-model strategy costs and the
 * square-merge transform selection (src/ac_strategy.rs,
 * src/ac_strategy/selection.rs, src/inflated_cost.rs, src/coeff_order.rs),
 * plus the whole-front-end CPU driver used as the benchmark baseline.
 * TEST INFRASTRUCTURE ONLY (see jxo.h).
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "jxo.h"
#include "jxo_internal.h"

float jxo_b_qm_mul(uint32_t b_qm_scale);
float jxo_x_qm_mul(uint32_t x_qm_scale);
void jxo_forward_quantize_rows(const float *xyb, int width, int height, const uint8_t *cells,
                               const uint8_t *qf, const int8_t *ytox, const int8_t *ytob,
                               const float *tables, const jxg_frame_params *p, int32_t *levels[3],
                               float *coeffs[3], int16_t *dc[3], int by0, int by1);

/* ---- quant-table blob layout (include/jxg.h) ---- */
static const uint32_t QT_SIZES[13] = {64, 128, 256, 1024, 512, 256, 4096, 2048, 64, 64, 64, 64, 64};
uint32_t jxo_qt_size(int kind) { return QT_SIZES[kind]; }
uint32_t jxo_qt_offset(int kind) {
  uint32_t off = 0;
  for (int k = 0; k < kind; k++) off += 6 * QT_SIZES[k];
  return off;
}
int jxo_qt_kind(int s) {
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
  }
  return -1;
}

/* ---- natural coefficient orders (src/coeff_order.rs:206-248,344-415) ---- */
void jxo_natural_scan(int cx, int cy, uint32_t *out) {
  int size = cx * cy * 64, side = cx * 8;
  int xs = cx / cy, xsm = xs - 1, xss = 0;
  while ((1 << xss) < xs) xss++;
  int cur = cx * cy;
  for (int i = 0; i < side; i++) {
    for (int j = 0; j <= i; j++) {
      int x = j, y = i - j;
      if (i % 2 != 0) {
        int t = x;
        x = y;
        y = t;
      }
      if (y & xsm) continue;
      y >>= xss;
      int val;
      if (x < cx && y < cy)
        val = y * cx + x;
      else
        val = cur++;
      out[val] = (uint32_t)(y * side + x);
    }
  }
  for (int ip = side - 1; ip >= 1; ip--) {
    int i = ip - 1;
    for (int j = 0; j <= i; j++) {
      int x = side - 1 - (i - j), y = side - 1 - j;
      if (i % 2 != 0) {
        int t = x;
        x = y;
        y = t;
      }
      if (y & xsm) continue;
      y >>= xss;
      out[cur++] = (uint32_t)(y * side + x);
    }
  }
  (void)size;
}

void jxo_scan_pos_lut(int w, int h, uint32_t *lut) {
  int big = w > h ? w : h, small = w > h ? h : w;
  int size = w * h;
  uint32_t *order = (uint32_t *)malloc(size * sizeof(uint32_t));
  uint32_t *canon = (uint32_t *)malloc(size * sizeof(uint32_t));
  jxo_natural_scan(big / 8, small / 8, order);
  for (int k = 0; k < size; k++) canon[order[k]] = (uint32_t)k;
  if (w >= h) {
    memcpy(lut, canon, size * sizeof(uint32_t));
  } else {
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) lut[y * w + x] = canon[x * h + y];
  }
  free(order);
  free(canon);
}

/* ---- rate model (src/inflated_cost.rs:43-76,435-560) ---- */
static float g_rate_lut[1024];
static int g_rate_ready = 0;
static void rate_lut_init(void) {
  if (g_rate_ready) return;
  for (int i = 0; i < 1024; i++) g_rate_lut[i] = log2f(1.0f + (float)i);
  g_rate_ready = 1;
}
static float rate_log2(float qabs) {
  uint32_t k = qabs >= 4294967295.0f ? 0xffffffffu : (uint32_t)qabs;
  if (k < 1024) return g_rate_lut[k];
  return log2f(1.0f + qabs);
}

/* DequantBias::level_f32 (src/group.rs:748-768) */
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

typedef struct {
  float sse, mag_bits;
  uint32_t nzeros, max_scan;
} sse_rate;

static sse_rate sse_and_rate(const float *coeff, const float *inv, float q_scaled, float one,
                             float numerator, int width, int height, int cx, int cy,
                             const float thr[4], const uint32_t *scan_pos) {
  sse_rate r = {0.0f, 0.0f, 0, 0};
  int half = width / 2;
  for (int y = 0; y < height; y++) {
    int yfix = y >= height / 2 ? 2 : 0;
    for (int x = 0; x < width; x++) {
      if (x < cx && y < cy) continue;
      int i = y * width + x;
      float threshold = x >= half ? thr[yfix + 1] : thr[yfix];
      float a = inv[i] * q_scaled * coeff[i];
      float q = fabsf(a) >= threshold ? roundf(a) : 0.0f;
      float d = a - level_f32(one, numerator, q);
      r.sse += d * d;
      if (q != 0.0f) {
        r.nzeros++;
        r.mag_bits += rate_log2(fabsf(q));
        if (scan_pos[i] > r.max_scan) r.max_scan = scan_pos[i];
      }
    }
  }
  return r;
}

static float model_bits(uint32_t nzeros, float mag_bits, uint32_t max_scan, int cx, int cy) {
  float header = 0.4f * rate_log2((float)nzeros);
  float visited = 0.0f;
  if (nzeros != 0) {
    int64_t span = (int64_t)max_scan + 1 - (int64_t)(cx * cy + nzeros);
    visited = span > 0 ? (float)span : 0.0f;
  }
  return (float)nzeros * 1.6f + 1.0f * mag_bits + header + 0.5f * visited;
}

/* erased_detail / selection_quantizer::scale
 * (src/ac_strategy/selection_quantizer.rs:46-120).  The gate `enabled`
 * holds at the Fast tiers for 1 <= d <= 4 with the SSIMULACRA2 arbiter. */
static int selection_quantizer_enabled(const jxg_frame_params *p) {
  if (p->arbiter != JXG_ARBITER_SSIMULACRA2 || !(p->distance >= 1.0f && p->distance <= 4.0f)) return 0;
  return p->speed == JXG_SPEED_FAST || p->speed == JXG_SPEED_FASTEST;
}

static float selection_scale(int s, const float *luma, const float *inv_y, float qac,
                             const jxg_frame_params *p, float meta_r) {
  if (meta_r == 0.0f || !(s == JXG_DCT32X32 || s == JXG_DCT32X16 || s == JXG_DCT16X32) ||
      !selection_quantizer_enabled(p))
    return 1.0f;
  int cxs = jxo_covered_x(s), cys = jxo_covered_y(s);
  int cx = cxs > cys ? cxs : cys, cy = cxs > cys ? cys : cxs;
  int width = cx * 8;
  double energy = 0.0;
  for (int y = 0; y < cy * 8; y++)
    for (int x = 0; x < width; x++) {
      if (x < cx * 2 && y < cy * 2) continue;
      int i = y * width + x;
      if (fabsf(luma[i] * inv_y[i] * qac) < 0.5f) energy += (double)luma[i] * (double)luma[i];
    }
  float erased = (float)energy;
  float d = p->distance;
  float t = erased / (2e-4f * d * d) - 1.0f;
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  return 1.0f + 0.6f * t;
}

#define RD_LAMBDA 0.08086717f
#define META_R 4.0f

static uint32_t *g_scan_luts[64];
static pthread_mutex_t g_scan_mu = PTHREAD_MUTEX_INITIALIZER;
static const uint32_t *scan_lut(int w, int h) {
  int key = (w / 8 - 1) * 8 + (h / 8 - 1);
  pthread_mutex_lock(&g_scan_mu);
  if (!g_scan_luts[key]) {
    uint32_t *l = (uint32_t *)malloc((size_t)w * h * sizeof(uint32_t));
    jxo_scan_pos_lut(w, h, l);
    g_scan_luts[key] = l;
  }
  pthread_mutex_unlock(&g_scan_mu);
  return g_scan_luts[key];
}

float jxo_strategy_cost(const float *xyb, int width, int height, int s, int px, int py,
                        float qac, float meta_r, const float cmap[3], const float *tables,
                        const jxg_frame_params *p) {
  rate_lut_init();
  size_t plane = (size_t)width * height;
  static __thread float co[3][64 * 64];
  int size = 0;
  for (int c = 0; c < 3; c++)
    size = jxo_forward_transform(s, xyb + c * plane, width, height, px, py, co[c]);
  float ncx = -cmap[0], ncb = -cmap[2];
  for (int i = 0; i < size; i++) {
    co[0][i] = jxo_fmla(ncx, co[1][i], co[0][i]);
    co[2][i] = jxo_fmla(ncb, co[1][i], co[2][i]);
  }
  int cxs = jxo_covered_x(s), cys = jxo_covered_y(s);
  int cx = cxs > cys ? cxs : cys, cy = cxs > cys ? cys : cxs;
  int kind = jxo_qt_kind(s);
  uint32_t n = jxo_qt_size(kind), off = jxo_qt_offset(kind);
  const float *inv[3] = {tables + off + 3 * n, tables + off + 4 * n, tables + off + 5 * n};
  qac = qac * selection_scale(s, co[1], inv[1], qac, p, meta_r);
  /* Footprint::visible_fraction (src/inflated_cost.rs:650-660) */
  int fw = cxs * 8, fh = cys * 8;
  int vw = width - px < fw ? width - px : fw;
  int vh = height - py < fh ? height - py : fh;
  float visible = (vw == fw && vh == fh) ? 1.0f : (float)(vw * vh) / (float)(fw * fh);
  float qm_mult[3] = {jxo_x_qm_mul(p->x_qm_scale), 1.0f, jxo_b_qm_mul(p->b_qm_scale)};
  const uint32_t *scan = scan_lut(cx * 8, cy * 8);
  float d_total = 0.0f, r_total = 0.0f;
  for (int c = 0; c < 3; c++) {
    float thr[4];
    jxo_ac_thresholds(c, cx, cy, p->distance, qm_mult[c], thr);
    sse_rate r = sse_and_rate(co[c], inv[c], qac * qm_mult[c], p->dequant_one[c],
                              p->dequant_numerator, cx * 8, cy * 8, cx, cy, thr, scan);
    d_total += p->channel_weight[c] * r.sse;
    r_total += model_bits(r.nzeros, r.mag_bits, r.max_scan, cx, cy);
  }
  /* strategy_cost64 (src/ac_strategy.rs:646-720) adds without fmla */
  if (s == JXG_DCT64X64 || s == JXG_DCT64X32 || s == JXG_DCT32X64)
    return d_total * visible + RD_LAMBDA * (r_total + meta_r);
  return jxo_fmla(RD_LAMBDA, r_total + meta_r, d_total * visible);
}

/* ---- merge tuning (src/ac_strategy.rs:87-350) ---- */
typedef struct {
  float bias_16x16, bias_32x32, accept_16, accept_32, risk_k, site_k;
} merge_tuning;

static float clampf_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

static float banded_at(float hq, float base, float d) {
  float t = clampf_((d - 0.5f) / (1.5f - 0.5f), 0.0f, 1.0f);
  return jxo_fmla(t, base - hq, hq);
}

static float merge_margin(float hq, float base, float d) {
  float fade = clampf_((d - 1.0f) / (4.0f - 1.0f), 0.0f, 1.0f);
  return banded_at(hq, base, d) * (1.0f - fade * (1.0f - 0.82f));
}

static merge_tuning merge_tuning_new(const jxg_frame_params *p) {
  float d = p->distance;
  merge_tuning t;
  t.site_k = (p->speed == JXG_SPEED_FAST || p->speed == JXG_SPEED_FASTEST) ? 0.2f : 0.0f;
  t.bias_16x16 = banded_at(1.2059275f, 1.05f, d);
  t.bias_32x32 = banded_at(1.2059275f, 1.06f, d);
  t.accept_16 = 1.0f - merge_margin(0.19243888f, 0.048f, d);
  t.accept_32 = 1.0f - merge_margin(0.21398508f, 0.234f, d);
  t.risk_k = jxo_aq_dampen(d) > 0.0f ? 1.44269504f : 0.0f;
  float vlq_t = clampf_((d - 4.0f) / (7.5f - 4.0f), 0.0f, 1.0f);
  if (vlq_t > 0.0f) {
    t.bias_16x16 = jxo_fmla(vlq_t, 0.880f - t.bias_16x16, t.bias_16x16);
    t.bias_32x32 = jxo_fmla(vlq_t, 0.867f - t.bias_32x32, t.bias_32x32);
    t.accept_16 = jxo_fmla(vlq_t, (1.0f - 0.067f) - t.accept_16, t.accept_16);
    t.accept_32 = jxo_fmla(vlq_t, (1.0f - 0.02925f) - t.accept_32, t.accept_32);
  }
  if (d >= 1.5f) t.bias_32x32 *= 0.94f;
  return t;
}

/* site_variance_rect / MergeTuning::site_bias (src/ac_strategy.rs:224-245,269-282) */
static float site_bias(const merge_tuning *t, const float *xyb, int width, int height, int x0,
                       int y0, int n) {
  if (t->site_k == 0.0f) return 1.0f;
  const float *py_ = xyb + (size_t)width * height;
  double s1 = 0.0, s2 = 0.0;
  size_t cnt = 0;
  int ye = y0 + n < height ? y0 + n : height;
  int xe = x0 + n < width ? x0 + n : width;
  for (int y = y0; y < ye; y++)
    for (int x = x0; x < xe; x++) {
      double v = (double)py_[(size_t)y * width + x];
      s1 += v;
      s2 += v * v;
      cnt++;
    }
  float var = 0.0f;
  if (cnt) {
    double mean = s1 / (double)cnt;
    double v = s2 / (double)cnt - mean * mean;
    var = (float)(v > 0.0 ? v : 0.0);
  }
  if (var < 1e-9f) var = 1e-9f;
  float s = clampf_(jxo_dirty_log2f(var / 0.003f) / 8.0f, -1.0f, 1.0f);
  return jxo_fast_exp(t->site_k * s);
}

static float risk_gated(float k, float accept, float q_min, float q_max, float area_scale) {
  if (q_min <= 0.0f || k <= 0.0f) return accept;
  float spread = q_max / q_min - 1.0f;
  return accept * jxo_fast_exp2(-k * area_scale * spread);
}

static float mean_max_mixer(float d) {
  float v = 1.0f - fmaxf(d - 1.54138f, 0.0f) * 0.56391f;
  return fmaxf(v, 0.0f);
}

static uint8_t aggregate_quant(uint8_t max_q, uint32_t sum, int covered, float d) {
  if (covered < 4) return max_q;
  float mixer = mean_max_mixer(d);
  float mean = (float)sum / (float)covered;
  float mixed = (float)max_q * mixer + mean * (1.0f - mixer);
  float v = clampf_(mixed + 0.5f, 1.0f, 255.0f);
  return (uint8_t)v;
}

static float aggregate_qac_2x2(const float q[2][2], float scale, float d) {
  float max_q = fmaxf(fmaxf(fmaxf(q[0][0], q[0][1]), q[1][0]), q[1][1]);
  float mixer = mean_max_mixer(d);
  float mean = (q[0][0] + q[0][1] + q[1][0] + q[1][1]) * 0.25f;
  float mixed_q = (max_q * mixer + mean * (1.0f - mixer)) / scale;
  return scale * floorf(clampf_(mixed_q + 0.5f, 1.0f, 255.0f));
}

/* can_place_strategy (src/dc_group_data.rs:171-202), group-relative. */
static int can_place(int x, int y, int s, int xsize, int ysize) {
  int cx = jxo_covered_x(s), cy = jxo_covered_y(s);
  if (x + cx > xsize || y + cy > ysize) return 0;
  if (x / 32 != (x + cx - 1) / 32) return 0;
  if (y / 32 != (y + cy - 1) / 32) return 0;
  if (x / 8 != (x + cx - 1) / 8) return 0;
  if (y / 8 != (y + cy - 1) / 8) return 0;
  return 1;
}

typedef struct {
  const float *xyb;
  int width, height;
  int gx0, gy0; /* DC group origin in blocks */
  int xsize, ysize;
  const uint8_t *qf; /* whole-image field */
  int qf_stride;
  const int8_t *ytox, *ytob;
  int tiles_x, tiles_y;
  const float *tables;
  const jxg_frame_params *p;
  merge_tuning mt;
  uint8_t *cells; /* whole image */
  float *cost8, *cost16, *cost32;
} group_ctx;

static void cmap_factors(const group_ctx *g, int bx, int by, float out[3]) {
  /* group-relative block -> this DC group's tile, clamped to its tile grid */
  int gtx = g->gx0 / 8, gty = g->gy0 / 8;
  int ntx = (g->xsize + 7) / 8, nty = (g->ysize + 7) / 8;
  int tx = bx / 8 < ntx - 1 ? bx / 8 : ntx - 1;
  int ty = by / 8 < nty - 1 ? by / 8 : nty - 1;
  size_t i = (size_t)(gty + ty) * g->tiles_x + gtx + tx;
  float inv_cf = 1.0f / g->p->cfl_color_factor;
  out[0] = jxo_fmla((float)g->ytox[i], inv_cf, g->p->cfl_base_x);
  out[1] = 0.0f;
  out[2] = jxo_fmla((float)g->ytob[i], inv_cf, g->p->cfl_base_b);
}

static uint8_t qf_at(const group_ctx *g, int bx, int by) {
  return g->qf[(size_t)(g->gy0 + by) * g->qf_stride + g->gx0 + bx];
}

static void set_first(group_ctx *g, int x, int y, int s) {
  int cx = jxo_covered_x(s), cy = jxo_covered_y(s);
  for (int iy = 0; iy < cy; iy++)
    for (int ix = 0; ix < cx; ix++)
      g->cells[(size_t)(g->gy0 + y + iy) * g->qf_stride + g->gx0 + x + ix] =
          (uint8_t)((s << 1) | (ix == 0 && iy == 0 ? 1 : 0));
}

static float cost_at(group_ctx *g, int s, int bx, int by, float qac, const float cmap[3]) {
  return jxo_strategy_cost(g->xyb, g->width, g->height, s, (g->gx0 + bx) * 8, (g->gy0 + by) * 8,
                           qac, META_R, cmap, g->tables, g->p);
}

/* select_super_block with SearchScope::for_speed at tiers without
 * rectangles (src/ac_strategy/selection.rs:238-455): rectangles price as
 * +inf, so only DCT8 and DCT16X16 compete. */
static void select_super_block(group_ctx *g, int bx0, int by0, float *chosen, float *dct8) {
  float qac[2][2];
  for (int dy = 0; dy < 2; dy++)
    for (int dx = 0; dx < 2; dx++) qac[dy][dx] = g->p->scale * (float)qf_at(g, bx0 + dx, by0 + dy);
  float cmap[3];
  cmap_factors(g, bx0, by0, cmap);
  float c8[2][2];
  for (int dy = 0; dy < 2; dy++)
    for (int dx = 0; dx < 2; dx++) {
      c8[dy][dx] = cost_at(g, JXG_DCT, bx0 + dx, by0 + dy, qac[dy][dx], cmap);
      if (g->cost8)
        g->cost8[(size_t)(g->gy0 + by0 + dy) * g->qf_stride + g->gx0 + bx0 + dx] = c8[dy][dx];
    }
  float dct8_left = c8[0][0] + c8[1][0];
  float dct8_right = c8[0][1] + c8[1][1];
  float dct8_top = c8[0][0] + c8[0][1];
  float dct8_bottom = c8[1][0] + c8[1][1];
  float total_dct8 = dct8_left + dct8_right;
  float c16_raw = cost_at(g, JXG_DCT16X16, bx0, by0, aggregate_qac_2x2(qac, g->p->scale, g->p->distance),
                          cmap);
  if (g->cost16) g->cost16[(size_t)(g->gy0 + by0) * g->qf_stride + g->gx0 + bx0] = c16_raw;
  float c16 = g->mt.bias_16x16 *
              site_bias(&g->mt, g->xyb, g->width, g->height, (g->gx0 + bx0) * 8, (g->gy0 + by0) * 8,
                        16) *
              c16_raw;
  float cost_16x8 = dct8_left + dct8_right;
  float cost_8x16 = dct8_top + dct8_bottom;
  float best_rect = fminf(cost_16x8, cost_8x16);
  float q_min = INFINITY, q_max = 0.0f;
  for (int dy = 0; dy < 2; dy++)
    for (int dx = 0; dx < 2; dx++) {
      q_min = fminf(q_min, qac[dy][dx]);
      q_max = fmaxf(q_max, qac[dy][dx]);
    }
  int pick16 = can_place(bx0, by0, JXG_DCT16X16, g->xsize, g->ysize) && c16 < best_rect &&
               c16 < total_dct8 * risk_gated(g->mt.risk_k, g->mt.accept_16, q_min, q_max, 1.0f);
  if (pick16) {
    set_first(g, bx0, by0, JXG_DCT16X16);
    *chosen = c16;
  } else {
    *chosen = cost_16x8 <= cost_8x16 ? cost_16x8 : cost_8x16;
  }
  *dct8 = total_dct8;
}

static uint8_t region_quant(const group_ctx *g, int bx, int by, int w, int h) {
  uint8_t max_q = 1;
  uint32_t sum = 0;
  for (int y = by; y < by + h; y++)
    for (int x = bx; x < bx + w; x++) {
      uint8_t q = qf_at(g, x, y);
      if (q > max_q) max_q = q;
      sum += q;
    }
  return aggregate_quant(max_q, sum, w * h, g->p->distance);
}

/* The eight partitions of four children (src/ac_strategy/selection.rs:86-89):
 * none, the square, either or both halves in either direction.  Bits: square,
 * left, right, top, bottom. */
static const uint8_t MERGE_PARTITIONS[8] = {0, 1, 2, 4, 6, 8, 16, 24};

/* partition_cost (src/ac_strategy/selection.rs:91-108) */
float jxo_partition_cost(const float children[2][2], const float merges[5], int mask) {
  if (mask & 1) return merges[0];
  float cost = 0.0f;
  for (int k = 1; k < 5; k++)
    if (mask & (1 << k)) cost += merges[k];
  for (int y = 0; y < 2; y++)
    for (int x = 0; x < 2; x++)
      if ((mask & ((1 << (1 + x)) | (1 << (3 + y)))) == 0) cost += children[y][x];
  return cost;
}

/* choose_merge_partition (src/ac_strategy/selection.rs:115-150): the
 * cheapest partition among those whose merged cost is below both the
 * children it replaces and their accepted baseline. */
int jxo_choose_merge_partition(const float children[2][2], const float baseline[2][2],
                               const float merges[5], const float accept[5]) {
  float cc[5], bc[5];
  const float (*t[2])[2] = {children, baseline};
  float *out[2] = {cc, bc};
  for (int i = 0; i < 2; i++) {
    const float (*c)[2] = t[i];
    out[i][0] = ((c[0][0] + c[0][1]) + c[1][0]) + c[1][1];
    out[i][1] = c[0][0] + c[1][0];
    out[i][2] = c[0][1] + c[1][1];
    out[i][3] = c[0][0] + c[0][1];
    out[i][4] = c[1][0] + c[1][1];
  }
  int best = 0;
  float best_cost = cc[0];
  for (int m = 1; m < 8; m++) {
    const int mask = MERGE_PARTITIONS[m];
    float merged = 0.0f, incumbent = 0.0f, limit = 0.0f;
    for (int k = 0; k < 5; k++)
      if (mask & (1 << k)) {
        merged += merges[k];
        incumbent += cc[k];
        limit += bc[k] * accept[k];
      }
    if (!(merged < incumbent && merged < limit)) continue;
    const float cost = jxo_partition_cost(children, merges, mask);
    if (cost < best_cost) {
      best = mask;
      best_cost = cost;
    }
  }
  return best;
}

/* choose_merge_partition restricted to the square merge, the Fast tier's
 * scope: rectangles price +inf, so no mask but the square can pass the
 * `merged < incumbent` gate, and the square wins when it passes. */
static int choose_square32(const float children[2][2], const float baseline[2][2], float merge,
                           float accept) {
  float child_total = children[0][0] + children[0][1] + children[1][0] + children[1][1];
  float base_total = baseline[0][0] + baseline[0][1] + baseline[1][0] + baseline[1][1];
  float limit = base_total * accept;
  if (!(merge < child_total && merge < limit)) return 0;
  return merge < child_total;
}

/* test hooks for the decision pieces */
int jxo_choose_square32(const float children[2][2], const float baseline[2][2], float merge,
                        float accept) {
  return choose_square32(children, baseline, merge, accept);
}
int jxo_aggregate_quant(int max_q, uint32_t sum, int covered, float distance) {
  return aggregate_quant((uint8_t)max_q, sum, covered, distance);
}
float jxo_aggregate_qac_2x2(const float qac[2][2], float scale, float distance) {
  return aggregate_qac_2x2(qac, scale, distance);
}
/* cmap_factors (src/ac_strategy.rs:1841) of one tile's map entries */
void jxo_cmap_factors(int ytox, int ytob, const jxg_frame_params *p, float out[3]) {
  const float inv_cf = 1.0f / p->cfl_color_factor;
  out[0] = jxo_fmla((float)ytox, inv_cf, p->cfl_base_x);
  out[1] = 0.0f;
  out[2] = jxo_fmla((float)ytob, inv_cf, p->cfl_base_b);
}

/* select_band over one DC group (src/ac_strategy/selection.rs:1000-1180),
 * square scope, followed by adjust_quant_field (:503). */
static void select_group(group_ctx *g) {
  const jxg_frame_params *p = g->p;
  int xsize = g->xsize, ysize = g->ysize;
  for (int by = 0; by < ysize; by++)
    for (int bx = 0; bx < xsize; bx++) set_first(g, bx, by, JXG_DCT);
  int by = 0;
  while (by + 1 < ysize) {
    int four_row = (by % 4 == 0) && by + 4 <= ysize;
    int bx = 0;
    while (bx + 1 < xsize) {
      int four_col = bx % 4 == 0 && bx + 4 <= xsize;
      if (four_row && four_col && can_place(bx, by, JXG_DCT32X32, xsize, ysize)) {
        float children[2][2], baseline[2][2];
        for (int sy = 0; sy < 2; sy++)
          for (int sx = 0; sx < 2; sx++)
            select_super_block(g, bx + sx * 2, by + sy * 2, &children[sy][sx], &baseline[sy][sx]);
        float qac32 = p->scale * (float)region_quant(g, bx, by, 4, 4);
        float cmap[3];
        cmap_factors(g, bx, by, cmap);
        float cost32 = cost_at(g, JXG_DCT32X32, bx, by, qac32, cmap);
        if (g->cost32) g->cost32[(size_t)(g->gy0 + by) * g->qf_stride + g->gx0 + bx] = cost32;
        float m32 =
            site_bias(&g->mt, g->xyb, g->width, g->height, (g->gx0 + bx) * 8, (g->gy0 + by) * 8, 32);
        float merged = g->mt.bias_32x32 * m32 * cost32;
        uint8_t q_min = 255, q_max = 0;
        for (int iy = 0; iy < 4; iy++)
          for (int ix = 0; ix < 4; ix++) {
            uint8_t q = qf_at(g, bx + ix, by + iy);
            if (q < q_min) q_min = q;
            if (q > q_max) q_max = q;
          }
        float gate = risk_gated(g->mt.risk_k, g->mt.accept_32, (float)q_min, (float)q_max, 2.0f);
        if (choose_square32(children, baseline, merged, gate)) set_first(g, bx, by, JXG_DCT32X32);
        bx += 4;
      } else if (four_row) {
        float ch, d8;
        select_super_block(g, bx, by, &ch, &d8);
        select_super_block(g, bx, by + 2, &ch, &d8);
        bx += 2;
      } else {
        float ch, d8;
        select_super_block(g, bx, by, &ch, &d8);
        bx += 2;
      }
    }
    by += four_row ? 4 : 2;
  }
}

/* ------------------------------------------------------------------ */
/* Coefficient-model re-tiling of the merge search's layout: the Fast and */
/* Fastest tiers' layout_dp::retile_coefficient                           */
/* (src/ac_strategy/selection/layout_dp.rs:1737-2430, called from         */
/* src/ac_strategy/selection.rs:2194-2208).  Every 8x8-block tile of the  */
/* DC group is re-planned exactly over DCT8 leaves, 8x16 / 16x8 pairs and */
/* 16x16 squares (plan_small, a broken-profile DP), around or over the    */
/* incumbent 32-family blocks, and the new pieces are installed when they */
/* price below the pieces they replace.  Fast adds unaligned pairs and    */
/* squares and a second pass on the grid shifted by (4, 4) blocks         */
/* (layout_dp_unaligned, layout_crossing: src/effort.rs:322-349).         */

enum { SM_PRE = 0, SM_LEAF = 1, SM_WIDE = 2, SM_TALL = 3, SM_SQUARE = 4 };
#define SM_STATES 512 /* 1 << (8 + 1) */

static int small_strategy(int choice) {
  switch (choice) {
    case SM_LEAF: return JXG_DCT;
    case SM_WIDE: return JXG_DCT8X16; /* two blocks wide */
    case SM_TALL: return JXG_DCT16X8; /* two blocks tall */
    default: return JXG_DCT16X16;
  }
}

/* SmallCosts::live_masks (layout_dp.rs:1860-1916): bit 0 wide, bit 1 tall,
 * bit 2 square, each kept only when it prices below the smaller pieces
 * tiling the same cells. */
void jxo_live_masks(const float *leaf, const float *wide, const float *tall, const float *square,
                    int w, int h, uint8_t *live) {
  memset(live, 0, 64);
  if (w * h == 0) return;
  for (int y = 0; y < h; y++) {
    const int s = y * w;
    if (y + 1 == h) {
      for (int x = 0; x + 1 < w; x++) live[s + x] = wide[s + x] < leaf[s + x] + leaf[s + x + 1];
      continue;
    }
    const int b = s + w;
    for (int x = 0; x + 1 < w; x++) {
      const float top_leaves = leaf[s + x] + leaf[s + x + 1];
      const float leaves = (top_leaves + leaf[b + x]) + leaf[b + x + 1];
      const float wides = wide[s + x] + wide[b + x];
      const float talls = tall[s + x] + tall[s + x + 1];
      const float best = fminf(fminf(leaves, wides), talls);
      live[s + x] = (uint8_t)((wide[s + x] < top_leaves) |
                              ((tall[s + x] < leaf[s + x] + leaf[b + x]) << 1) |
                              ((square[s + x] < best) << 2));
    }
    const int r = w - 1;
    live[s + r] = (uint8_t)((tall[s + r] < leaf[s + r] + leaf[b + r]) << 1);
  }
}

typedef struct {
  float cost[SM_STATES], next[SM_STATES];
  uint16_t active[SM_STATES], next_active[SM_STATES];
  int n_active;
  uint8_t choices[64 * SM_STATES];
  int plan_n;
  uint8_t plan_cell[64], plan_choice[64];
} small_planner;

static void small_planner_init(small_planner *s) {
  for (int i = 0; i < SM_STATES; i++) s->cost[i] = s->next[i] = INFINITY;
  s->n_active = 0;
  s->plan_n = 0;
}

/* plan_small (layout_dp.rs:1918-2019): the cells are visited in row-major
 * order; a state holds the filled flags of the current cell and the w cells
 * after it.  Predecessors are relaxed in the order the states entered the
 * layer, each with the leaf, wide, tall, square choices in that order, and a
 * state keeps the first strictly cheaper arrival, so ties resolve as in the
 * reference.  Returns the optimum; when it is finite and below
 * replace_below, the pieces of the plan are saved (last cell first). */
static float plan_small(const float *leaf, const float *wide, const float *tall,
                        const float *square, const uint8_t *live, uint64_t fixed, int w, int h,
                        small_planner *s, float replace_below) {
  const int cells = w * h;
  const int states = 1 << (w + 1);
  const int mask = states - 1;
  const int below_bit = 1 << w, corner_bit = 2 << w;
  const int added[5] = {0, 0, 2, below_bit, 2 | below_bit | corner_bit};
  const int start = (int)(fixed & (uint64_t)mask);
  s->cost[start] = 0.0f;
  s->active[0] = (uint16_t)start;
  s->n_active = 1;
  uint64_t incoming = fixed >> (w + 1);
  for (int i = 0; i < cells; i++) {
    const int corner = (int)(incoming & 1u) << w;
    incoming >>= 1;
    const uint8_t shapes = live[i];
    uint8_t *from = s->choices + (size_t)i * states;
    int next_len = 0;
    for (int a = 0; a < s->n_active; a++) {
      const int state = s->active[a];
      const float here = s->cost[state];
      s->cost[state] = INFINITY;
#define RELAX(filled, c, ch)                                  \
  do {                                                        \
    const int to_ = ((((filled) >> 1) | corner) & (SM_STATES - 1)); \
    const float cc_ = (c);                                    \
    if (cc_ < s->next[to_]) {                                 \
      if (s->next[to_] == INFINITY) s->next_active[next_len++] = (uint16_t)to_; \
      s->next[to_] = cc_;                                     \
      from[to_] = (uint8_t)(ch);                              \
    }                                                         \
  } while (0)
      if (state & 1) {
        RELAX(state, here, SM_PRE);
        continue;
      }
      RELAX(state, here + leaf[i], SM_LEAF);
      const int right = !(state & 2), below = !(state & below_bit);
      if (right && (shapes & 1)) RELAX(state | 2, here + wide[i], SM_WIDE);
      if (below && (shapes & 2)) RELAX(state | below_bit, here + tall[i], SM_TALL);
      if (right && below && corner == 0 && (shapes & 4))
        RELAX(state | added[SM_SQUARE], here + square[i], SM_SQUARE);
#undef RELAX
    }
    for (int k = 0; k < SM_STATES; k++) {
      float t = s->cost[k];
      s->cost[k] = s->next[k];
      s->next[k] = t;
    }
    memcpy(s->active, s->next_active, sizeof(uint16_t) * (size_t)next_len);
    s->n_active = next_len;
  }
  const float total = s->cost[0];
  for (int a = 0; a < s->n_active; a++) s->cost[s->active[a]] = INFINITY;
  if (!(isfinite(total) && total < replace_below)) return total;
  s->plan_n = 0;
  int state = 0;
  for (int i = cells - 1; i >= 0; i--) {
    const int choice = s->choices[(size_t)i * states + state];
    if (choice != SM_PRE) {
      s->plan_cell[s->plan_n] = (uint8_t)i;
      s->plan_choice[s->plan_n] = (uint8_t)choice;
      s->plan_n++;
    }
    state = (((state << 1) ^ added[choice]) & mask) | (choice == SM_PRE);
  }
  return total;
}

/* Exported for the planner's own known-answer and exhaustive checks: the
 * optimum and the saved pieces (cell index, choice), last cell first. */
float jxo_plan_small(const float *leaf, const float *wide, const float *tall, const float *square,
                     uint64_t fixed, int w, int h, float replace_below, int *n_out,
                     uint8_t *cells_out, uint8_t *choices_out) {
  static __thread small_planner *sp;
  if (!sp) {
    sp = (small_planner *)malloc(sizeof(small_planner));
    small_planner_init(sp);
  }
  uint8_t live[64];
  jxo_live_masks(leaf, wide, tall, square, w, h, live);
  float t = plan_small(leaf, wide, tall, square, live, fixed, w, h, sp, replace_below);
  *n_out = sp->plan_n;
  memcpy(cells_out, sp->plan_cell, (size_t)sp->plan_n);
  memcpy(choices_out, sp->plan_choice, (size_t)sp->plan_n);
  return t;
}

/* can_place_strategy_crossing (src/dc_group_data.rs:207-218): inside the
 * group and one 256-px AC group; 64-px tiles may be crossed. */
static int can_place_crossing(int x, int y, int s, int xsize, int ysize) {
  int cx = jxo_covered_x(s), cy = jxo_covered_y(s);
  return x + cx <= xsize && y + cy <= ysize && x / 32 == (x + cx - 1) / 32 &&
         y / 32 == (y + cy - 1) / 32;
}

static int is_coefficient_strategy(int s) {
  return s == JXG_DCT || s == JXG_DCT16X8 || s == JXG_DCT8X16 || s == JXG_DCT16X16;
}

static uint64_t piece_mask(int x, int y, int s, int stride) {
  uint64_t row = (1ull << jxo_covered_x(s)) - 1ull, m = 0;
  for (int r = 0; r < jxo_covered_y(s); r++) m |= row << (y * stride + x + r * stride);
  return m;
}

/* footprint_is_self_contained (src/ac_strategy/selection.rs:3235-3258) */
static int footprint_closed(const group_ctx *g, int bx, int by, int w, int h) {
  int area = 0;
  for (int y = by; y < by + h; y++)
    for (int x = bx; x < bx + w; x++) {
      uint8_t cell = g->cells[(size_t)(g->gy0 + y) * g->qf_stride + g->gx0 + x];
      if (!(cell & 1)) continue;
      int s = cell >> 1, cx = jxo_covered_x(s), cy = jxo_covered_y(s);
      if (x + cx > bx + w || y + cy > by + h) return 0;
      area += cx * cy;
    }
  return area == w * h;
}

typedef struct {
  float *memo; /* [5][ysize][xsize]: DCT8, 16X8, 8X16, 16X16, 32X32; NaN = not yet */
} retile_memo;

static int memo_slot(int s) {
  switch (s) {
    case JXG_DCT: return 0;
    case JXG_DCT16X8: return 1;
    case JXG_DCT8X16: return 2;
    case JXG_DCT16X16: return 3;
    case JXG_DCT32X32: return 4;
  }
  return -1;
}

/* CoefficientScores::score (layout_dp.rs:1746-1810): strategy_cost at the
 * piece with region_qac of the (pre-adjustment) field and META_R.  The
 * reference's caches hold the same values. */
static float piece_score(group_ctx *g, retile_memo *m, int s, int x, int y) {
  int slot = memo_slot(s);
  size_t k = (size_t)y * g->xsize + x;
  float *cell = slot >= 0 ? &m->memo[(size_t)slot * g->xsize * g->ysize + k] : NULL;
  if (cell && !isnan(*cell)) return *cell;
  float cmap[3];
  cmap_factors(g, x, y, cmap);
  float qac = g->p->scale *
              (float)region_quant(g, x, y, jxo_covered_x(s), jxo_covered_y(s));
  float c = cost_at(g, s, x, y, qac, cmap);
  if (cell) *cell = c;
  return c;
}

typedef struct {
  int x, y, s;
} tpiece;

static int tpiece_cmp(const void *a, const void *b) {
  const tpiece *p = (const tpiece *)a, *q = (const tpiece *)b;
  if (p->y != q->y) return p->y - q->y;
  if (p->x != q->x) return p->x - q->x;
  return p->s - q->s;
}

static int tpiece_in(const tpiece *set, int n, tpiece p) {
  for (int i = 0; i < n; i++)
    if (set[i].x == p.x && set[i].y == p.y && set[i].s == p.s) return 1;
  return 0;
}

/* search_tile_coefficient (layout_dp.rs:2033-2199) on the w x h tile at
 * group block (bx, by); the accepted pieces are installed at once (tiles of
 * one pass are disjoint and read only their own cells). */
static void retile_tile(group_ctx *g, retile_memo *m, small_planner *sp, int bx, int by, int w,
                        int h, int crossing, int unaligned) {
  tpiece cur[64];
  int ncur = 0;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      uint8_t cell = g->cells[(size_t)(g->gy0 + by + y) * g->qf_stride + g->gx0 + bx + x];
      if (!(cell & 1)) continue;
      int s = cell >> 1;
      if (!crossing || (x + jxo_covered_x(s) <= w && y + jxo_covered_y(s) <= h)) {
        tpiece p = {x, y, s};
        cur[ncur++] = p;
      }
    }
  uint64_t blocked = 0;
  if (crossing) {
    uint64_t full = (w * h == 64) ? ~0ull : ((1ull << (w * h)) - 1ull);
    blocked = full;
    for (int i = 0; i < ncur; i++) blocked &= ~piece_mask(cur[i].x, cur[i].y, cur[i].s, w);
  }
  for (int i = 0; i < ncur; i++)
    if (cur[i].s == JXG_IDENTITY || cur[i].s == JXG_DCT2X2) return;
  float costs[4][64];
  for (int k = 0; k < 4; k++)
    for (int i = 0; i < 64; i++) costs[k][i] = INFINITY;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const int i = y * w + x;
      for (int choice = SM_LEAF; choice <= SM_SQUARE; choice++) {
        const int s = small_strategy(choice);
        if (x + jxo_covered_x(s) > w || y + jxo_covered_y(s) > h) continue;
        if (blocked && (piece_mask(x, y, s, w) & blocked)) continue;
        const int gx = bx + x, gy = by + y;
        const int off_grid = choice == SM_WIDE   ? (gx % 2 != 0)
                             : choice == SM_TALL ? (gy % 2 != 0)
                             : choice == SM_SQUARE ? (gx % 2 != 0 || gy % 2 != 0)
                                                   : 0;
        if (off_grid && !unaligned) continue;
        const int legal = crossing ? can_place_crossing(gx, gy, s, g->xsize, g->ysize)
                                   : can_place(gx, gy, s, g->xsize, g->ysize);
        if (legal) costs[choice - 1][i] = piece_score(g, m, s, gx, gy);
      }
    }
  uint64_t fixed = 0;
  float fixed_cost = 0.0f;
  for (int i = 0; i < ncur; i++) {
    if (is_coefficient_strategy(cur[i].s)) continue;
    fixed |= piece_mask(cur[i].x, cur[i].y, cur[i].s, w);
    fixed_cost += piece_score(g, m, cur[i].s, bx + cur[i].x, by + cur[i].y);
  }
  uint8_t live[64];
  jxo_live_masks(costs[0], costs[1], costs[2], costs[3], w, h, live);
  const float around = plan_small(costs[0], costs[1], costs[2], costs[3], live, fixed | blocked, w,
                                  h, sp, INFINITY) +
                       fixed_cost;
  int keep_fixed = 1;
  float winner = around;
  if (fixed) {
    const float over =
        plan_small(costs[0], costs[1], costs[2], costs[3], live, blocked, w, h, sp, around);
    if (over < around) {
      winner = over;
      keep_fixed = 0;
    }
  }
  if (!isfinite(winner)) return;
  tpiece lay[64];
  int nlay = 0;
  for (int k = 0; k < sp->plan_n; k++) {
    tpiece p = {sp->plan_cell[k] % w, sp->plan_cell[k] / w, small_strategy(sp->plan_choice[k])};
    lay[nlay++] = p;
  }
  if (keep_fixed)
    for (int i = 0; i < ncur; i++)
      if (!is_coefficient_strategy(cur[i].s)) lay[nlay++] = cur[i];
  qsort(lay, (size_t)nlay, sizeof(tpiece), tpiece_cmp);
  float old_cost = 0.0f, new_cost = 0.0f;
  for (int i = 0; i < ncur; i++)
    if (!tpiece_in(lay, nlay, cur[i]))
      old_cost += piece_score(g, m, cur[i].s, bx + cur[i].x, by + cur[i].y);
  for (int i = 0; i < nlay; i++)
    if (!tpiece_in(cur, ncur, lay[i]))
      new_cost += piece_score(g, m, lay[i].s, bx + lay[i].x, by + lay[i].y);
  if (!(new_cost < old_cost)) return;
  for (int i = 0; i < nlay; i++)
    if (!tpiece_in(cur, ncur, lay[i])) set_first(g, bx + lay[i].x, by + lay[i].y, lay[i].s);
}

/* layout_dp_coefficient / _unaligned / layout_crossing per speed
 * (src/effort.rs:311-349). */
static void retile_flags(const jxg_frame_params *p, int *on, int *unaligned, int *crossing) {
  *on = p->speed == JXG_SPEED_FASTEST || p->speed == JXG_SPEED_FAST;
  *unaligned = p->speed == JXG_SPEED_FAST;
  *crossing = p->speed == JXG_SPEED_FAST;
}

/* One re-tiling pass over a DC group: the aligned grid (crossing = 0, only
 * tiles whose transforms stay inside them) or the grid shifted by (4, 4)
 * blocks (find_coefficient, layout_dp.rs:2340-2379). */
static void retile_pass(group_ctx *g, retile_memo *m, small_planner *sp, int crossing,
                        int unaligned) {
  const int o = crossing ? 4 : 0;
  for (int by = o; by < g->ysize; by += 8)
    for (int bx = o; bx < g->xsize; bx += 8) {
      const int w = g->xsize - bx < 8 ? g->xsize - bx : 8;
      const int h = g->ysize - by < 8 ? g->ysize - by : 8;
      if (w * h < 2 || (!crossing && !footprint_closed(g, bx, by, w, h))) continue;
      retile_tile(g, m, sp, bx, by, w, h, crossing, unaligned);
    }
}

static void retile_coefficient(group_ctx *g, int passes_mask) {
  int on, unaligned, crossing;
  retile_flags(g->p, &on, &unaligned, &crossing);
  if (!on) return;
  retile_memo m;
  size_t n = (size_t)g->xsize * g->ysize * 5;
  m.memo = (float *)malloc(n * sizeof(float));
  for (size_t i = 0; i < n; i++) m.memo[i] = NAN;
  small_planner *sp = (small_planner *)malloc(sizeof(small_planner));
  small_planner_init(sp);
  if (passes_mask & 1) retile_pass(g, &m, sp, 0, unaligned);
  if ((passes_mask & 2) && crossing) retile_pass(g, &m, sp, 1, unaligned);
  free(sp);
  free(m.memo);
}

static void adjust_quant_field(group_ctx *g, uint8_t *qf_out) {
  for (int by = 0; by < g->ysize; by++)
    for (int bx = 0; bx < g->xsize; bx++) {
      size_t i = (size_t)(g->gy0 + by) * g->qf_stride + g->gx0 + bx;
      uint8_t cell = g->cells[i];
      if (!(cell & 1)) continue;
      int s = cell >> 1;
      int cx = jxo_covered_x(s), cy = jxo_covered_y(s);
      if (cx == 1 && cy == 1) continue;
      uint8_t max_q = 0;
      uint32_t sum = 0;
      for (int iy = 0; iy < cy; iy++)
        for (int ix = 0; ix < cx; ix++) {
          uint8_t q = qf_at(g, bx + ix, by + iy);
          if (q > max_q) max_q = q;
          sum += q;
        }
      uint8_t val = aggregate_quant(max_q, sum, cx * cy, g->p->distance);
      for (int iy = 0; iy < cy; iy++)
        for (int ix = 0; ix < cx; ix++)
          qf_out[(size_t)(g->gy0 + by + iy) * g->qf_stride + g->gx0 + bx + ix] = val;
    }
}

static void group_init(group_ctx *g, const float *xyb, int width, int height, int gx, int gy,
                       const uint8_t *qf, const int8_t *ytox, const int8_t *ytob,
                       const float *tables, const jxg_frame_params *p, uint8_t *cells) {
  memset(g, 0, sizeof(*g));
  g->xyb = xyb;
  g->width = width;
  g->height = height;
  int x0 = gx * 2048, y0 = gy * 2048;
  int xs = width - x0 < 2048 ? width - x0 : 2048;
  int ys = height - y0 < 2048 ? height - y0 : 2048;
  g->gx0 = x0 / 8;
  g->gy0 = y0 / 8;
  g->xsize = (xs + 7) / 8;
  g->ysize = (ys + 7) / 8;
  g->qf = qf;
  g->qf_stride = (width + 7) / 8;
  g->ytox = ytox;
  g->ytob = ytob;
  g->tiles_x = (width + 63) / 64;
  g->tiles_y = (height + 63) / 64;
  g->tables = tables;
  g->p = p;
  g->mt = merge_tuning_new(p);
  g->cells = cells;
}

static void refine_quant_field(group_ctx *g, uint8_t *qf_out);
static void rerank_large_transforms(group_ctx *g);

/* fill_ac_strategy returns before scoping when the speed has no square
 * merges or the distance takes the DCT8-only shortcut
 * (src/ac_strategy/selection.rs:1930-1935, src/ac_strategy.rs:71-79). */
static int selection_runs(const jxg_frame_params *p) {
  if (p->distance <= 0.056713393f) return 0;
  return p->speed >= JXG_SPEED_FASTEST;
}

/* The Fast / Fastest decision of one DC group (fill_ac_strategy's stages,
 * src/ac_strategy/selection.rs:1917-2481): the merge search, the re-tiling,
 * the rerank (Fast, d <= 1.709), adjust_quant_field into qf_out, then the
 * quant-field refinement (Fast, d >= 1.5). */
static void decide_group(group_ctx *g, uint8_t *qf_out) {
  select_group(g);
  retile_coefficient(g, 3);
  rerank_large_transforms(g);
  adjust_quant_field(g, qf_out);
  refine_quant_field(g, qf_out);
}

void jxo_select_strategy(const float *xyb, int width, int height, uint8_t *qf,
                         const int8_t *ytox, const int8_t *ytob, const float *tables,
                         const jxg_frame_params *p, uint8_t *cells) {
  int xb = (width + 7) / 8, yb = (height + 7) / 8;
  if (!selection_runs(p)) {
    memset(cells, JXG_CELL(JXG_DCT, 1), (size_t)xb * yb);
    return;
  }
  uint8_t *qf_in = (uint8_t *)malloc((size_t)xb * yb);
  memcpy(qf_in, qf, (size_t)xb * yb);
  for (int gy = 0; gy * 2048 < height; gy++)
    for (int gx = 0; gx * 2048 < width; gx++) {
      group_ctx g;
      group_init(&g, xyb, width, height, gx, gy, qf_in, ytox, ytob, tables, p, cells);
      decide_group(&g, qf);
    }
  free(qf_in);
}

/* ------------------------------------------------------------------ */
/* refine_quant_field with qf_trellis (src/ac_strategy/selection.rs:534-896): */
/* after adjust_quant_field, the Fast tier tries one or two coarser steps of  */
/* each transform's quantizer, prices every candidate by its reconstruction   */
/* (quantize, inverse transform, SSIM deficit; jxo_recon.c), and picks the    */
/* field along the DC group's raster chain of first blocks with a Viterbi     */
/* search that also prices the QF residual stream (two rounds, the prices     */
/* re-estimated from the chain each time).                                    */

int jxo_recon_coeffs(const float *xyb, int width, int height, int s, int px, int py,
                     const float cmap[3], float *co);
float jxo_recon_cost(const float *xyb, int width, int height, int s, int px, int py, float qac,
                     const float cmap[3], const float *co, const float *tables,
                     const jxg_frame_params *p, float *dist_out, float *rate_out);
float jxo_reconstruction_lambda(float d);

/* quant_refinement_steps (:534-542) */
static int refinement_steps(float d) { return d < 1.5f ? 0 : (d >= 2.0f ? 2 : 1); }

/* STRATEGY_CODE_LUT (src/dc_group_data.rs:73-78) */
static const uint8_t STRATEGY_CODE[24] = {0, 6, 7, 4, 5, 3, 12, 13, 10, 11, 14, 15,
                                          16, 17, 18, 19, 20, 1, 2, 21, 22, 23, 8, 9};

/* pack_signed (src/entropy/token.rs:159) */
static uint32_t pack_signed(int32_t v) {
  int64_t x = v;
  return (uint32_t)((x << 1) ^ (x >> 63));
}

/* qf_symbol (:725-733) with ac_metadata_context (src/frame/metadata.rs:46) */
static void qf_symbol(int left, int qf, int *c, int *v) {
  *c = left > 11 ? 0 : (left > 5 ? 1 : (left > 3 ? 2 : 3));
  uint32_t pk = pack_signed(qf - left);
  *v = pk < 511 ? (int)pk : 511;
}

/* qf_residual_prices (:707-723) */
static void qf_prices(int first_left, const uint8_t *chain, int n, float prices[4][512]) {
  static __thread uint32_t counts[4][512];
  uint32_t totals[4] = {0, 0, 0, 0};
  memset(counts, 0, sizeof(counts));
  int left = first_left;
  for (int i = 0; i < n; i++) {
    int qf = (int)chain[i] - 1, c, v;
    qf_symbol(left, qf, &c, &v);
    counts[c][v]++;
    totals[c]++;
    left = qf;
  }
  for (int c = 0; c < 4; c++) {
    const float denom = (float)totals[c] + 2.0f;
    for (int v = 0; v < 512; v++) {
      const float prior = 2.0f * jxo_fast_exp2(-fminf((float)v + 1.0f, 100.0f));
      prices[c][v] = -jxo_dirty_log2f(((float)counts[c][v] + prior) / denom);
    }
  }
}

typedef struct {
  int len;
  uint8_t q[4];
  float cost[4];
} qrefine;

/* qf_trellis (:737-800) */
static void qf_trellis(int first_left, const uint8_t *fixed, const qrefine *const *states, int n,
                       float lambda_bits, const float prices[4][512], uint8_t *chain,
                       uint8_t (*back)[4]) {
  uint8_t prev_q[4] = {0, 0, 0, 0};
  float prev_cost[4] = {0, 0, 0, 0};
  int prev_len = 0;
  for (int i = 0; i < n; i++) {
    uint8_t cq[4];
    float cc[4];
    int len;
    if (states[i]) {
      len = states[i]->len;
      for (int s = 0; s < 4; s++) {
        cq[s] = states[i]->q[s];
        cc[s] = states[i]->cost[s];
      }
    } else {
      len = 1;
      cq[0] = fixed[i];
      cc[0] = 0.0f;
    }
    float cost[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
    back[i][0] = back[i][1] = back[i][2] = back[i][3] = 0;
    for (int s = 0; s < len; s++) {
      const int qf = (int)cq[s] - 1;
      int c, v;
      if (i == 0) {
        qf_symbol(first_left, qf, &c, &v);
        cost[s] = cc[s] + lambda_bits * prices[c][v];
        continue;
      }
      for (int pp = 0; pp < prev_len; pp++) {
        qf_symbol((int)prev_q[pp] - 1, qf, &c, &v);
        const float total = prev_cost[pp] + cc[s] + lambda_bits * prices[c][v];
        if (total < cost[s]) {
          cost[s] = total;
          back[i][s] = (uint8_t)pp;
        }
      }
    }
    float floor_ = INFINITY;
    for (int s = 0; s < len; s++) floor_ = fminf(floor_, cost[s]);
    for (int s = 0; s < len; s++) {
      prev_q[s] = cq[s];
      cost[s] -= floor_;
    }
    for (int s = 0; s < 4; s++) prev_cost[s] = cost[s];
    prev_len = len;
  }
  if (n == 0) return;
  int s = 0;
  for (int k = 1; k < prev_len; k++)
    if (prev_cost[k] < prev_cost[s]) s = k; /* first minimum (all >= +0) */
  for (int i = n - 1; i >= 0; i--) {
    chain[i] = states[i] ? states[i]->q[s] : fixed[i];
    s = back[i][s];
  }
}

/* Test entry points: the QF price table of a chain, and the trellis over
 * n blocks whose states (up to 4 (q, cost) pairs; len 0 = a fixed block) are
 * given in arrays. */
void jxo_qf_prices(int first_left, const uint8_t *chain, int n, float *prices /* 4 x 512 */) {
  qf_prices(first_left, chain, n, (float(*)[512])prices);
}
void jxo_qf_trellis(int first_left, const uint8_t *fixed, const int *lens, const uint8_t *qs,
                    const float *costs, int n, float lambda_bits, const float *prices,
                    uint8_t *chain) {
  qrefine *refs = (qrefine *)malloc(sizeof(qrefine) * (n ? n : 1));
  const qrefine **st = (const qrefine **)malloc(sizeof(qrefine *) * (n ? n : 1));
  uint8_t(*back)[4] = (uint8_t(*)[4])malloc(4 * (size_t)(n ? n : 1));
  for (int i = 0; i < n; i++) {
    st[i] = NULL;
    if (lens[i] > 0) {
      refs[i].len = lens[i];
      for (int k = 0; k < 4; k++) {
        refs[i].q[k] = qs[i * 4 + k];
        refs[i].cost[k] = costs[i * 4 + k];
      }
      st[i] = &refs[i];
    }
  }
  qf_trellis(first_left, fixed, st, n, lambda_bits, (const float(*)[512])prices, chain, back);
  free(refs);
  free(st);
  free(back);
}

/* refine_quant_field (:802-896) at the Fast tier (no learned prices, no
 * cached rerank costs, no X-heavy upward step) on the adjusted field. */
static void refine_quant_field(group_ctx *g, uint8_t *qf_out) {
  const jxg_frame_params *p = g->p;
  const int steps = refinement_steps(p->distance);
  if (!steps || p->speed != JXG_SPEED_FAST) return;
  const int xs = g->xsize, ys = g->ysize;
  const int cap = xs * ys;
  int *fbx = (int *)malloc(sizeof(int) * cap), *fby = (int *)malloc(sizeof(int) * cap);
  uint8_t *fs = (uint8_t *)malloc(cap), *fixed = (uint8_t *)malloc(cap),
          *chain = (uint8_t *)malloc(cap);
  qrefine *refs = (qrefine *)malloc(sizeof(qrefine) * cap);
  const qrefine **states = (const qrefine **)malloc(sizeof(qrefine *) * cap);
  uint8_t(*back)[4] = (uint8_t(*)[4])malloc(4 * (size_t)cap);
  static __thread float co[3 * 4096];
  int n = 0;
  for (int by = 0; by < ys; by++)
    for (int bx = 0; bx < xs; bx++) {
      const size_t i = (size_t)(g->gy0 + by) * g->qf_stride + g->gx0 + bx;
      const uint8_t cell = g->cells[i];
      if (!(cell & 1)) continue;
      const int s = cell >> 1;
      fbx[n] = bx;
      fby[n] = by;
      fs[n] = (uint8_t)s;
      fixed[n] = qf_out[i];
      states[n] = NULL;
      const uint8_t q = qf_out[i];
      if (!(s == JXG_DCT64X64 || s == JXG_DCT64X32 || s == JXG_DCT32X64) && q > 1) {
        float cmap[3];
        cmap_factors(g, bx, by, cmap);
        const int px = (g->gx0 + bx) * 8, py = (g->gy0 + by) * 8;
        jxo_recon_coeffs(g->xyb, g->width, g->height, s, px, py, cmap, co);
        qrefine *r = &refs[n];
        r->len = 1;
        r->q[0] = q;
        r->cost[0] = jxo_recon_cost(g->xyb, g->width, g->height, s, px, py,
                                    p->scale * (float)q, cmap, co, g->tables, p, NULL, NULL);
        for (int k = 1; k < 4; k++) {
          r->q[k] = q;
          r->cost[k] = r->cost[0];
        }
        const int cand[3] = {q - 1, steps == 2 ? q - 2 : q, q};
        for (int k = 0; k < 3; k++) {
          const int cq = cand[k] < 0 ? 0 : cand[k];
          int dup = cq == 0;
          for (int m = 0; m < r->len && !dup; m++) dup = r->q[m] == cq;
          if (dup) continue;
          r->q[r->len] = (uint8_t)cq;
          r->cost[r->len] = jxo_recon_cost(g->xyb, g->width, g->height, s, px, py,
                                           p->scale * (float)cq, cmap, co, g->tables, p, NULL,
                                           NULL);
          r->len++;
        }
        if (r->len > 1) states[n] = r;
      }
      n++;
    }
  /* independent_choice (:695-704) */
  for (int i = 0; i < n; i++) {
    if (!states[i]) {
      chain[i] = fixed[i];
      continue;
    }
    int b = 0;
    for (int k = 1; k < states[i]->len; k++)
      if (states[i]->cost[k] < states[i]->cost[b]) b = k;
    chain[i] = states[i]->q[b];
  }
  if (n > 0) {
    const int first_left = STRATEGY_CODE[g->cells[(size_t)g->gy0 * g->qf_stride + g->gx0] >> 1];
    const float lambda_bits = jxo_fmla(jxo_reconstruction_lambda(p->distance), 1.0f + 0.0f, 0.0f);
    static __thread float prices[4][512];
    for (int round = 0; round < 2; round++) {
      qf_prices(first_left, chain, n, prices);
      qf_trellis(first_left, fixed, states, n, lambda_bits, (const float(*)[512])prices, chain, back);
    }
  }
  for (int i = 0; i < n; i++) {
    if (chain[i] == fixed[i]) continue;
    const int cw = jxo_covered_x(fs[i]), ch = jxo_covered_y(fs[i]);
    for (int y = 0; y < ch; y++)
      for (int x = 0; x < cw; x++)
        qf_out[(size_t)(g->gy0 + fby[i] + y) * g->qf_stride + g->gx0 + fbx[i] + x] = chain[i];
  }
  free(fbx);
  free(fby);
  free(fs);
  free(fixed);
  free(chain);
  free(refs);
  free(states);
  free(back);
}

/* One re-tiling pass (aligned: crossing = 0; shifted: 1) over every DC group
 * of a given layout, the field left as it is (test entry point). */
void jxo_retile_pass(const float *xyb, int width, int height, const uint8_t *qf,
                     const int8_t *ytox, const int8_t *ytob, const float *tables,
                     const jxg_frame_params *p, uint8_t *cells, int crossing) {
  jxo_set_fused(p->fused);
  for (int gy = 0; gy * 2048 < height; gy++)
    for (int gx = 0; gx * 2048 < width; gx++) {
      group_ctx g;
      group_init(&g, xyb, width, height, gx, gy, qf, ytox, ytob, tables, p, cells);
      retile_coefficient(&g, crossing ? 2 : 1);
    }
}

void jxo_strategy_costs(const float *xyb, int width, int height, const uint8_t *qf,
                        const int8_t *ytox, const int8_t *ytob, const float *tables,
                        const jxg_frame_params *p, float *cost8, float *cost16, float *cost32) {
  int xb = (width + 7) / 8, yb = (height + 7) / 8;
  size_t n = (size_t)xb * yb;
  for (size_t i = 0; i < n; i++) cost8[i] = cost16[i] = cost32[i] = NAN;
  uint8_t *cells = (uint8_t *)malloc(n);
  for (int gy = 0; gy * 2048 < height; gy++)
    for (int gx = 0; gx * 2048 < width; gx++) {
      group_ctx g;
      group_init(&g, xyb, width, height, gx, gy, qf, ytox, ytob, tables, p, cells);
      g.cost8 = cost8;
      g.cost16 = cost16;
      g.cost32 = cost32;
      select_group(&g);
    }
  free(cells);
}

/* ------------------------------------------------------------------ */
/* Whole front end with a thread pool over rows of DC groups / bands.  The */
/* stage order is the reference's: XYB, AQ per DC group, selection per DC  */
/* group, then transform + quantize.                                      */

typedef struct {
  int kind; /* 0 xyb rows, 1 aq group, 2 select group, 3 quantize rows, 4 cmap tile row */
  const uint8_t *rgb;
  size_t row_bytes;
  int channels, width, height;
  const jxg_frame_params *p;
  const float *tables;
  int8_t *ytox, *ytob;
  float *xyb;
  uint8_t *qf, *qf_in, *cells;
  int32_t **levels;
  float **coeffs;
  int16_t **dc;
  int next; /* work counter */
  int total;
  pthread_mutex_t mu;
} fe_job;

static int take(fe_job *j) {
  pthread_mutex_lock(&j->mu);
  int i = j->next < j->total ? j->next++ : -1;
  pthread_mutex_unlock(&j->mu);
  return i;
}

static void *fe_worker(void *arg) {
  fe_job *j = (fe_job *)arg;
  int W = j->width, H = j->height;
  int xb = (W + 7) / 8;
  int gxs = (W + 2047) / 2048;
  size_t plane = (size_t)W * H;
  float lut[256];
  jxo_srgb_lut(lut);
  /* the frame's XybMatrix::fwd and XYB kernel flavour */
  float m[9];
  memcpy(m, j->p->xyb_fwd, sizeof(m));
  const int order = j->p->xyb_order;
  int i;
  while ((i = take(j)) >= 0) {
    if (j->kind == 0) {
      for (int y = i * 8; y < (i + 1) * 8 && y < H; y++) {
        const uint8_t *row = j->rgb + (size_t)y * j->row_bytes;
        for (int x = 0; x < W; x++) {
          const uint8_t *px = row + (size_t)x * j->channels;
          float o[3];
          if (j->channels <= 2) { /* gray: linearize_gray, snapped after the alpha fill */
            jxo_xyb_pixel(m, order, lut[px[0]], lut[px[0]], lut[px[0]], o);
          } else {
            jxo_xyb_pixel(m, order, lut[px[0]], lut[px[1]], lut[px[2]], o);
          }
          size_t k = (size_t)y * W + x;
          j->xyb[k] = o[0];
          j->xyb[plane + k] = o[1];
          j->xyb[2 * plane + k] = o[2];
        }
      }
    } else if (j->kind == 1) {
      int gx = i % gxs, gy = i / gxs;
      int x0 = gx * 2048, y0 = gy * 2048;
      int xs = W - x0 < 2048 ? W - x0 : 2048, ys = H - y0 < 2048 ? H - y0 : 2048;
      int gxb = (xs + 7) / 8, gyb = (ys + 7) / 8;
      uint8_t *tmp = (uint8_t *)malloc((size_t)gxb * gyb);
      jxo_fill_quant_field_group(j->xyb, W, H, x0, y0, gxb, gyb, j->p->distance,
                                 1.0f / j->p->scale, tmp);
      for (int by = 0; by < gyb; by++)
        memcpy(j->qf + (size_t)(y0 / 8 + by) * xb + x0 / 8, tmp + (size_t)by * gxb, gxb);
      free(tmp);
    } else if (j->kind == 2) {
      int gx = i % gxs, gy = i / gxs;
      group_ctx g;
      group_init(&g, j->xyb, W, H, gx, gy, j->qf_in, j->ytox, j->ytob, j->tables, j->p, j->cells);
      decide_group(&g, j->qf);
    } else if (j->kind == 4) {
      const int tiles_x = (W + 63) / 64;
      for (int tx = 0; tx < tiles_x; tx++) {
        int x, b;
        jxo_cmap_tile(j->xyb, W, H, tx, i, j->tables, j->p, &x, &b);
        j->ytox[i * tiles_x + tx] = (int8_t)x;
        j->ytob[i * tiles_x + tx] = (int8_t)b;
      }
    } else if (j->kind == 3) {
      int yb = (H + 7) / 8;
      int by1 = (i + 1) * 8 < yb ? (i + 1) * 8 : yb;
      jxo_forward_quantize_rows(j->xyb, W, H, j->cells, j->qf, j->ytox, j->ytob, j->tables, j->p,
                                j->levels, j->coeffs, j->dc, i * 8, by1);
    }
  }
  return NULL;
}

static void run_pool(fe_job *j, int kind, int total, int threads) {
  j->kind = kind;
  j->next = 0;
  j->total = total;
  pthread_t tids[256];
  if (threads > 256) threads = 256;
  for (int t = 0; t < threads; t++) pthread_create(&tids[t], NULL, fe_worker, j);
  for (int t = 0; t < threads; t++) pthread_join(tids[t], NULL);
}

void jxo_front_end(const uint8_t *rgb, size_t row_bytes, int channels, int width, int height,
                   const jxg_frame_params *p, const float *tables, int8_t *ytox, int8_t *ytob,
                   int threads, uint8_t *qf, uint8_t *cells,
                   int32_t *levels[3], float *coeffs[3], int16_t *dc[3], uint8_t *alpha_out) {
  jxo_set_fused(p->fused);
  fe_job j;
  memset(&j, 0, sizeof(j));
  pthread_mutex_init(&j.mu, NULL);
  j.rgb = rgb;
  j.row_bytes = row_bytes;
  j.channels = channels;
  j.width = width;
  j.height = height;
  j.p = p;
  j.tables = tables;
  j.ytox = ytox;
  j.ytob = ytob;
  size_t plane = (size_t)width * height;
  int xb = (width + 7) / 8, yb = (height + 7) / 8;
  j.xyb = (float *)malloc(3 * plane * sizeof(float));
  j.qf = qf;
  j.cells = cells;
  if (threads < 1) threads = 1;
  int groups = ((width + 2047) / 2048) * ((height + 2047) / 2048);
  run_pool(&j, 0, (height + 7) / 8, threads);
  /* Alpha of 2- and 4-channel input: fringe clearing (encode_image.rs:3650,
   * 3980), then fill_invisible and snap_achromatic_xyb in the order of
   * encode_frame (src/frame/encoder.rs:152-157). */
  uint8_t *alpha = NULL;
  float *factors = NULL;
  if (channels == 2 || channels == 4) {
    alpha = (uint8_t *)malloc(plane);
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++)
        alpha[(size_t)y * width + x] = rgb[(size_t)y * row_bytes + (size_t)x * channels + channels - 1];
    if (p->flags & JXG_FLAG_LOSSY_ALPHA) jxo_alpha_fringe_clear(alpha, width, height);
    if (p->flags & JXG_FLAG_FILL_INVISIBLE) {
      for (int c = 0; c < 3; c++)
        jxo_fill_invisible_plane(j.xyb + c * plane, (size_t)width, width, height, alpha);
      if (!jxo_alpha_is_constant(alpha, plane)) {
        factors = (float *)malloc((size_t)xb * yb * sizeof(float));
        jxo_alpha_block_factors(alpha, width, height, factors);
      }
    }
    if (alpha_out) memcpy(alpha_out, alpha, plane);
  }
  if (channels <= 2)
    for (size_t k = 0; k < plane; k++) {
      j.xyb[k] = 0.0f;
      j.xyb[2 * plane + k] = j.xyb[plane + k];
    }
  run_pool(&j, 1, groups, threads);
  /* chroma-HF protection after the AQ field (src/frame/dc_groups.rs:159-167) */
  if (p->flags & JXG_FLAG_CHROMA_HF) jxo_chroma_hf_protection(j.xyb, width, height, qf, p->distance);
  run_pool(&j, 4, (height + 63) / 64, threads);
  if (selection_runs(p)) {
    j.qf_in = (uint8_t *)malloc((size_t)xb * yb);
    memcpy(j.qf_in, qf, (size_t)xb * yb);
    /* selection parallelises over DC groups in this driver; the reference
     * also splits each group into row bands, which is decision-neutral
     * (src/ac_strategy/selection.rs:943-948). */
    run_pool(&j, 2, groups, threads);
    free(j.qf_in);
  } else {
    memset(cells, JXG_CELL(JXG_DCT, 1), (size_t)xb * yb);
  }
  /* after every layout and quant-field decision (src/frame/dc_groups.rs:302-312) */
  if (factors) jxo_apply_alpha_factors(qf, cells, xb, yb, factors);
  free(factors);
  free(alpha);
  j.levels = levels;
  j.coeffs = coeffs;
  j.dc = dc;
  run_pool(&j, 3, (yb + 7) / 8, threads);
  free(j.xyb);
  pthread_mutex_destroy(&j.mu);
}

/* ------------------------------------------------------------------ */
/* The SSIM reconstruction rerank at the Fast tier (scope.rerank:        */
/* d <= FAST_RERANK_MAX_DISTANCE, src/ac_strategy.rs:71-73,410-412;      */
/* rerank_large_transforms, src/ac_strategy/selection.rs:3568-3611, and  */
/* find_rerank_downgrades, :2718-3232).  Every merged transform (2 or    */
/* more blocks) of the re-tiled layout is priced in the reconstruction   */
/* domain against its DCT8 tiling and its best subdivision into aligned  */
/* square children; where either is cheaper the merge is split.  At Fast */
/* the search scope has no rectangles and no fine mosaic, the saved      */
/* child layouts hold no 1x1 leaf other than DCT8 (those arms add        */
/* nothing), no learned prices exist, and the upgrade passes need        */
/* d >= 2.  The rerank's cached base costs, which the refinement reuses, */
/* equal the refinement's own evaluation of the same transform and       */
/* quantizer, so the refinement recomputes them.                         */

#define RERANK_META_R 2.6465739748323758f
#define FAST_RERANK_MAX_DISTANCE 1.7090991223128462f
#define NO_CHILD 0xFF

/* reconstruction_strategy_cost_and_base's cost (src/ac_strategy.rs:
 * 774-1052, unbounded): the selection quantizer's scale on qac, the
 * reconstruction distortion and rate, rd_cost with the metadata charge. */
static float rerank_cost(group_ctx *g, int s, int bx, int by, float qac) {
  static __thread float co[3 * 4096];
  float cmap[3];
  cmap_factors(g, bx, by, cmap);
  const int px = (g->gx0 + bx) * 8, py = (g->gy0 + by) * 8;
  jxo_recon_coeffs(g->xyb, g->width, g->height, s, px, py, cmap, co);
  const int kind = jxo_qt_kind(s);
  const uint32_t n = jxo_qt_size(kind), off = jxo_qt_offset(kind);
  qac *= selection_scale(s, co + 4096, g->tables + off + 4 * n, qac, g->p, RERANK_META_R);
  float dist, rate;
  jxo_recon_cost(g->xyb, g->width, g->height, s, px, py, qac, cmap, co, g->tables, g->p, &dist,
                 &rate);
  return jxo_fmla(jxo_reconstruction_lambda(g->p->distance), rate + RERANK_META_R, dist);
}

typedef struct {
  int x, y, w, h;
  float cost;
} rr_tile;

typedef struct {
  float cost;
  uint8_t tiles[16];
} rr_plan;

static int ilog2i(int v) {
  int r = 0;
  while ((1 << (r + 1)) <= v) r++;
  return r;
}

static rr_plan rr_join(rr_plan a, rr_plan b) {
  for (int i = 0; i < 16; i++)
    if (b.tiles[i] != NO_CHILD) a.tiles[i] = b.tiles[i];
  a.cost = a.cost + b.cost;
  return a;
}

/* best_reconstruction_plan (src/ac_strategy/selection.rs:2636-2716): the
 * cheapest tiling of a width x height block footprint by aligned
 * power-of-two tiles, halving horizontally first, then vertically, the
 * single tile last; strict comparisons keep the first candidate. */
static rr_plan best_reconstruction_plan(const rr_tile *tiles, int nt, int width, int height) {
  rr_plan empty;
  empty.cost = INFINITY;
  memset(empty.tiles, NO_CHILD, 16);
  rr_plan plans[9 * 16];
  for (int i = 0; i < 9 * 16; i++) plans[i] = empty;
#define RR_INDEX(lx, ly, x, y) (((ly) * 3 + (lx)) * 16 + (y) * 4 + (x))
  for (int i = 0; i < nt; i++) {
    const rr_tile t = tiles[i];
    if ((t.w & (t.w - 1)) || (t.h & (t.h - 1)) || t.w > width || t.h > height || t.x >= width ||
        t.y >= height || t.x % t.w || t.y % t.h)
      continue;
    const int slot = RR_INDEX(ilog2i(t.w), ilog2i(t.h), t.x, t.y);
    if (t.cost < plans[slot].cost) {
      plans[slot].cost = t.cost;
      plans[slot].tiles[t.y * 4 + t.x] = (uint8_t)i;
    }
  }
  const int lw = ilog2i(width), lh = ilog2i(height);
  for (int ly = 0; ly <= lh; ly++) {
    const int h = 1 << ly;
    for (int lx = 0; lx <= lw; lx++) {
      const int w = 1 << lx;
      for (int y = 0; y < height; y += h)
        for (int x = 0; x < width; x += w) {
          rr_plan best = empty;
          if (lx > 0)
            best = rr_join(plans[RR_INDEX(lx - 1, ly, x, y)], plans[RR_INDEX(lx - 1, ly, x + w / 2, y)]);
          if (ly > 0) {
            rr_plan split =
                rr_join(plans[RR_INDEX(lx, ly - 1, x, y)], plans[RR_INDEX(lx, ly - 1, x, y + h / 2)]);
            if (split.cost < best.cost) best = split;
          }
          const int slot = RR_INDEX(lx, ly, x, y);
          if (plans[slot].cost < best.cost) best = plans[slot];
          plans[slot] = best;
        }
    }
  }
  return plans[RR_INDEX(lw, lh, 0, 0)];
#undef RR_INDEX
}

typedef struct {
  int bx, by, cx, cy;
  uint8_t restore[16];
} rr_downgrade;

/* The decision for the merge s at (bx, by): 1 and the restore grid when it
 * is split. */
static int rerank_merge(group_ctx *g, int s, int bx, int by, float margin, uint8_t restore[16]) {
  const jxg_frame_params *p = g->p;
  const int cxb = jxo_covered_x(s), cyb = jxo_covered_y(s);
  const float j_big = rerank_cost(g, s, bx, by, p->scale * (float)region_quant(g, bx, by, cxb, cyb));
  rr_tile tiles[24];
  uint8_t strat[24];
  int nt = 0;
  float j_dct8 = 0.0f;
  for (int iy = 0; iy < cyb; iy++)
    for (int ix = 0; ix < cxb; ix++) {
      const float c = rerank_cost(g, JXG_DCT, bx + ix, by + iy,
                                  p->scale * (float)region_quant(g, bx + ix, by + iy, 1, 1));
      tiles[nt] = (rr_tile){ix, iy, 1, 1, c};
      strat[nt++] = JXG_DCT;
      j_dct8 += c;
    }
  /* the children: squares only (no rectangles in this scope), aligned
   * inside the footprint and smaller than it */
  for (int iy = 0; iy < cyb; iy++)
    for (int ix = 0; ix < cxb; ix++) {
      const int cs = JXG_DCT16X16, ccx = 2, ccy = 2;
      if (ix + ccx > cxb || iy + ccy > cyb || ix % ccx || iy % ccy || (ccx == cxb && ccy == cyb) ||
          !can_place(bx + ix, by + iy, cs, g->xsize, g->ysize))
        continue;
      const float c = rerank_cost(g, cs, bx + ix, by + iy,
                                  p->scale * (float)region_quant(g, bx + ix, by + iy, ccx, ccy));
      tiles[nt] = (rr_tile){ix, iy, ccx, ccy, c};
      strat[nt++] = (uint8_t)cs;
    }
  const rr_plan plan = best_reconstruction_plan(tiles, nt, cxb, cyb);
  const int child_wins = plan.cost < j_dct8;
  const float j_alt = child_wins ? plan.cost : j_dct8;
  if (!(j_alt < j_big * margin)) return 0;
  memset(restore, NO_CHILD, 16);
  for (int iy = 0; iy < cyb; iy++)
    for (int ix = 0; ix < cxb; ix++) {
      if (child_wins) {
        const uint8_t t = plan.tiles[iy * 4 + ix];
        if (t != NO_CHILD) restore[iy * 4 + ix] = strat[t];
      } else {
        restore[iy * 4 + ix] = JXG_DCT;
      }
    }
  return 1;
}

static int rerank_runs(const jxg_frame_params *p) {
  return p->speed == JXG_SPEED_FAST && p->distance <= FAST_RERANK_MAX_DISTANCE;
}

static int g_rerank_off; /* tests: the decision without the rerank */
void jxo_rerank_enable(int on) { g_rerank_off = !on; }

static void rerank_large_transforms(group_ctx *g) {
  if (g_rerank_off || !rerank_runs(g->p)) return;
  /* MergeTuning::rerank_margin: 0 without the arbiter's rerank downgrades
   * (src/ac_strategy.rs:320,346-348; src/arbiter.rs:50-62) */
  const float margin = g->p->arbiter == JXG_ARBITER_SSIMULACRA2 ? 1.0f : 0.0f;
  rr_downgrade *dg = (rr_downgrade *)malloc(sizeof(rr_downgrade) * (size_t)g->xsize * g->ysize);
  int nd = 0;
  /* every merge is judged on the layout before any split */
  for (int by = 0; by < g->ysize; by++)
    for (int bx = 0; bx < g->xsize; bx++) {
      const uint8_t cell = g->cells[(size_t)(g->gy0 + by) * g->qf_stride + g->gx0 + bx];
      if (!(cell & 1)) continue;
      const int s = cell >> 1;
      if (s == JXG_DCT64X64 || s == JXG_DCT64X32 || s == JXG_DCT32X64) continue; /* recon64, not at Fast */
      if (jxo_covered_x(s) * jxo_covered_y(s) <= 1) continue;
      if (rerank_merge(g, s, bx, by, margin, dg[nd].restore)) {
        dg[nd].bx = bx;
        dg[nd].by = by;
        dg[nd].cx = jxo_covered_x(s);
        dg[nd].cy = jxo_covered_y(s);
        nd++;
      }
    }
  for (int i = 0; i < nd; i++)
    for (int iy = 0; iy < dg[i].cy; iy++)
      for (int ix = 0; ix < dg[i].cx; ix++) {
        const uint8_t r = dg[i].restore[iy * 4 + ix];
        if (r != NO_CHILD) set_first(g, dg[i].bx + ix, dg[i].by + iy, r);
      }
  free(dg);
}

/* Test entry points. */
float jxo_rerank_plan(const float *tile_xywh_cost, int nt, int width, int height,
                      uint8_t *tiles_out) {
  rr_tile t[64];
  for (int i = 0; i < nt && i < 64; i++)
    t[i] = (rr_tile){(int)tile_xywh_cost[5 * i], (int)tile_xywh_cost[5 * i + 1],
                     (int)tile_xywh_cost[5 * i + 2], (int)tile_xywh_cost[5 * i + 3],
                     tile_xywh_cost[5 * i + 4]};
  rr_plan r = best_reconstruction_plan(t, nt < 64 ? nt : 64, width, height);
  memcpy(tiles_out, r.tiles, 16);
  return r.cost;
}

/* The rerank pass (Fast, d <= 1.709) over a given layout and field, in
 * place. */
void jxo_rerank_pass(const float *xyb, int width, int height, const uint8_t *qf,
                     const int8_t *ytox, const int8_t *ytob, const float *tables,
                     const jxg_frame_params *p, uint8_t *cells) {
  jxo_set_fused(p->fused);
  for (int gy = 0; gy * 2048 < height; gy++)
    for (int gx = 0; gx * 2048 < width; gx++) {
      group_ctx g;
      group_init(&g, xyb, width, height, gx, gy, qf, ytox, ytob, tables, p, cells);
      rerank_large_transforms(&g);
    }
}
