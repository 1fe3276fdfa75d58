/*
 * jxo_cfl.c - oracle: chroma-from-luma slope maps.  TEST INFRASTRUCTURE ONLY
 * (see jxo.h).
 *
 * fill_cmap (src/color_correlation.rs:901-960) on its regression path
 * (compute_cmap_tile :441-508), the one XYB frames take at Fast and faster
 * (cfl_slope_search :965-972 is false there).  Tiles are 8x8 blocks; the
 * regression path reads no neighbouring tile, so the reference's anti-diagonal
 * wavefront equals this raster order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "jxo.h"
#include "jxo_internal.h"

static const float K_INV_COLOR_FACTOR = 1.0f / 84.0f; /* :46 */
static const float K_DISTANCE_MULTIPLIER_AC = 1e-3f;  /* :48 */

/* cfl_deadzone (:383-394) */
static float cfl_deadzone(float distance) {
  const float AMOUNT = 1.5f, LO = 1.0f, HI = 2.0f;
  if (distance <= LO) return AMOUNT;
  if (distance >= HI) return 0.0f;
  const float SCALE = AMOUNT / (HI - LO);
  return SCALE * (HI - distance);
}

/* solve_multiplier (:397-409) */
static int solve_multiplier(float ca, float cb, int num, float distance_mul, float dz,
                            float units) {
  if (num == 0) return 0;
  float x = -cb / jxo_fmla((float)num * distance_mul, 0.5f, ca);
  if (fabsf(x) < dz) x = 0.0f;
  float r = roundf(x * units);
  if (r < -128.0f) r = -128.0f;
  if (r > 127.0f) r = 127.0f;
  return (int)r;
}

/* cfl_regression_scalar (:333-357) */
static void cfl_regression(const float *y, const float *x, const float *b, const float *qm_x,
                           const float *qm_b, float sums[4]) {
  sums[0] = sums[1] = sums[2] = sums[3] = 0.0f;
  for (int i = 0; i < 64; i++) {
    const float m_x = y[i] * qm_x[i];
    const float s_x = x[i] * qm_x[i];
    const float a_x = K_INV_COLOR_FACTOR * m_x;
    const float b_x = -s_x;
    sums[0] = jxo_fmla(a_x, a_x, sums[0]);
    sums[1] = jxo_fmla(a_x, b_x, sums[1]);
    const float m_b = y[i] * qm_b[i];
    const float s_b = b[i] * qm_b[i];
    const float a_b = K_INV_COLOR_FACTOR * m_b;
    const float b_b = jxo_fmla(1.0f, m_b, -s_b);
    sums[2] = jxo_fmla(a_b, a_b, sums[2]);
    sums[3] = jxo_fmla(a_b, b_b, sums[3]);
  }
}

void jxo_cmap_tile(const float *xyb, int width, int height, int tx, int ty, const float *tables,
                   const jxg_frame_params *p, int *ytox, int *ytob) {
  const size_t plane = (size_t)width * height;
  const int xb = (width + 7) / 8, yb = (height + 7) / 8;
  const int bx0 = tx * 8, by0 = ty * 8;
  const int bx_count = xb - bx0 < 8 ? xb - bx0 : 8;
  const int by_count = yb - by0 < 8 ? yb - by0 : 8;
  const int kind = jxo_qt_kind(JXG_DCT);
  const float *qm_x = tables + jxo_qt_offset(kind) + 3 * 64; /* inv_matrix(0) */
  const float *qm_b = tables + jxo_qt_offset(kind) + 5 * 64; /* inv_matrix(2) */
  float ca_x = 0.0f, cb_x = 0.0f, ca_b = 0.0f, cb_b = 0.0f;
  int num = 0;
  for (int by = 0; by < by_count; by++) {
    for (int bx = 0; bx < bx_count; bx++) {
      const int px = (bx0 + bx) * 8, py = (by0 + by) * 8;
      if (px + 8 > width || py + 8 > height) continue;
      float blk_y[64], blk_x[64], blk_b[64], sums[4];
      jxo_forward_transform(JXG_DCT, xyb + plane, width, height, px, py, blk_y);
      jxo_forward_transform(JXG_DCT, xyb, width, height, px, py, blk_x);
      jxo_forward_transform(JXG_DCT, xyb + 2 * plane, width, height, px, py, blk_b);
      blk_y[0] = blk_x[0] = blk_b[0] = 0.0f;
      /* shift_to_xyb_bases (:413-429) */
      const float sx = -p->cfl_base_x, sb = 1.0f - p->cfl_base_b;
      if (sx != 0.0f)
        for (int i = 0; i < 64; i++) blk_x[i] = jxo_fmla(sx, blk_y[i], blk_x[i]);
      if (sb != 0.0f)
        for (int i = 0; i < 64; i++) blk_b[i] = jxo_fmla(sb, blk_y[i], blk_b[i]);
      cfl_regression(blk_y, blk_x, blk_b, qm_x, qm_b, sums);
      ca_x += sums[0];
      cb_x += sums[1];
      ca_b += sums[2];
      cb_b += sums[3];
      num += 64;
    }
  }
  const float dz = cfl_deadzone(p->distance);
  const float units = p->cfl_color_factor / 84.0f;
  *ytox = solve_multiplier(ca_x, cb_x, num, K_DISTANCE_MULTIPLIER_AC, dz, units);
  *ytob = solve_multiplier(ca_b, cb_b, num, K_DISTANCE_MULTIPLIER_AC, dz, units);
}

void jxo_fill_cmap(const float *xyb, int width, int height, const float *tables,
                   const jxg_frame_params *p, int8_t *ytox, int8_t *ytob) {
  const int tiles_x = (width + 63) / 64, tiles_y = (height + 63) / 64;
  for (int ty = 0; ty < tiles_y; ty++)
    for (int tx = 0; tx < tiles_x; tx++) {
      int x, b;
      jxo_cmap_tile(xyb, width, height, tx, ty, tables, p, &x, &b);
      ytox[ty * tiles_x + tx] = (int8_t)x;
      ytob[ty * tiles_x + tx] = (int8_t)b;
    }
}
