/*
 * jxo_alpha.c - oracle: the alpha-dependent stages of the RGBA / gray+alpha
 * front end.  TEST INFRASTRUCTURE ONLY (see jxo.h).
 *
 *   jxo_alpha_fringe_clear    alpha_fringe::clear (src/alpha_fringe.rs:77-103)
 *                             with sweep (:105-225), u8 samples, max 255
 *   jxo_fill_invisible_plane  fill_invisible (src/invisible.rs:34-146)
 *   jxo_alpha_block_factors   alpha_block_factors (src/adaptive_quant.rs:953-1030)
 *   jxo_apply_alpha_factors   apply_alpha_factors (src/adaptive_quant.rs:1036-1070)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "jxo.h"

/* ------------------------------------------------------------------ */
/* alpha_fringe::clear                                                  */

enum { FR_REACH = 4, FR_RADIUS = 4, FR_MEAN_DIVISOR = 2, FR_FILL_SHARE = 12 };

/* fringe_level of sweep (src/alpha_fringe.rs:113-117): the sample's
 * distance (its level, or max - level) when within reach, else 0. */
static uint32_t fringe_level(uint8_t a, int inverted, uint32_t reach) {
  const uint32_t level = inverted ? 255u - a : a;
  return level <= reach ? level : 0u;
}

/* sweep (src/alpha_fringe.rs:105-225): a running 9x9 window (rows
 * [y-4, y+5) and columns [x-4, x+5), clipped to the image) of fringe levels
 * over the rows between the first and last fringe sample.  A fringe sample
 * whose window mean is at most max / (2 * 255) is replaced, one row after it
 * has left the window, so every window sees original samples. */
static size_t fringe_sweep(uint8_t *alpha, int width, int height, int inverted,
                           uint8_t replacement) {
  const uint32_t reach = FR_REACH * 255u / 255u;
  const size_t n = (size_t)width * height;
  size_t first = n, last = 0;
  for (size_t i = 0; i < n; i++)
    if (fringe_level(alpha[i], inverted, reach) != 0) {
      first = i;
      break;
    }
  if (first == n) return 0;
  for (size_t i = n; i-- > first;)
    if (fringe_level(alpha[i], inverted, reach) != 0) {
      last = i;
      break;
    }
  const int first_row = (int)(first / width), last_row = (int)(last / width);
  const int pending_rows = FR_RADIUS + 1;
  uint32_t *columns = (uint32_t *)calloc((size_t)width, sizeof(uint32_t));
  uint8_t *pending = (uint8_t *)calloc((size_t)width, 1);
  int top = first_row > FR_RADIUS ? first_row - FR_RADIUS : 0;
  int bottom = first_row + FR_RADIUS + 1 < height ? first_row + FR_RADIUS + 1 : height;
  for (int y = top; y < bottom; y++)
    for (int x = 0; x < width; x++)
      columns[x] += fringe_level(alpha[(size_t)y * width + x], inverted, reach);
  size_t changed = 0;
  for (int y = first_row; y <= last_row; y++) {
    const uint8_t *row = alpha + (size_t)y * width;
    const uint8_t bit = (uint8_t)(1u << (y % pending_rows));
    int next_x = -1;
    uint32_t sum = 0;
    for (int x = 0; x < width; x++) {
      if (fringe_level(row[x], inverted, reach) == 0) continue;
      const int x0 = x > FR_RADIUS ? x - FR_RADIUS : 0;
      const int x1 = x + FR_RADIUS + 1 < width ? x + FR_RADIUS + 1 : width;
      if (x == next_x) {
        if (x > FR_RADIUS) sum -= columns[x0 - 1];
        if (x + FR_RADIUS < width) sum += columns[x1 - 1];
      } else {
        sum = 0;
        for (int k = x0; k < x1; k++) sum += columns[k];
      }
      next_x = x + 1;
      if ((uint64_t)sum * FR_MEAN_DIVISOR * 255u <=
          (uint64_t)(bottom - top) * (uint64_t)(x1 - x0) * 255u)
        pending[x] |= bit;
    }
    if (y == last_row) break;
    if (y >= FR_RADIUS) {
      const uint8_t tbit = (uint8_t)(1u << (top % pending_rows));
      uint8_t *trow = alpha + (size_t)top * width;
      for (int x = 0; x < width; x++) {
        columns[x] -= fringe_level(trow[x], inverted, reach);
        if (pending[x] & tbit) {
          pending[x] &= (uint8_t)~tbit;
          trow[x] = replacement;
          changed++;
        }
      }
      top++;
    }
    if (bottom < height) {
      const uint8_t *brow = alpha + (size_t)bottom * width;
      for (int x = 0; x < width; x++) columns[x] += fringe_level(brow[x], inverted, reach);
      bottom++;
    }
  }
  for (int y = top; y <= last_row; y++) {
    const uint8_t tbit = (uint8_t)(1u << (y % pending_rows));
    uint8_t *row = alpha + (size_t)y * width;
    for (int x = 0; x < width; x++)
      if (pending[x] & tbit) {
        row[x] = replacement;
        changed++;
      }
  }
  free(columns);
  free(pending);
  return changed;
}

void jxo_alpha_fringe_clear(uint8_t *alpha, int width, int height) {
  if (width <= 0 || height <= 0) return;
  const size_t n = (size_t)width * height;
  fringe_sweep(alpha, width, height, 0, 0);
  const uint32_t reach = FR_REACH * 255u / 255u;
  size_t near_opaque = 0;
  for (size_t i = 0; i < n; i++) {
    const uint32_t d = 255u - alpha[i];
    if (d >= 1 && d <= reach) near_opaque++;
  }
  if (near_opaque * FR_FILL_SHARE < n) return;
  uint8_t *filled = (uint8_t *)malloc(n);
  memcpy(filled, alpha, n);
  const size_t changed = fringe_sweep(filled, width, height, 1, 255);
  if (changed * FR_FILL_SHARE >= n) memcpy(alpha, filled, n);
  free(filled);
}

/* ------------------------------------------------------------------ */
/* fill_invisible: every alpha-0 sample, in row-major order, becomes the  */
/* weighted mean of its left neighbour (weight 1, 2 when visible), its    */
/* upper-right neighbour (1, plus 2 when visible) and the visible right,  */
/* lower-right, upper and lower neighbours (2 each); earlier fills feed   */
/* later ones (fill_row, src/invisible.rs:72-146).                        */

void jxo_fill_invisible_plane(float *plane, size_t stride, int width, int height,
                              const uint8_t *alpha) {
  if (width <= 0 || height <= 0) return;
  const size_t n = (size_t)width * height;
  size_t i;
  for (i = 0; i < n; i++)
    if (alpha[i] == 0) break;
  if (i == n) return;
  for (int y = 0; y < height; y++) {
    const int up = y > 0, down = y + 1 < height;
    float *row = plane + (size_t)y * stride;
    const float *urow = up ? row - stride : row;
    const float *drow = down ? row + stride : row;
    const uint8_t *ar = alpha + (size_t)y * width;
    const uint8_t *au = up ? ar - width : ar;
    const uint8_t *ad = down ? ar + width : ar;
    for (int x = 0; x < width; x++) {
      if (ar[x] != 0) continue;
      float sum = 0.0f, weight = 0.0f;
      if (x > 0) {
        sum += row[x - 1];
        weight += 1.0f;
        if (ar[x - 1] > 0) {
          sum += row[x - 1];
          weight += 1.0f;
        }
      }
      if (x + 1 < width) {
        if (up) {
          sum += urow[x + 1];
          weight += 1.0f;
        }
        if (ar[x + 1] > 0) {
          sum += 2.0f * row[x + 1];
          weight += 2.0f;
        }
        if (up && au[x + 1] > 0) {
          sum += 2.0f * urow[x + 1];
          weight += 2.0f;
        }
        if (down && ad[x + 1] > 0) {
          sum += 2.0f * drow[x + 1];
          weight += 2.0f;
        }
      }
      if (up && au[x] > 0) {
        sum += 2.0f * urow[x];
        weight += 2.0f;
      }
      if (down && ad[x] > 0) {
        sum += 2.0f * drow[x];
        weight += 2.0f;
      }
      row[x] = weight > 1.0f ? sum / weight : sum;
    }
  }
}

/* ------------------------------------------------------------------ */
/* alpha_block_factors / apply_alpha_factors                            */

#define ALPHA_AQ_FLOOR 0.1f

int jxo_alpha_is_constant(const uint8_t *alpha, size_t n) {
  for (size_t i = 1; i < n; i++)
    if (alpha[i] != alpha[0]) return 0;
  return 1;
}

void jxo_alpha_block_factors(const uint8_t *alpha, int width, int height, float *factors) {
  const int bw = (width + 7) / 8, bh = (height + 7) / 8;
  for (int by = 0; by < bh; by++)
    for (int bx = 0; bx < bw; bx++) {
      const int x = bx * 8, end = x + 8 < width ? x + 8 : width;
      float peak = 0.0f;
      for (int y = by * 8; y < by * 8 + 8 && y < height; y++) {
        uint8_t m = 0;
        for (int k = x; k < end; k++)
          if (alpha[(size_t)y * width + k] > m) m = alpha[(size_t)y * width + k];
        peak = fmaxf(peak, (float)m);
      }
      const float a = peak / 255.0f;
      factors[(size_t)by * bw + bx] = (a == 0.0f || a >= 1.0f) ? 1.0f : fmaxf(a, ALPHA_AQ_FLOOR);
    }
}

void jxo_apply_alpha_factors(uint8_t *qf, const uint8_t *cells, int xb, int yb,
                             const float *factors) {
  for (int by = 0; by < yb; by++)
    for (int bx = 0; bx < xb; bx++) {
      const uint8_t cell = cells[(size_t)by * xb + bx];
      if (!(cell & 1)) continue;
      const int raw = cell >> 1;
      const int cx = jxo_covered_x(raw), cy = jxo_covered_y(raw);
      float k;
      if (bx + cx > xb || by + cy > yb) {
        k = 1.0f;
      } else if (cx == 1 && cy == 1) {
        k = factors[(size_t)by * xb + bx];
      } else {
        float peak = 0.0f;
        for (int iy = 0; iy < cy && peak != 1.0f; iy++)
          for (int ix = 0; ix < cx; ix++) peak = fmaxf(peak, factors[(size_t)(by + iy) * xb + bx + ix]);
        k = peak;
      }
      int q = (int)roundf((float)qf[(size_t)by * xb + bx] * k);
      q = q < 1 ? 1 : (q > 255 ? 255 : q);
      for (int iy = 0; iy < cy; iy++)
        memset(qf + (size_t)(by + iy) * xb + bx, q, (size_t)cx);
    }
}
