/*
 * jxo_adjoint.c - oracle: FP64 model and adjoint of the quant-field tuning
 * objective.  TEST INFRASTRUCTURE ONLY (see jxo.h).
 *
 * The reference has no differentiable path; the objective is this build's
 * own, stated here so the GPU kernels can be checked against it:
 *
 *   for every transform T of the layout (first block (bx, by), footprint
 *   w x h px, covered blocks n_T) of the DCT8..DCT32 family whose footprint
 *   lies inside one aligned 32x32-px region (the others, such as the
 *   re-tiling's unaligned pieces, keep their field and are left out of L),
 *   theta_T = mean of the per-block field
 *   theta over its covered blocks, and per channel c:
 *     coefficients  C = F_h X F_w^T  (F_n = C_n / n, C_n[k][j] = c_k cos(pi k (2j+1) / 2n),
 *                                     c_0 = 1, c_k = sqrt 2: the front end's scaled DCT-II)
 *     step          s_k = 1 / (inv_c[k] * scale * theta_T * qm_c)   (the quantizer's step)
 *     soft level    Chat_k = C_k + s_k u_k, u_k uniform in [-1/2, 1/2) from jxo_tune_noise,
 *                   no noise on the lowest cov_y x cov_x frequencies (the DC image)
 *     reconstruction Xhat = C_h^T Chat C_w  (the exact inverse of the forward transform)
 *     distortion    D_T = sum_c w_c sum_{px of T inside the image} (Xhat - X)^2
 *     rate proxy    R_T = sum_c sum_{k not DC} log2(1 + |Chat_k / s_k|)
 *   loss L = sum_T D_T + lambda R_T, gradient dL/dtheta_b = dL/dtheta_T / n_T for
 *   every block b of T.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "jxo.h"
#include "jxo_internal.h"

float jxo_b_qm_mul(uint32_t b_qm_scale);
float jxo_x_qm_mul(uint32_t x_qm_scale);

/* The tuning noise: lowbias32 of the packed (seed, step, channel, first
 * block, coefficient), its top 24 bits as a uniform in [-1/2, 1/2). */
static uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

float jxo_tune_noise(uint32_t seed, uint32_t step, int c, uint32_t first_block, int k) {
  uint32_t h = lowbias32(seed ^ 0x9e3779b9U);
  h = lowbias32(h ^ step);
  h = lowbias32(h ^ (uint32_t)(c * 0x10000 + k));
  h = lowbias32(h ^ first_block);
  return (float)(h >> 8) * (1.0f / 16777216.0f) - 0.5f;
}

static double cmat(int n, int k, int j) {
  const double ck = k == 0 ? 1.0 : sqrt(2.0);
  return ck * cos(M_PI * k * (2 * j + 1) / (2.0 * n));
}

/* The DCT8 / 16 / 32 family takes part; other cells are left out. */
static int kind_of(int s) {
  switch (s) {
    case JXG_DCT:
    case JXG_DCT16X8:
    case JXG_DCT8X16:
    case JXG_DCT16X16:
    case JXG_DCT32X32:
    case JXG_DCT32X16:
    case JXG_DCT16X32:
    case JXG_DCT32X8:
    case JXG_DCT8X32: return jxo_qt_kind(s);
    default: return -1;
  }
}

/* Storage index of (vertical v, horizontal u) frequency, as the quantizer
 * stores coefficients (square 8x8 u-major, rows-first shapes v-major,
 * column-first shapes u-major). */
static int storage_of(int s, int w, int h, int v, int u) {
  if (s == JXG_DCT) return u * 8 + v;
  const int rows_first = s == JXG_DCT8X16 || s == JXG_DCT16X32 || s == JXG_DCT8X32;
  if (rows_first) return v * w + u;
  return u * h + v;
}

int jxo_tune_supported(int s) { return kind_of(s) >= 0; }

/* Loss of the whole image and, when grad != NULL, dL/dtheta per block
 * (xblocks x yblocks, zero where no transform covers).  xyb: 3 W x H planes. */
double jxo_tune_loss_grad(const float *xyb, int W, int H, const uint8_t *cells,
                          const double *theta, const float *tables, const jxg_frame_params *p,
                          double lambda, uint32_t seed, uint32_t step, double *grad) {
  const int xb = (W + 7) / 8, yb = (H + 7) / 8;
  const size_t plane = (size_t)W * H;
  const double cw[3] = {p->channel_weight[0], p->channel_weight[1], p->channel_weight[2]};
  const double qm[3] = {jxo_x_qm_mul(p->x_qm_scale), 1.0, jxo_b_qm_mul(p->b_qm_scale)};
  if (grad) memset(grad, 0, (size_t)xb * yb * sizeof(double));
  double loss = 0.0;
  double *X = (double *)malloc(32 * 32 * sizeof(double));
  double *Cf = (double *)malloc(32 * 32 * sizeof(double));
  double *Ch = (double *)malloc(32 * 32 * sizeof(double));
  double *Xh = (double *)malloc(32 * 32 * sizeof(double));
  double *R = (double *)malloc(32 * 32 * sizeof(double));
  double *G = (double *)malloc(32 * 32 * sizeof(double));
  double *T = (double *)malloc(32 * 32 * sizeof(double));
  for (int by = 0; by < yb; by++)
    for (int bx = 0; bx < xb; bx++) {
      const uint8_t cell = cells[(size_t)by * xb + bx];
      if (!(cell & 1)) continue;
      const int s = cell >> 1;
      if (kind_of(s) < 0) continue;
      const int cvx = jxo_covered_x(s), cvy = jxo_covered_y(s);
      if ((bx & 3) + cvx > 4 || (by & 3) + cvy > 4) continue;  /* crosses a 32x32 region */
      const int w = cvx * 8, h = cvy * 8;
      double th = 0.0;
      for (int iy = 0; iy < cvy; iy++)
        for (int ix = 0; ix < cvx; ix++) th += theta[(size_t)(by + iy) * xb + bx + ix];
      const int ncov = cvx * cvy;
      th /= ncov;
      const int kind = kind_of(s);
      const uint32_t qn = jxo_qt_size(kind), qoff = jxo_qt_offset(kind);
      const uint32_t first = (uint32_t)(by * xb + bx);
      double dth = 0.0;
      for (int c = 0; c < 3; c++) {
        const float *inv = tables + qoff + (3 + c) * qn;
        const double qs = (double)p->scale * th * qm[c];
        /* X: footprint with edge replication */
        for (int y = 0; y < h; y++)
          for (int x = 0; x < w; x++) {
            const int sy = by * 8 + y < H ? by * 8 + y : H - 1;
            const int sx = bx * 8 + x < W ? bx * 8 + x : W - 1;
            X[y * w + x] = xyb[c * plane + (size_t)sy * W + sx];
          }
        /* C = F_h X F_w^T */
        for (int y = 0; y < h; y++)
          for (int u = 0; u < w; u++) {
            double a = 0.0;
            for (int x = 0; x < w; x++) a += cmat(w, u, x) * X[y * w + x];
            T[y * w + u] = a / w;
          }
        for (int v = 0; v < h; v++)
          for (int u = 0; u < w; u++) {
            double a = 0.0;
            for (int y = 0; y < h; y++) a += cmat(h, v, y) * T[y * w + u];
            Cf[v * w + u] = a / h;
          }
        /* soft quantization and rate */
        double drate = 0.0;
        for (int v = 0; v < h; v++)
          for (int u = 0; u < w; u++) {
            const int i = v * w + u;
            if (v < cvy && u < cvx) {
              Ch[i] = Cf[i];
              continue;
            }
            const int k = storage_of(s, w, h, v, u);
            const double iv = inv[k];
            const double st = 1.0 / (iv * qs);
            const double un = jxo_tune_noise(seed, step, c, first, k);
            Ch[i] = Cf[i] + st * un;
            const double q = Ch[i] / st;
            loss += lambda * log2(1.0 + fabs(q));
            /* d/dtheta log2(1 + |q|), q = Cf * iv * qs + u, dq/dtheta = (q - u) / theta */
            drate += (q >= 0.0 ? 1.0 : -1.0) / ((1.0 + fabs(q)) * M_LN2) * (q - un) / th;
          }
        /* Xhat = C_h^T Chat C_w */
        for (int v = 0; v < h; v++)
          for (int x = 0; x < w; x++) {
            double a = 0.0;
            for (int u = 0; u < w; u++) a += cmat(w, u, x) * Ch[v * w + u];
            T[v * w + x] = a;
          }
        for (int y = 0; y < h; y++)
          for (int x = 0; x < w; x++) {
            double a = 0.0;
            for (int v = 0; v < h; v++) a += cmat(h, v, y) * T[v * w + x];
            Xh[y * w + x] = a;
          }
        /* distortion over in-image pixels; residual 2 w_c (Xhat - X) */
        for (int y = 0; y < h; y++)
          for (int x = 0; x < w; x++) {
            const int in = by * 8 + y < H && bx * 8 + x < W;
            const double d = Xh[y * w + x] - X[y * w + x];
            if (in) loss += cw[c] * d * d;
            R[y * w + x] = in ? 2.0 * cw[c] * d : 0.0;
          }
        /* dL/dChat = C_h R C_w^T */
        for (int y = 0; y < h; y++)
          for (int u = 0; u < w; u++) {
            double a = 0.0;
            for (int x = 0; x < w; x++) a += cmat(w, u, x) * R[y * w + x];
            T[y * w + u] = a;
          }
        for (int v = 0; v < h; v++)
          for (int u = 0; u < w; u++) {
            double a = 0.0;
            for (int y = 0; y < h; y++) a += cmat(h, v, y) * T[y * w + u];
            G[v * w + u] = a;
          }
        /* dChat/dtheta = u * ds/dtheta = -u s / theta */
        for (int v = 0; v < h; v++)
          for (int u = 0; u < w; u++) {
            if (v < cvy && u < cvx) continue;
            const int k = storage_of(s, w, h, v, u);
            const double st = 1.0 / ((double)inv[k] * qs);
            const double un = jxo_tune_noise(seed, step, c, first, k);
            dth += G[v * w + u] * (-un * st / th);
          }
        dth += lambda * drate;
      }
      if (grad)
        for (int iy = 0; iy < cvy; iy++)
          for (int ix = 0; ix < cvx; ix++) grad[(size_t)(by + iy) * xb + bx + ix] = dth / ncov;
    }
  free(X);
  free(Cf);
  free(Ch);
  free(Xh);
  free(R);
  free(G);
  free(T);
  return loss;
}
