/*
 * jxo_epf.c - oracle: the six-term EPF luma proxy (LumaProxy,
 * src/epf_proxy.rs:151-245, box_filter.rs, row_kernels.rs).  TEST
 * INFRASTRUCTURE ONLY (see jxo.h).
 *
 * The 5x5 box moments are computed in the form box_filter.rs's own tests
 * pin the streaming kernels against (original_box5, box_filter.rs:466-488):
 * a running vertical sum per column, then a running horizontal sum per row,
 * with mirrored edges.  Pooled sums run in pixel order in f32 like the
 * reference, or in f64 when `wide` is set (the accumulation the GPU form
 * uses; every per-pixel term is identical).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "jxo.h"
#include "jxo_internal.h"

struct jxo_epf {
  int w, h;
  float *src;
  int hw, hh, qw, qh;
  float *hv, *hm, *hvar; /* half scale: values, box mean, box variance */
  float *qv, *qm, *qvar; /* quarter scale */
};

/* mirror (src/epf_proxy.rs:17-28) */
static int mirror(int i, int n) {
  if (i >= 0 && i < n) return i;
  int m = i % (2 * n);
  if (m < 0) m += 2 * n;
  return m < n ? m : 2 * n - m - 1;
}

/* half / half_y (row_kernels.rs:60-70): 2x2 means, with +0.01 per tap for
 * the source-luma pyramid's first level. */
static void half(const float *src, int w, int h, int bias, float *dst) {
  const int ow = w / 2, oh = h / 2;
  for (int y = 0; y < oh; y++)
    for (int x = 0; x < ow; x++) {
      const float *r0 = src + (size_t)(2 * y) * w, *r1 = src + (size_t)(2 * y + 1) * w;
      float t00 = r0[2 * x], t10 = r1[2 * x], t01 = r0[2 * x + 1], t11 = r1[2 * x + 1];
      if (bias) {
        t00 += 0.01f;
        t10 += 0.01f;
        t01 += 0.01f;
        t11 += 0.01f;
      }
      dst[(size_t)y * ow + x] = 0.25f * (t00 + t10 + t01 + t11);
    }
}

/* original_box5 (box_filter.rs:466-488); sums start at -0.0 like the
 * streaming kernels' sums (:118) and Rust's f32 Sum. */
static void box5(const float *src, int w, int h, float *dst) {
  float *tmp = (float *)malloc((size_t)w * h * sizeof(float));
  for (int x = 0; x < w; x++) {
    float sum = -0.0f;
    for (int y = -2; y <= 2; y++) sum += src[(size_t)mirror(y, h) * w + x];
    for (int y = 0; y < h; y++) {
      tmp[(size_t)y * w + x] = sum * 0.2f;
      sum += src[(size_t)mirror(y + 3, h) * w + x] - src[(size_t)mirror(y - 2, h) * w + x];
    }
  }
  for (int y = 0; y < h; y++) {
    const float *row = tmp + (size_t)y * w;
    float sum = -0.0f;
    for (int x = -2; x <= 2; x++) sum += row[mirror(x, w)];
    for (int x = 0; x < w; x++) {
      dst[(size_t)y * w + x] = sum * 0.2f;
      sum += row[mirror(x + 3, w)] - row[mirror(x - 2, w)];
    }
  }
  free(tmp);
}

void jxo_epf_box5(const float *src, int w, int h, float *dst) { box5(src, w, h, dst); }

/* Scale::new (src/epf_proxy.rs:75-97): box mean and variance of a scale. */
static void scale_new(const float *v, int w, int h, float *mean, float *var) {
  const size_t n = (size_t)w * h;
  float *sq = (float *)calloc(n ? n : 1, sizeof(float)), *bsq = (float *)malloc(n * sizeof(float));
  for (size_t i = 0; i < n; i++) sq[i] = v[i] * v[i];
  box5(v, w, h, mean);
  box5(sq, w, h, bsq);
  for (size_t i = 0; i < n; i++) { /* row_kernels.rs:80-84 */
    const float d = bsq[i] - mean[i] * mean[i];
    var[i] = d > 0.0f ? d : 0.0f;
  }
  free(sq);
  free(bsq);
}

/* Scale::compare (src/epf_proxy.rs:99-128) with pool (row_kernels.rs:86-103). */
static void scale_compare(const float *ref_v, const float *ref_mean, const float *ref_var,
                          const float *test, int w, int h, int wide, float out[2]) {
  const size_t n = (size_t)w * h;
  float *sq = (float *)calloc(n ? n : 1, sizeof(float)), *cr = (float *)calloc(n ? n : 1, sizeof(float));
  float *bm = (float *)malloc(n * sizeof(float)), *bsq = (float *)malloc(n * sizeof(float));
  float *bcr = (float *)malloc(n * sizeof(float));
  for (size_t i = 0; i < n; i++) {
    sq[i] = test[i] * test[i];
    cr[i] = ref_v[i] * test[i];
  }
  box5(test, w, h, bm);
  box5(sq, w, h, bsq);
  box5(cr, w, h, bcr);
  float s0 = 0.0f, s1 = 0.0f;
  double d0 = 0.0, d1 = 0.0;
  for (size_t i = 0; i < n; i++) {
    const float a = ref_mean[i], b = bm[i];
    float var = bsq[i] - b * b;
    if (!(var > 0.0f)) var = 0.0f;
    const float cov = (ref_var[i] == 0.0f || var == 0.0f) ? 0.0f : bcr[i] - a * b;
    float s = 1.0f - (1.0f - (a - b) * (a - b)) * (2.0f * cov + 0.0009f) / (ref_var[i] + var + 0.0009f);
    if (!(s > 0.0f)) s = 0.0f;
    const float s4 = s * s * s * s;
    s0 += s;
    s1 += s4;
    d0 += s;
    d1 += s4;
  }
  if (wide) {
    s0 = (float)d0;
    s1 = (float)d1;
  }
  const float nf = (float)n;
  out[0] = s0 / nf;
  out[1] = sqrtf(sqrtf(s1 / nf));
  free(sq);
  free(cr);
  free(bm);
  free(bsq);
  free(bcr);
}

jxo_epf *jxo_epf_new(const float *y, int w, int h) {
  jxo_epf *p = (jxo_epf *)calloc(1, sizeof(jxo_epf));
  p->w = w;
  p->h = h;
  p->src = (float *)malloc((size_t)w * h * sizeof(float));
  memcpy(p->src, y, (size_t)w * h * sizeof(float));
  p->hw = w / 2;
  p->hh = h / 2;
  p->qw = p->hw / 2;
  p->qh = p->hh / 2;
  const size_t hn = (size_t)p->hw * p->hh, qn = (size_t)p->qw * p->qh;
  p->hv = (float *)malloc(hn * sizeof(float));
  p->hm = (float *)malloc(hn * sizeof(float));
  p->hvar = (float *)malloc(hn * sizeof(float));
  p->qv = (float *)malloc(qn * sizeof(float));
  p->qm = (float *)malloc(qn * sizeof(float));
  p->qvar = (float *)malloc(qn * sizeof(float));
  half(y, w, h, 1, p->hv);
  half(p->hv, p->hw, p->hh, 0, p->qv);
  scale_new(p->hv, p->hw, p->hh, p->hm, p->hvar);
  scale_new(p->qv, p->qw, p->qh, p->qm, p->qvar);
  return p;
}

void jxo_epf_free(jxo_epf *p) {
  if (!p) return;
  free(p->src);
  free(p->hv);
  free(p->hm);
  free(p->hvar);
  free(p->qv);
  free(p->qm);
  free(p->qvar);
  free(p);
}

/* LumaProxy::compare (src/epf_proxy.rs:173-215) with gradients (:218-245,
 * row_kernels.rs:104-114). */
float jxo_epf_compare(const jxo_epf *p, const float *y, int wide, float terms[6]) {
  const int w = p->w, h = p->h;
  float *hy = (float *)malloc((size_t)p->hw * p->hh * sizeof(float));
  float *qy = (float *)malloc((size_t)p->qw * p->qh * sizeof(float));
  half(y, w, h, 1, hy);
  half(hy, p->hw, p->hh, 0, qy); /* half_in_place gives the same values */
  float hs[2], qs[2];
  scale_compare(p->hv, p->hm, p->hvar, hy, p->hw, p->hh, wide, hs);
  scale_compare(p->qv, p->qm, p->qvar, qy, p->qw, p->qh, wide, qs);
  float g[3] = {0.0f, 0.0f, 0.0f};
  double gd[3] = {0.0, 0.0, 0.0};
  for (int pass = 0; pass < 2; pass++) {
    const int rows = pass == 0 ? h : h - 1, cols = pass == 0 ? w - 1 : w;
    for (int r = 0; r < rows; r++)
      for (int x = 0; x < cols; x++) {
        const size_t i0 = (size_t)r * w + x, i1 = pass == 0 ? i0 + 1 : i0 + w;
        const float a = p->src[i1] - p->src[i0];
        const float b = y[i1] - y[i0];
        const float e = fabsf(b - a);
        float ex = fabsf(b) - fabsf(a);
        if (!(ex > 0.0f)) ex = 0.0f;
        float loss = fabsf(a) - fabsf(b);
        if (!(loss > 0.0f)) loss = 0.0f;
        const float l4 = loss * loss * loss * loss;
        g[0] += e;
        g[1] += ex;
        g[2] += l4;
        gd[0] += e;
        gd[1] += ex;
        gd[2] += l4;
      }
  }
  if (wide)
    for (int i = 0; i < 3; i++) g[i] = (float)gd[i];
  const float n = (float)((size_t)h * (w - 1) + (size_t)(h - 1) * w);
  terms[0] = qs[0];
  terms[1] = hs[0];
  terms[2] = hs[1];
  terms[3] = g[0] / n;
  terms[4] = g[1] / n;
  terms[5] = sqrtf(sqrtf(g[2] / n));
  static const float WEIGHTS[6] = {109.4163228369586f, 4.685925102200196f, 2.491781270704224f,
                                   2.167275675983448f, 4.55227651921194f,  2.847714678987977f};
  float score = -0.0f;
  for (int i = 0; i < 6; i++) score += terms[i] * WEIGHTS[i];
  free(hy);
  free(qy);
  return score;
}
