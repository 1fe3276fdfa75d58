/*
 * jxo_filters.c - oracle: whole-plane filters of the front end (gaborish
 * pre-inversion and the EPF luma proxy).  TEST INFRASTRUCTURE ONLY (see jxo.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "jxo.h"
#include "jxo_internal.h"

/* ------------------------------------------------------------------ */
/* gaborish_inverse (src/gaborish.rs:44-77): normalised symmetric 5x5   */
/* kernel, mirror-reflected edges, every output from source pixels.     */

static const float K_GAB[5] = {
    -0.09495815671340026f, -0.041031725066768575f, 0.013710004822696948f,
    0.006510206083837737f, -0.0014789063378272242f,
};

/* reflect (src/gaborish.rs:80-91). */
static int gab_reflect(int i, int len) {
  int j = i;
  while (j < 0 || j >= len) {
    if (j < 0)
      j = -j - 1;
    else
      j = 2 * len - j - 1;
  }
  return j;
}

void jxo_gaborish_weights(float mul, float w[6]) {
  const float w_r = mul * K_GAB[0];
  const float w_d = mul * K_GAB[1];
  const float w_r2 = mul * K_GAB[2];
  const float w_l = mul * K_GAB[3];
  const float w_d2 = mul * K_GAB[4];
  float sum = 1.0f + mul * 4.0f * (K_GAB[0] + K_GAB[1] + K_GAB[2] + K_GAB[4] + 2.0f * K_GAB[3]);
  if (sum < 1e-5f) sum = 1e-5f;
  const float norm = 1.0f / sum;
  w[0] = norm;        /* centre                      */
  w[1] = norm * w_r;  /* 4 cardinal neighbours at 1  */
  w[2] = norm * w_d;  /* 4 diagonal at sqrt 2        */
  w[3] = norm * w_r2; /* 4 cardinal at 2             */
  w[4] = norm * w_l;  /* 8 knight moves at sqrt 5    */
  w[5] = norm * w_d2; /* 4 diagonal at 2 sqrt 2      */
}

/* symmetric5_row (src/gaborish.rs:137-279) for one output pixel. */
static float gab_pixel(const float *p, int w, int h, int x, int y, const float k[6]) {
  const float *r0 = p + (size_t)y * w;
  const float *rm1 = p + (size_t)gab_reflect(y - 1, h) * w;
  const float *rp1 = p + (size_t)gab_reflect(y + 1, h) * w;
  const float *rm2 = p + (size_t)gab_reflect(y - 2, h) * w;
  const float *rp2 = p + (size_t)gab_reflect(y + 2, h) * w;
  const int xm1 = gab_reflect(x - 1, w), xp1 = gab_reflect(x + 1, w);
  const int xm2 = gab_reflect(x - 2, w), xp2 = gab_reflect(x + 2, w);
  const float s_r = rm1[x] + rp1[x] + r0[xm1] + r0[xp1];
  const float s_d = rm1[xm1] + rm1[xp1] + rp1[xm1] + rp1[xp1];
  const float s_r2 = rm2[x] + rp2[x] + r0[xm2] + r0[xp2];
  const float s_d2 = rm2[xm2] + rm2[xp2] + rp2[xm2] + rp2[xp2];
  const float s_l = rm2[xm1] + rm2[xp1] + rm1[xm2] + rm1[xp2] + rp1[xm2] + rp1[xp2] + rp2[xm1] +
                    rp2[xp1];
  float acc = k[0] * r0[x];
  acc = jxo_fmla(k[1], s_r, acc);
  acc = jxo_fmla(k[2], s_d, acc);
  acc = jxo_fmla(k[3], s_r2, acc);
  acc = jxo_fmla(k[5], s_d2, acc);
  acc = jxo_fmla(k[4], s_l, acc);
  return acc;
}

void jxo_gaborish_inverse(float *xyb, int width, int height, float mul) {
  float k[6];
  jxo_gaborish_weights(mul, k);
  const size_t n = (size_t)width * height;
  float *tmp = (float *)malloc(n * sizeof(float));
  for (int c = 0; c < 3; c++) {
    float *p = xyb + c * n;
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++) tmp[(size_t)y * width + x] = gab_pixel(p, width, height, x, y, k);
    memcpy(p, tmp, n * sizeof(float));
  }
  free(tmp);
}
