/*
 * jxo_tables.c - oracle restatement of DequantMatrices::new_fast
 * (src/quant_weights.rs:2294-2337), the quant tables of the reference's
 * default effort.  TEST INFRASTRUCTURE ONLY (see jxo.h).
 *
 * Written function by function after the reference (compute_*_matrix,
 * default_dct8_override, coarse_dct8_override, scaled_override, derive_8x32,
 * tables_4x8, identity_pair, dct2_pair, compute_afv_matrix_from), with the
 * f16 wire rounding done by the F16C instructions, as the reference does on
 * x86 (f32_to_f16_bits_f16c, src/util.rs:414-424), so it shares no code with
 * the product's software conversion.
 */
#include <immintrin.h>
#include <math.h>
#include <string.h>

#include "jxo.h"
#include "jxo_internal.h"

/* ---- constants (src/quant_weights.rs) ---- */
static const float DCT8_BANDS[3][6] = {{3150.0f, 0.0f, -0.4f, -0.4f, -0.4f, -2.0f},
                                       {560.0f, 0.0f, -0.3f, -0.3f, -0.3f, -0.3f},
                                       {512.0f, -2.0f, -1.0f, 0.0f, -1.0f, -2.0f}}; /* :123 */
static const float FLAT_B8_BANDS_HQ[6] = {512.0f, -0.5f, -0.25f, 0.0f, -0.25f, -0.5f}; /* :129 */
static const float FLAT_B8_BANDS_MID[6] = {512.0f, -0.13f, -0.235f, -0.276f, -0.609f,
                                           -0.339f}; /* :130 */
static const float DCT16X16_BANDS[3][7] = {
    {8996.872571181412f, -1.3000777393353804f, -0.49424529824571225f, -0.43909377445710344f,
     -0.6350101832695744f, -0.9017726405082761f, -1.6162099239887414f},
    {3191.4836629684423f, -0.67424582104194355f, -0.807458134284710f, -0.4492583748484344f,
     -0.35865440981033403f, -0.31322389111877305f, -0.37615025315725483f},
    {1157.504081454872f, -2.0531423165804414f, -1.4f, -0.506871300333784f,
     -0.42708730624733904f, -1.4856834539296244f, -4.920914288440160f}}; /* :927 */
static const float DCT32X32_BANDS[3][8] = {
    {15718.40830982519f, -1.025f, -0.98f, -0.9012f, -0.4f, -0.48819395f, -0.421064f, -0.27f},
    {7305.763681069598f, -0.80419582f, -0.76330365f, -0.5566038f, -0.49785305f, -0.43699593f,
     -0.40180867f, -0.27321683f},
    {3803.53173721215f, -3.0607336f, -2.041327f, -2.023565f, -0.54953895f, -0.4f, -0.4f,
     -0.3f}}; /* :963 */
static const float DCT16X8_BANDS[3][7] = {
    {7240.7734393502f, -0.7f, -0.7f, -0.2f, -0.2f, -0.2f, -0.5f},
    {1448.15468787004f, -0.5f, -0.5f, -0.5f, -0.2f, -0.2f, -0.2f},
    {506.854002029f, -1.4f, -0.2f, -0.5f, -0.5f, -1.5f, -3.6f}}; /* :1006 */
static const float DCT16X32_BANDS[3][8] = {
    {13844.97076442300573f, -0.97113799999999995f, -0.658f, -0.42026f, -0.22712f, -0.2206f,
     -0.226f, -0.6f},
    {4798.964084220744293f, -0.61125308982767057f, -0.83770786552491361f, -0.79014862079498627f,
     -0.2692727459704829f, -0.38272769465388551f, -0.22924222653091453f, -0.20719098826199578f},
    {1807.236946760964614f, -1.2f, -1.2f, -0.7f, -0.7f, -0.7f, -0.4f, -0.5f}}; /* :1015 */
static const float DCT4X4_BANDS[3][4] = {{2200.0f, 0.0f, 0.0f, 0.0f},
                                         {392.0f, 0.0f, 0.0f, 0.0f},
                                         {112.0f, -0.25f, -0.25f, -0.5f}}; /* :1056 */
static const float DCT4X8_BANDS[3][4] = {
    {2198.050556016380522f, -0.96269623020744692f, -0.76194253026666783f, -0.6551140670773547f},
    {764.3655248643528689f, -0.92630200888366945f, -0.9675229603596517f, -0.27845290869168118f},
    {527.107573587542228f, -1.4594385811273854f, -1.450082094097871593f,
     -1.5843722511996204f}}; /* :1066 */
static const float DCT64X64_BANDS[3][8] = {
    {0.9f * 26629.073922049845f, -1.025f, -0.78f, -0.65012f, -0.19041574084286472f,
     -0.20819395464f, -0.421064f, -0.3273384553584867f},
    {0.9f * 9311.323871001005f, -0.3041958212306401f, -0.3633036457487539f,
     -0.35660379990111464f, -0.3443074455424403f, -0.33699592683512467f, -0.3018086652624211f,
     -0.27321683125358037f},
    {0.9f * 4992.248644553863f, -1.2f, -1.2f, -0.8f, -0.7f, -0.7f, -0.4f, -0.5f}}; /* :1087 */
static const float DCT32X64_BANDS[3][8] = {
    {0.65f * 23629.073922049845f, -1.025f, -0.78f, -0.65012f, -0.19041574084286472f,
     -0.20819395464f, -0.421064f, -0.3273384553584867f},
    {0.65f * 8611.323871001005f, -0.3041958212306401f, -0.3633036457487539f,
     -0.35660379990111464f, -0.3443074455424403f, -0.33699592683512467f, -0.3018086652624211f,
     -0.27321683125358037f},
    {0.65f * 4492.248644553863f, -1.2f, -1.2f, -0.8f, -0.7f, -0.7f, -0.4f, -0.5f}}; /* :1121 */
static const float AFV_BANDS[3][9] = {
    {3072.0f, 3072.0f, 256.0f, 256.0f, 256.0f, 414.0f, 0.0f, 0.0f, 0.0f},
    {1024.0f, 1024.0f, 50.0f, 50.0f, 50.0f, 58.0f, 0.0f, 0.0f, 0.0f},
    {384.0f, 384.0f, 12.0f, 12.0f, 12.0f, 22.0f, -0.25f, -0.25f, -0.25f}}; /* :1347 */
#define BAD ((float)0xBAD)
static const float AFV_FREQS[16] = {BAD, BAD, 0.8517778890324296f, 5.37778436506804f,
                                    BAD, BAD, 4.734747904497923f, 5.449245381693219f,
                                    1.6598270267479331f, 4.0f, 7.275749096817861f,
                                    10.423227632456525f, 2.662932286148962f,
                                    7.630657783650829f, 8.962388608184032f,
                                    12.97166202570235f}; /* :1358 */
#define AFV_FREQ_LO 0.8517778890324296f
static const float QM_4X8_SCALE[3] = {1.2f, 1.0f, 0.45f};          /* :1807 */
static const float IDENTITY_WEIGHTS[3][3] = {{280.0f, 3160.0f, 3160.0f},
                                             {60.0f, 864.0f, 864.0f},
                                             {18.0f, 200.0f, 200.0f}}; /* :1842 */
static const float DCT2_WEIGHTS[3][6] = {{3840.0f, 2560.0f, 1280.0f, 640.0f, 480.0f, 300.0f},
                                         {960.0f, 640.0f, 320.0f, 180.0f, 140.0f, 120.0f},
                                         {640.0f, 320.0f, 128.0f, 64.0f, 32.0f, 16.0f}}; /* :1848 */

/* The static library tables the reference uses verbatim (:246, :455); the
 * tests pass them in from the reference's source text
 * (tests/golden/reference_tables.json) through jxo_set_static_tables. */
static float g_static_8x8[3][64];
static float g_static_16x8[3][128];
static int g_static_set = 0;

void jxo_set_static_tables(const float *m8x8, const float *m16x8) {
  memcpy(g_static_8x8, m8x8, sizeof(g_static_8x8));
  memcpy(g_static_16x8, m16x8, sizeof(g_static_16x8));
  g_static_set = 1;
}

/* ---- f16 wire rounding (F16C, round to nearest even) ---- */
__attribute__((target("f16c"))) static float f16_roundtrip(float v) {
  const __m128i h = _mm_cvtps_ph(_mm_set_ss(v), _MM_FROUND_TO_NEAREST_INT);
  return _mm_cvtss_f32(_mm_cvtph_ps(h));
}

typedef struct {
  int num_bands;
  float bands[3][16];
} band_override;

static float band_mult(float v) { return v > 0.0f ? 1.0f + v : 1.0f / (1.0f - v); } /* :1002 */

static float interpolate_vec_bands(float scaled_pos, const float *bands) { /* :1372 */
  const float idx_f = floorf(scaled_pos);
  const float frac = scaled_pos - idx_f;
  const size_t idx = (size_t)idx_f;
  const float a = bands[idx];
  const float b = bands[idx + 1];
  return powf(b / a, frac) * a;
}

/* default_dct8_override / coarse_dct8_override (:140-200) */
static band_override coarse_dct8_override(void) {
  const float hf_step = powf(1.25f, 1.0f / 5.0f);
  band_override out;
  memset(&out, 0, sizeof(out));
  out.num_bands = 6;
  for (int c = 0; c < 3; c++) memcpy(out.bands[c], DCT8_BANDS[c], 6 * sizeof(float));
  for (int i = 1; i < 6; i++) {
    const float ratio = band_mult(DCT8_BANDS[1][i]) * hf_step;
    out.bands[1][i] = ratio > 1.0f ? ratio - 1.0f : 1.0f - 1.0f / ratio;
  }
  for (int c = 0; c < 3; c++)
    for (int i = 0; i < 6; i++) {
      const float v = out.bands[c][i];
      out.bands[c][i] = i == 0 ? f16_roundtrip(v / 64.0f) * 64.0f : f16_roundtrip(v);
    }
  return out;
}

static band_override default_dct8_override(int use_coarse, const float flat_b8[6]) {
  band_override out;
  if (use_coarse) {
    out = coarse_dct8_override();
  } else {
    memset(&out, 0, sizeof(out));
    out.num_bands = 6;
    for (int c = 0; c < 3; c++) memcpy(out.bands[c], DCT8_BANDS[c], 6 * sizeof(float));
  }
  memcpy(out.bands[2], flat_b8, 6 * sizeof(float));
  for (int c = 0; c < 3; c++)
    for (int i = 0; i < 6; i++) {
      const float v = out.bands[c][i];
      out.bands[c][i] = i == 0 ? f16_roundtrip(v / 64.0f) * 64.0f : f16_roundtrip(v);
    }
  return out;
}

/* scaled_override (:202-224) */
static band_override scaled_override(const float *defaults, int n, float scale) {
  band_override out;
  memset(&out, 0, sizeof(out));
  out.num_bands = n;
  for (int c = 0; c < 3; c++)
    for (int i = 0; i < n; i++) {
      const float v = i == 0 ? defaults[c * n] * scale : defaults[c * n + i];
      out.bands[c][i] = i == 0 ? f16_roundtrip(v / 64.0f) * 64.0f : f16_roundtrip(v);
    }
  return out;
}

/* compute_dct8x8_matrix, compute_dct16x16_matrix, compute_dct32x32_matrix,
 * compute_dct16x8_matrix, compute_dct32x16_matrix, compute_dct64x64_matrix,
 * compute_dct64x32_matrix (:1281-1712): one rows x cols radial table. */
static void compute_matrix(const float *src, int num_bands, int rows, int cols, float *out) {
  float bands[16];
  bands[0] = src[0];
  for (int i = 1; i < num_bands; i++) bands[i] = bands[i - 1] * band_mult(src[i]);
  const float scale = ((float)num_bands - 1.0f) / ((float)M_SQRT2 + 1e-6f);
  const float rcprow = scale / (float)(rows - 1);
  const float rcpcol = scale / (float)(cols - 1);
  for (int y = 0; y < rows; y++) {
    const float dy = (float)y * rcprow;
    const float dy2 = dy * dy;
    for (int x = 0; x < cols; x++) {
      const float dx = (float)x * rcpcol;
      const float dist = sqrtf(jxo_fmla(dx, dx, dy2));
      out[y * cols + x] = 1.0f / interpolate_vec_bands(dist, bands);
    }
  }
}

/* fill_ac_reciprocals (:3136) */
static void fill_ac_reciprocals(const float *matrix, float *inv, int n) {
  inv[0] = 0.0f;
  for (int i = 1; i < n; i++) inv[i] = 1.0f / matrix[i];
}

static float *kind_matrix(float *blob, int kind, int c) {
  return blob + jxo_qt_offset(kind) + c * jxo_qt_size(kind);
}
static float *kind_inverse(float *blob, int kind, int c) {
  return blob + jxo_qt_offset(kind) + (3 + c) * jxo_qt_size(kind);
}

static void put_override(float *blob, int kind, int rows, int cols, const band_override *o) {
  for (int c = 0; c < 3; c++) {
    compute_matrix(o->bands[c], o->num_bands, rows, cols, kind_matrix(blob, kind, c));
    fill_ac_reciprocals(kind_matrix(blob, kind, c), kind_inverse(blob, kind, c), rows * cols);
  }
}

static void put_library(float *blob, int kind, int rows, int cols, const float *rows_src,
                        int num_bands) {
  for (int c = 0; c < 3; c++) {
    compute_matrix(rows_src + c * num_bands, num_bands, rows, cols, kind_matrix(blob, kind, c));
    fill_ac_reciprocals(kind_matrix(blob, kind, c), kind_inverse(blob, kind, c), rows * cols);
  }
}

/* derive_8x32 (:1729-1789) */
static void derive_8x32(const float pair[3][7], float *blob) {
  for (int c = 0; c < 3; c++) {
    float pair_curve[7];
    pair_curve[0] = pair[c][0];
    for (int i = 1; i < 7; i++) pair_curve[i] = pair_curve[i - 1] * band_mult(pair[c][i]);
    float nodes[8];
    for (int i = 0; i < 8; i++) {
      const float pos = (float)i * 6.0f / 7.0f;
      const float sample = pos >= 6.0f ? pair_curve[6] : interpolate_vec_bands(pos, pair_curve);
      nodes[i] = sample * (float)M_SQRT2;
    }
    float params[8];
    params[0] = f16_roundtrip(nodes[0] / 64.0f) * 64.0f;
    for (int i = 1; i < 8; i++) {
      const float r = nodes[i] / nodes[i - 1];
      params[i] = f16_roundtrip(r >= 1.0f ? r - 1.0f : 1.0f - 1.0f / r);
    }
    float bands[8];
    bands[0] = params[0];
    for (int i = 1; i < 8; i++) bands[i] = bands[i - 1] * band_mult(params[i]);
    const float scale = 7.0f / ((float)M_SQRT2 + 1e-6f);
    const float rcprow = scale / 7.0f, rcpcol = scale / 31.0f;
    float *m = kind_matrix(blob, JXG_QT_DCT8X32, c);
    for (int y = 0; y < 8; y++) {
      const float dy = (float)y * rcprow;
      for (int x = 0; x < 32; x++) {
        const float dx = (float)x * rcpcol;
        m[y * 32 + x] = 1.0f / interpolate_vec_bands(sqrtf(jxo_fmla(dx, dx, dy * dy)), bands);
      }
    }
    fill_ac_reciprocals(m, kind_inverse(blob, JXG_QT_DCT8X32, c), 256);
  }
}

/* compute_dct4x8_matrix (:1386-1424) from 4-band parameters */
static void compute_dct4x8_matrix(const float src[3][4], float *blob) {
  for (int c = 0; c < 3; c++) {
    float bands[4];
    bands[0] = src[c][0];
    for (int i = 1; i < 4; i++) bands[i] = bands[i - 1] * band_mult(src[c][i]);
    const float scale = 3.0f / ((float)M_SQRT2 + 1e-6f);
    const float rcprow = scale / 3.0f, rcpcol = scale / 7.0f;
    float w4x8[32];
    for (int y = 0; y < 4; y++) {
      const float dy = (float)y * rcprow;
      const float dy2 = dy * dy;
      for (int x = 0; x < 8; x++) {
        const float dx = (float)x * rcpcol;
        w4x8[y * 8 + x] = interpolate_vec_bands(sqrtf(jxo_fmla(dx, dx, dy2)), bands);
      }
    }
    float *m = kind_matrix(blob, JXG_QT_DCT4X8, c);
    for (int y = 0; y < 8; y++)
      for (int x = 0; x < 8; x++) m[y * 8 + x] = 1.0f / w4x8[(y / 2) * 8 + x];
    fill_ac_reciprocals(m, kind_inverse(blob, JXG_QT_DCT4X8, c), 64);
  }
}

/* compute_dct4x4_matrix_from (:1430-1468) */
static void compute_dct4x4_matrix(float *blob) {
  for (int c = 0; c < 3; c++) {
    float bands[4];
    bands[0] = DCT4X4_BANDS[c][0];
    for (int i = 1; i < 4; i++) bands[i] = bands[i - 1] * band_mult(DCT4X4_BANDS[c][i]);
    const float scale = 3.0f / ((float)M_SQRT2 + 1e-6f);
    const float rcp = scale / 3.0f;
    float w4x4[16];
    for (int y = 0; y < 4; y++) {
      const float dy = (float)y * rcp;
      const float dy2 = dy * dy;
      for (int x = 0; x < 4; x++) {
        const float dx = (float)x * rcp;
        w4x4[y * 4 + x] = interpolate_vec_bands(sqrtf(jxo_fmla(dx, dx, dy2)), bands);
      }
    }
    float *m = kind_matrix(blob, JXG_QT_DCT4X4, c);
    for (int y = 0; y < 8; y++)
      for (int x = 0; x < 8; x++) m[y * 8 + x] = 1.0f / w4x4[(y / 2) * 4 + (x / 2)];
    fill_ac_reciprocals(m, kind_inverse(blob, JXG_QT_DCT4X4, c), 64);
  }
}

/* compute_afv_matrix_from(AFV_BANDS, DCT4X8_BANDS, DCT4X4_BANDS) (:1474-1562) */
static void compute_afv_matrix(float *blob) {
  const float afv_freq_hi = 12.97166202570235f - AFV_FREQ_LO + 1e-6f;
  for (int c = 0; c < 3; c++) {
    float bands[4];
    bands[0] = DCT4X8_BANDS[c][0];
    for (int i = 1; i < 4; i++) bands[i] = bands[i - 1] * band_mult(DCT4X8_BANDS[c][i]);
    const float scale = 3.0f / ((float)M_SQRT2 + 1e-6f);
    float w4x8[32];
    for (int y = 0; y < 4; y++) {
      const float dy = (float)y * (scale / 3.0f);
      const float dy2 = dy * dy;
      for (int x = 0; x < 8; x++) {
        const float dx = (float)x * (scale / 7.0f);
        w4x8[y * 8 + x] = interpolate_vec_bands(sqrtf(jxo_fmla(dx, dx, dy2)), bands);
      }
    }
    bands[0] = DCT4X4_BANDS[c][0];
    for (int i = 1; i < 4; i++) bands[i] = bands[i - 1] * band_mult(DCT4X4_BANDS[c][i]);
    float w4x4[16];
    for (int y = 0; y < 4; y++) {
      const float dy = (float)y * (scale / 3.0f);
      const float dy2 = dy * dy;
      for (int x = 0; x < 4; x++) {
        const float dx = (float)x * (scale / 3.0f);
        w4x4[y * 4 + x] = interpolate_vec_bands(sqrtf(jxo_fmla(dx, dx, dy2)), bands);
      }
    }
    bands[0] = AFV_BANDS[c][5];
    for (int i = 1; i < 4; i++) bands[i] = bands[i - 1] * band_mult(AFV_BANDS[c][i + 5]);
    float weights[64] = {0};
    weights[0] = 1.0f;
    weights[8] = AFV_BANDS[c][0];
    weights[1] = AFV_BANDS[c][1];
    weights[16] = AFV_BANDS[c][2];
    weights[2] = AFV_BANDS[c][3];
    weights[18] = AFV_BANDS[c][4];
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        if (x < 2 && y < 2) continue;
        const float pos = (AFV_FREQS[y * 4 + x] - AFV_FREQ_LO) * 3.0f / afv_freq_hi;
        weights[(2 * y) * 8 + 2 * x] = interpolate_vec_bands(pos, bands);
      }
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 8; x++) {
        if (x == 0 && y == 0) continue;
        weights[(2 * y + 1) * 8 + x] = w4x8[y * 8 + x];
      }
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        if (x == 0 && y == 0) continue;
        weights[(2 * y) * 8 + 2 * x + 1] = w4x4[y * 4 + x];
      }
    float *m = kind_matrix(blob, JXG_QT_AFV, c);
    for (int k = 0; k < 64; k++) m[k] = 1.0f / weights[k];
    fill_ac_reciprocals(m, kind_inverse(blob, JXG_QT_AFV, c), 64);
  }
}

/* identity_pair (:1906-1922) and dct2_pair (:1924-1955): the inverse holds
 * the weight itself, not the reciprocal of the step. */
static void fine_pairs(float *blob) {
  for (int c = 0; c < 3; c++) {
    float w[64];
    for (int k = 0; k < 64; k++) w[k] = IDENTITY_WEIGHTS[c][0];
    w[1] = IDENTITY_WEIGHTS[c][1];
    w[8] = IDENTITY_WEIGHTS[c][1];
    w[9] = IDENTITY_WEIGHTS[c][2];
    for (int k = 0; k < 64; k++) {
      kind_matrix(blob, JXG_QT_IDENTITY, c)[k] = 1.0f / w[k];
      kind_inverse(blob, JXG_QT_IDENTITY, c)[k] = k != 0 ? w[k] : 0.0f;
    }
    const float *dw = DCT2_WEIGHTS[c];
    for (int k = 0; k < 64; k++) w[k] = BAD;
    w[1] = dw[0];
    w[8] = dw[0];
    w[9] = dw[1];
    for (int y = 0; y < 2; y++)
      for (int x = 0; x < 2; x++) {
        w[y * 8 + x + 2] = dw[2];
        w[(y + 2) * 8 + x] = dw[2];
        w[(y + 2) * 8 + x + 2] = dw[3];
      }
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        w[y * 8 + x + 4] = dw[4];
        w[(y + 4) * 8 + x] = dw[4];
        w[(y + 4) * 8 + x + 4] = dw[5];
      }
    for (int k = 0; k < 64; k++) {
      kind_matrix(blob, JXG_QT_DCT2X2, c)[k] = 1.0f / w[k];
      kind_inverse(blob, JXG_QT_DCT2X2, c)[k] = k != 0 ? w[k] : 0.0f;
    }
  }
}

int jxo_fast_tables(float distance, int arbiter, float *blob) {
  if (!g_static_set) return -1;
  memset(blob, 0, JXG_QT_BLOB_FLOATS * sizeof(float));
  const float ss2_from = arbiter == 1 ? 1.75f : 2.25f; /* Tuning::qm_ss2_distance */
  int use_ss2 = 0, coarse = 0;
  const float *flat = NULL;
  if (distance < 0.3f) {
    /* new(distance) -> new_untilted's HQ set: spec tables */
  } else if (distance >= ss2_from) {
    use_ss2 = 1;
    coarse = distance >= 3.5f;
    flat = FLAT_B8_BANDS_MID;
  } else {
    flat = distance >= 1.25f ? FLAT_B8_BANDS_MID : FLAT_B8_BANDS_HQ;
  }
  /* DCT8: the signalled override, or the stored library table */
  if (flat) {
    const band_override o8 = default_dct8_override(coarse, flat);
    put_override(blob, JXG_QT_DCT8, 8, 8, &o8);
  } else {
    for (int c = 0; c < 3; c++) {
      memcpy(kind_matrix(blob, JXG_QT_DCT8, c), g_static_8x8[c], 64 * sizeof(float));
      fill_ac_reciprocals(g_static_8x8[c], kind_inverse(blob, JXG_QT_DCT8, c), 64);
    }
  }
  /* large_tables (:2113-2192) */
  if (use_ss2) {
    const band_override o16 = scaled_override(&DCT16X16_BANDS[0][0], 7, 0.78f);
    const band_override o32 = scaled_override(&DCT32X32_BANDS[0][0], 8, 0.89f);
    const band_override o3216 = scaled_override(&DCT16X32_BANDS[0][0], 8, 1.20f);
    put_override(blob, JXG_QT_DCT16X16, 16, 16, &o16);
    put_override(blob, JXG_QT_DCT32X32, 32, 32, &o32);
    put_override(blob, JXG_QT_DCT32X16, 16, 32, &o3216);
  } else {
    put_library(blob, JXG_QT_DCT16X16, 16, 16, &DCT16X16_BANDS[0][0], 7);
    put_library(blob, JXG_QT_DCT32X32, 32, 32, &DCT32X32_BANDS[0][0], 8);
    put_library(blob, JXG_QT_DCT32X16, 16, 32, &DCT16X32_BANDS[0][0], 8);
  }
  /* shared_tables (:1986-2064), tables_4x8 (:1810-1836), derive_8x32 */
  for (int c = 0; c < 3; c++) {
    memcpy(kind_matrix(blob, JXG_QT_DCT16X8, c), g_static_16x8[c], 128 * sizeof(float));
    fill_ac_reciprocals(g_static_16x8[c], kind_inverse(blob, JXG_QT_DCT16X8, c), 128);
  }
  derive_8x32(DCT16X8_BANDS, blob);
  put_library(blob, JXG_QT_DCT64X64, 64, 64, &DCT64X64_BANDS[0][0], 8);
  put_library(blob, JXG_QT_DCT64X32, 32, 64, &DCT32X64_BANDS[0][0], 8);
  fine_pairs(blob);
  compute_dct4x4_matrix(blob);
  {
    float o[3][4];
    for (int c = 0; c < 3; c++) {
      o[c][0] = f16_roundtrip(DCT4X8_BANDS[c][0] * QM_4X8_SCALE[c] / 64.0f) * 64.0f;
      for (int i = 1; i < 4; i++) o[c][i] = f16_roundtrip(DCT4X8_BANDS[c][i]);
    }
    compute_dct4x8_matrix((const float(*)[4])o, blob);
  }
  compute_afv_matrix(blob);
  return 0;
}
