// jxg_tables.cpp - the dequantization tables the reference's default effort
// codes with: DequantMatrices::new_fast(distance, arbiter)
// (src/quant_weights.rs:2294-2337), built on the host in the reference's
// float operation order and laid out as the quant-matrix blob of jxg.h.
//
// Compiled with -ffp-contract=off: every a*b+c stays unfused unless written
// through fmla(), whose flavour (`fused`) is the reference's build-dependent
// fmla (src/dct.rs:33-53).
//
// Tiers of new_fast (qm_ss2_distance is 2.25 for the SSIMULACRA2 arbiter,
// 1.75 for CVVDP, src/arbiter.rs:50-62):
//   0  d < 0.3            the spec tables (new -> new_untilted's HQ set)
//   1  0.3 <= d < 1.25    DCT8 with the FLAT_B8_BANDS_HQ B row
//   2  1.25 <= d < ss2    DCT8 with the FLAT_B8_BANDS_MID B row
//   3  ss2 <= d < 3.5     MID DCT8 row; DCT16 x0.78, DCT32 x0.89, 32x16 x1.20
//   4  d >= max(ss2, 3.5) as 3 with the coarse DCT8 luma tilt
// Every tier shares the spec IDENTITY / DCT2X2 / DCT4X4 / AFV / DCT64
// tables, the stored DCT16X8 table, the refitted DCT4X8 table and the
// DCT8X32 table derived from the DCT16X8 bands (src/quant_weights.rs:1986-2050).
#include <math.h>
#include <string.h>

#include <cmath>

#include "jxg_internal.h"

namespace {

inline float fmla(bool fused, float a, float b, float c) {
  return fused ? std::fma(a, b, c) : a * b + c;
}

inline float bits_to_f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// f16 round trip of a relative band parameter, and of an absolute weight as
// the wire carries it (w / 64 in f16).
inline float f16r(float v) { return jxg_f16_to_f32(jxg_f32_to_f16(v)); }
inline float wire_abs(float v) { return f16r(v / 64.0f) * 64.0f; }

const uint32_t kSizes[JXG_QT_NUM_KINDS] = {64, 128, 256, 1024, 512, 256, 4096,
                                           2048, 64, 64, 64, 64, 64};

// ---- band parameters (src/quant_weights.rs:123-138,927-1230,1347-1366) ----
// JPEG XL library rows (X, Y, B): the first entry is the weight at radius 0,
// the rest the multiplicative steps between radial bands.
const float kDct8Bands[3][6] = {{3150.0f, 0.0f, -0.4f, -0.4f, -0.4f, -2.0f},
                                {560.0f, 0.0f, -0.3f, -0.3f, -0.3f, -0.3f},
                                {512.0f, -2.0f, -1.0f, 0.0f, -1.0f, -2.0f}};
const float kFlatB8Hq[6] = {512.0f, -0.5f, -0.25f, 0.0f, -0.25f, -0.5f};
const float kFlatB8Mid[6] = {512.0f, -0.13f, -0.235f, -0.276f, -0.609f, -0.339f};
const float kDct16x16Bands[3][7] = {
    {8996.872571181412f, -1.3000777393353804f, -0.49424529824571225f, -0.43909377445710344f,
     -0.6350101832695744f, -0.9017726405082761f, -1.6162099239887414f},
    {3191.4836629684423f, -0.67424582104194355f, -0.80745813428471f, -0.4492583748484344f,
     -0.35865440981033403f, -0.31322389111877305f, -0.37615025315725483f},
    {1157.504081454872f, -2.0531423165804414f, -1.4f, -0.506871300333784f,
     -0.42708730624733904f, -1.4856834539296244f, -4.92091428844016f}};
const float kDct32x32Bands[3][8] = {
    {15718.40830982519f, -1.025f, -0.98f, -0.9012f, -0.4f, -0.48819395f, -0.421064f, -0.27f},
    {7305.763681069598f, -0.80419582f, -0.76330365f, -0.5566038f, -0.49785305f, -0.43699593f,
     -0.40180867f, -0.27321683f},
    {3803.53173721215f, -3.0607336f, -2.041327f, -2.023565f, -0.54953895f, -0.4f, -0.4f, -0.3f}};
const float kDct16x8Bands[3][7] = {{7240.7734393502f, -0.7f, -0.7f, -0.2f, -0.2f, -0.2f, -0.5f},
                                   {1448.15468787004f, -0.5f, -0.5f, -0.5f, -0.2f, -0.2f, -0.2f},
                                   {506.854002029f, -1.4f, -0.2f, -0.5f, -0.5f, -1.5f, -3.6f}};
const float kDct16x32Bands[3][8] = {
    {13844.97076442300573f, -0.97113799999999995f, -0.658f, -0.42026f, -0.22712f, -0.2206f,
     -0.226f, -0.6f},
    {4798.964084220744293f, -0.61125308982767057f, -0.83770786552491361f, -0.79014862079498627f,
     -0.2692727459704829f, -0.38272769465388551f, -0.22924222653091453f, -0.20719098826199578f},
    {1807.236946760964614f, -1.2f, -1.2f, -0.7f, -0.7f, -0.7f, -0.4f, -0.5f}};
const float kDct64x64Bands[3][8] = {
    {0.9f * 26629.073922049845f, -1.025f, -0.78f, -0.65012f, -0.19041574084286472f,
     -0.20819395464f, -0.421064f, -0.3273384553584867f},
    {0.9f * 9311.323871001005f, -0.3041958212306401f, -0.3633036457487539f,
     -0.35660379990111464f, -0.3443074455424403f, -0.33699592683512467f, -0.3018086652624211f,
     -0.27321683125358037f},
    {0.9f * 4992.248644553863f, -1.2f, -1.2f, -0.8f, -0.7f, -0.7f, -0.4f, -0.5f}};
const float kDct32x64Bands[3][8] = {
    {0.65f * 23629.073922049845f, -1.025f, -0.78f, -0.65012f, -0.19041574084286472f,
     -0.20819395464f, -0.421064f, -0.3273384553584867f},
    {0.65f * 8611.323871001005f, -0.3041958212306401f, -0.3633036457487539f,
     -0.35660379990111464f, -0.3443074455424403f, -0.33699592683512467f, -0.3018086652624211f,
     -0.27321683125358037f},
    {0.65f * 4492.248644553863f, -1.2f, -1.2f, -0.8f, -0.7f, -0.7f, -0.4f, -0.5f}};
const float kDct4x4Bands[3][4] = {
    {2200.0f, 0.0f, 0.0f, 0.0f}, {392.0f, 0.0f, 0.0f, 0.0f}, {112.0f, -0.25f, -0.25f, -0.5f}};
const float kDct4x8Bands[3][4] = {
    {2198.050556016380522f, -0.96269623020744692f, -0.76194253026666783f, -0.6551140670773547f},
    {764.3655248643528689f, -0.92630200888366945f, -0.9675229603596517f, -0.27845290869168118f},
    {527.107573587542228f, -1.4594385811273854f, -1.450082094097871593f, -1.5843722511996204f}};
// DCT4X8 chroma refit of XYB frames (QM_4X8_SCALE, src/quant_weights.rs:1807).
const float kQm4x8Scale[3] = {1.2f, 1.0f, 0.45f};
const float kAfvBands[3][9] = {{3072.0f, 3072.0f, 256.0f, 256.0f, 256.0f, 414.0f, 0.0f, 0.0f, 0.0f},
                               {1024.0f, 1024.0f, 50.0f, 50.0f, 50.0f, 58.0f, 0.0f, 0.0f, 0.0f},
                               {384.0f, 384.0f, 12.0f, 12.0f, 12.0f, 22.0f, -0.25f, -0.25f, -0.25f}};
const float kBad = 2989.0f;  // libjxl's 0xBAD sentinel of the unused AFV / DCT2X2 slots
const float kAfvFreqs[16] = {kBad, kBad, 0.8517778890324296f, 5.37778436506804f,
                             kBad, kBad, 4.734747904497923f, 5.449245381693219f,
                             1.6598270267479331f, 4.0f, 7.275749096817861f, 10.423227632456525f,
                             2.662932286148962f, 7.630657783650829f, 8.962388608184032f,
                             12.97166202570235f};
const float kAfvFreqLo = 0.8517778890324296f;
const float kAfvFreqHi = 12.97166202570235f - kAfvFreqLo + 1e-6f;
const float kIdentityWeights[3][3] = {{280.0f, 3160.0f, 3160.0f}, {60.0f, 864.0f, 864.0f},
                                      {18.0f, 200.0f, 200.0f}};
const float kDct2Weights[3][6] = {{3840.0f, 2560.0f, 1280.0f, 640.0f, 480.0f, 300.0f},
                                  {960.0f, 640.0f, 320.0f, 180.0f, 140.0f, 120.0f},
                                  {640.0f, 320.0f, 128.0f, 64.0f, 32.0f, 16.0f}};

// The stored DCT8 and DCT16X8 step tables of the JPEG XL library
// (DEQUANT_MATRIX_8X8 / DEQUANT_MATRIX_16X8, src/quant_weights.rs:246,455), as
// f32 bit patterns.  The reference uses them instead of the band formula,
// from which they differ by a few ulp.
#include "jxg_static_tables.inc"

const float kSsScale16 = 0.78f, kSsScale32 = 0.89f, kSsScale16x32 = 1.20f;
const float kDct8YHfScale = 1.25f;
const float kFlatB8MinDistance = 0.3f, kFlatB8MidMinDistance = 1.25f;
const float kOriginalCoarseDct8Distance = 3.5f;

float band_mult(float v) { return v > 0.0f ? 1.0f + v : 1.0f / (1.0f - v); }

float interpolate_bands(float pos, const float *bands) {
  float idx_f = floorf(pos);
  float frac = pos - idx_f;
  int idx = (int)idx_f;
  float a = bands[idx], b = bands[idx + 1];
  return powf(b / a, frac) * a;
}

// A signalled band override (BandOverride, src/quant_weights.rs:97-101).
struct Bands {
  int nb;
  float v[3][16];
};

// default_dct8_override / coarse_dct8_override (src/quant_weights.rs:140-200):
// the spec rows (coarse: the luma steps tilted by 1.25^(1/5) each), the B row
// replaced by `flat_b8`, every parameter f16-rounded as signalled.
Bands dct8_override(bool coarse, const float flat_b8[6]) {
  Bands o;
  memset(&o, 0, sizeof(o));
  o.nb = 6;
  for (int c = 0; c < 3; c++)
    for (int i = 0; i < 6; i++) o.v[c][i] = kDct8Bands[c][i];
  auto round_all = [&]() {
    for (int c = 0; c < 3; c++)
      for (int i = 0; i < 6; i++) o.v[c][i] = i == 0 ? wire_abs(o.v[c][i]) : f16r(o.v[c][i]);
  };
  if (coarse) {
    const float hf_step = powf(kDct8YHfScale, 1.0f / 5.0f);
    for (int i = 1; i < 6; i++) {
      const float ratio = band_mult(kDct8Bands[1][i]) * hf_step;
      o.v[1][i] = ratio > 1.0f ? ratio - 1.0f : 1.0f - 1.0f / ratio;
    }
    round_all();
  }
  for (int i = 0; i < 6; i++) o.v[2][i] = flat_b8[i];
  round_all();
  return o;
}

// scaled_override (src/quant_weights.rs:202-224): the base weight scaled,
// every parameter f16-rounded.
Bands scaled(const float *rows, int nb, float scale) {
  Bands o;
  memset(&o, 0, sizeof(o));
  o.nb = nb;
  for (int c = 0; c < 3; c++)
    for (int i = 0; i < nb; i++)
      o.v[c][i] = i == 0 ? wire_abs(rows[c * nb] * scale) : f16r(rows[c * nb + i]);
  return o;
}

Bands plain(const float *rows, int nb) {
  Bands o;
  memset(&o, 0, sizeof(o));
  o.nb = nb;
  for (int c = 0; c < 3; c++)
    for (int i = 0; i < nb; i++) o.v[c][i] = rows[c * nb + i];
  return o;
}

// GetQuantWeights-style radial table of one channel: rows x cols step sizes
// (the compute_dct*_matrix family, src/quant_weights.rs:1281-1712).
void band_matrix(bool fused, const float *src, int nb, int rows, int cols, float *out) {
  float bands[16];
  bands[0] = src[0];
  for (int i = 1; i < nb; i++) bands[i] = bands[i - 1] * band_mult(src[i]);
  const float scale = ((float)nb - 1.0f) / (1.41421356237309504880f + 1e-6f);
  const float rcprow = scale / (float)(rows - 1);
  const float rcpcol = scale / (float)(cols - 1);
  for (int y = 0; y < rows; y++) {
    const float dy = (float)y * rcprow;
    const float dy2 = dy * dy;
    for (int x = 0; x < cols; x++) {
      const float dx = (float)x * rcpcol;
      out[y * cols + x] = 1.0f / interpolate_bands(sqrtf(fmla(fused, dx, dx, dy2)), bands);
    }
  }
}

// Radial weights (not steps) of a 4 x cols grid, as compute_dct4x8_matrix and
// compute_afv_matrix_from build them (src/quant_weights.rs:1386-1545).
void weights4(bool fused, const float *src, int cols, float *w) {
  float bands[4];
  bands[0] = src[0];
  for (int i = 1; i < 4; i++) bands[i] = bands[i - 1] * band_mult(src[i]);
  const float scale = 3.0f / (1.41421356237309504880f + 1e-6f);
  const float rcprow = scale / 3.0f, rcpcol = scale / (float)(cols - 1);
  for (int y = 0; y < 4; y++) {
    const float dy = (float)y * rcprow;
    const float dy2 = dy * dy;
    for (int x = 0; x < cols; x++) {
      const float dx = (float)x * rcpcol;
      w[y * cols + x] = interpolate_bands(sqrtf(fmla(fused, dx, dx, dy2)), bands);
    }
  }
}

// fill_ac_reciprocals (src/quant_weights.rs:3136): inv[0] = 0, inv[k] = 1/m[k].
void reciprocals(const float *m, float *inv, uint32_t n) {
  inv[0] = 0.0f;
  for (uint32_t i = 1; i < n; i++) inv[i] = 1.0f / m[i];
}

uint32_t offset_of(int kind) {
  uint32_t off = 0;
  for (int k = 0; k < kind; k++) off += 6 * kSizes[k];
  return off;
}

float *mat(float *blob, int kind, int c) { return blob + offset_of(kind) + c * kSizes[kind]; }
float *inv(float *blob, int kind, int c) { return blob + offset_of(kind) + (3 + c) * kSizes[kind]; }

void put_bands(bool fused, float *blob, int kind, int rows, int cols, const Bands &b) {
  const uint32_t n = kSizes[kind];
  for (int c = 0; c < 3; c++) {
    band_matrix(fused, b.v[c], b.nb, rows, cols, mat(blob, kind, c));
    reciprocals(mat(blob, kind, c), inv(blob, kind, c), n);
  }
}

void put_static(float *blob, int kind, const uint32_t *bits, uint32_t n) {
  for (int c = 0; c < 3; c++) {
    float *m = mat(blob, kind, c);
    for (uint32_t i = 0; i < n; i++) m[i] = bits_to_f(bits[c * n + i]);
    reciprocals(m, inv(blob, kind, c), n);
  }
}

// derive_8x32 (src/quant_weights.rs:1729-1789): the DCT16X8 pair's band curve
// sampled at the 8-band grid's normalised frequencies, x sqrt(2), signalled
// (f16) and decoded again into the 8 x 32 table.
void put_8x32(bool fused, float *blob, const float pair[3][7]) {
  const float sqrt2 = 1.41421356237309504880f;
  for (int c = 0; c < 3; c++) {
    float curve[7];
    curve[0] = pair[c][0];
    for (int i = 1; i < 7; i++) curve[i] = curve[i - 1] * band_mult(pair[c][i]);
    float nodes[8], params[8], bands[8];
    for (int i = 0; i < 8; i++) {
      const float pos = (float)i * 6.0f / 7.0f;
      const float s = pos >= 6.0f ? curve[6] : interpolate_bands(pos, curve);
      nodes[i] = s * sqrt2;
    }
    params[0] = f16r(nodes[0] / 64.0f) * 64.0f;
    for (int i = 1; i < 8; i++) {
      const float r = nodes[i] / nodes[i - 1];
      params[i] = f16r(r >= 1.0f ? r - 1.0f : 1.0f - 1.0f / r);
    }
    bands[0] = params[0];
    for (int i = 1; i < 8; i++) bands[i] = bands[i - 1] * band_mult(params[i]);
    const float scale = 7.0f / (sqrt2 + 1e-6f);
    const float rcprow = scale / 7.0f, rcpcol = scale / 31.0f;
    float *m = mat(blob, JXG_QT_DCT8X32, c);
    for (int y = 0; y < 8; y++) {
      const float dy = (float)y * rcprow;
      for (int x = 0; x < 32; x++) {
        const float dx = (float)x * rcpcol;
        m[y * 32 + x] = 1.0f / interpolate_bands(sqrtf(fmla(fused, dx, dx, dy * dy)), bands);
      }
    }
    reciprocals(m, inv(blob, JXG_QT_DCT8X32, c), 256);
  }
}

// The tables shared by every tier (shared_tables / tables_4x8 /
// build_fine_matrices, src/quant_weights.rs:1810-1836,1897-2064).
void put_shared(bool fused, float *blob) {
  put_static(blob, JXG_QT_DCT16X8, &kStatic16x8[0][0], 128);
  put_8x32(fused, blob, kDct16x8Bands);
  put_bands(fused, blob, JXG_QT_DCT64X64, 64, 64, plain(&kDct64x64Bands[0][0], 8));
  put_bands(fused, blob, JXG_QT_DCT64X32, 32, 64, plain(&kDct32x64Bands[0][0], 8));
  for (int c = 0; c < 3; c++) {
    // IDENTITY: step 1/w, inverse w itself (identity_pair, :1906-1922)
    float w[64];
    for (int k = 0; k < 64; k++) w[k] = kIdentityWeights[c][0];
    w[1] = kIdentityWeights[c][1];
    w[8] = kIdentityWeights[c][1];
    w[9] = kIdentityWeights[c][2];
    float *m = mat(blob, JXG_QT_IDENTITY, c), *iv = inv(blob, JXG_QT_IDENTITY, c);
    for (int k = 0; k < 64; k++) {
      m[k] = 1.0f / w[k];
      iv[k] = k == 0 ? 0.0f : w[k];
    }
    // DCT2X2: the recursive band layout (dct2_pair, :1924-1955)
    const float *d2 = kDct2Weights[c];
    for (int k = 0; k < 64; k++) w[k] = kBad;
    w[1] = d2[0];
    w[8] = d2[0];
    w[9] = d2[1];
    for (int y = 0; y < 2; y++)
      for (int x = 0; x < 2; x++) {
        w[y * 8 + x + 2] = d2[2];
        w[(y + 2) * 8 + x] = d2[2];
        w[(y + 2) * 8 + x + 2] = d2[3];
      }
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        w[y * 8 + x + 4] = d2[4];
        w[(y + 4) * 8 + x] = d2[4];
        w[(y + 4) * 8 + x + 4] = d2[5];
      }
    m = mat(blob, JXG_QT_DCT2X2, c);
    iv = inv(blob, JXG_QT_DCT2X2, c);
    for (int k = 0; k < 64; k++) {
      m[k] = 1.0f / w[k];
      iv[k] = k == 0 ? 0.0f : w[k];
    }
    // DCT4X4: 4x4 radial weights, each on a 2x2 cell (:1426-1468)
    float w44[16], w48[32];
    weights4(fused, kDct4x4Bands[c], 4, w44);
    m = mat(blob, JXG_QT_DCT4X4, c);
    for (int y = 0; y < 8; y++)
      for (int x = 0; x < 8; x++) m[y * 8 + x] = 1.0f / w44[(y / 2) * 4 + x / 2];
    reciprocals(m, inv(blob, JXG_QT_DCT4X4, c), 64);
    // DCT4X8: the refitted signalled table (tables_4x8, :1810-1836)
    float b48[4];
    b48[0] = f16r(kDct4x8Bands[c][0] * kQm4x8Scale[c] / 64.0f) * 64.0f;
    for (int i = 1; i < 4; i++) b48[i] = f16r(kDct4x8Bands[c][i]);
    weights4(fused, b48, 8, w48);
    m = mat(blob, JXG_QT_DCT4X8, c);
    for (int y = 0; y < 8; y++)
      for (int x = 0; x < 8; x++) m[y * 8 + x] = 1.0f / w48[(y / 2) * 8 + x];
    reciprocals(m, inv(blob, JXG_QT_DCT4X8, c), 64);
    // AFV: the library DCT4X8 and DCT4X4 weights around the AFV corner
    // (compute_afv_matrix_from, :1474-1562)
    weights4(fused, kDct4x8Bands[c], 8, w48);
    weights4(fused, kDct4x4Bands[c], 4, w44);
    float ab[4];
    ab[0] = kAfvBands[c][5];
    for (int i = 1; i < 4; i++) ab[i] = ab[i - 1] * band_mult(kAfvBands[c][i + 5]);
    w[0] = 1.0f;
    w[8] = kAfvBands[c][0];
    w[1] = kAfvBands[c][1];
    w[16] = kAfvBands[c][2];
    w[2] = kAfvBands[c][3];
    w[18] = kAfvBands[c][4];
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        if (x < 2 && y < 2) continue;
        const float pos = (kAfvFreqs[y * 4 + x] - kAfvFreqLo) * 3.0f / kAfvFreqHi;
        w[(2 * y) * 8 + 2 * x] = interpolate_bands(pos, ab);
      }
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 8; x++) {
        if (x == 0 && y == 0) continue;
        w[(2 * y + 1) * 8 + x] = w48[y * 8 + x];
      }
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        if (x == 0 && y == 0) continue;
        w[(2 * y) * 8 + 2 * x + 1] = w44[y * 4 + x];
      }
    m = mat(blob, JXG_QT_AFV, c);
    for (int k = 0; k < 64; k++) m[k] = 1.0f / w[k];
    reciprocals(m, inv(blob, JXG_QT_AFV, c), 64);
  }
}

}  // namespace

extern "C" {

int jxg_quant_tier(float distance, int arbiter) {
  const float ss2_from = arbiter == JXG_ARBITER_CVVDP ? 1.75f : 2.25f;
  if (!(distance >= kFlatB8MinDistance)) return 0;
  if (distance >= ss2_from) return distance >= kOriginalCoarseDct8Distance ? 4 : 3;
  return distance >= kFlatB8MidMinDistance ? 2 : 1;
}

int jxg_default_quant_tables(float distance, int arbiter, int fused, float *blob) {
  if (!blob) return JXG_ERR_INVALID_ARG;
  if (!(distance > 0.0f) || !std::isfinite(distance)) return JXG_ERR_INVALID_DISTANCE;
  if (arbiter != JXG_ARBITER_SSIMULACRA2 && arbiter != JXG_ARBITER_CVVDP)
    return JXG_ERR_INVALID_ARG;
  const bool f = fused != 0;
  memset(blob, 0, JXG_QT_BLOB_FLOATS * sizeof(float));
  const int tier = jxg_quant_tier(distance, arbiter);
  // DCT8 (compute_with_ordinary_fit without the ordinary fit, :2727-2760)
  if (tier == 0)
    put_static(blob, JXG_QT_DCT8, &kStatic8x8[0][0], 64);
  else
    put_bands(f, blob, JXG_QT_DCT8, 8, 8,
              dct8_override(tier == 4, tier == 1 ? kFlatB8Hq : kFlatB8Mid));
  // DCT16X16 / DCT32X32 / DCT32X16 (large_tables, dct16/32_override and
  // rect32_override, :2068-2192): the SS2 scales from tier 3 on.
  const bool ss2 = tier >= 3;
  put_bands(f, blob, JXG_QT_DCT16X16, 16, 16,
            ss2 ? scaled(&kDct16x16Bands[0][0], 7, kSsScale16) : plain(&kDct16x16Bands[0][0], 7));
  put_bands(f, blob, JXG_QT_DCT32X32, 32, 32,
            ss2 ? scaled(&kDct32x32Bands[0][0], 8, kSsScale32) : plain(&kDct32x32Bands[0][0], 8));
  put_bands(f, blob, JXG_QT_DCT32X16, 16, 32,
            ss2 ? scaled(&kDct16x32Bands[0][0], 8, kSsScale16x32)
                : plain(&kDct16x32Bands[0][0], 8));
  put_shared(f, blob);
  return JXG_OK;
}

}  // extern "C"
