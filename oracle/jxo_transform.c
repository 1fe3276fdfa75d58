/*
 * jxo_transform.c - oracle: forward transforms (scalar butterflies and 2-D
 * assemblies of src/dct.rs), DC extraction from the lowest frequencies, and
 * the first-pass quantization of write_ac_group (src/group.rs).
 * TEST INFRASTRUCTURE ONLY (see jxo.h).
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "jxo.h"
#include "jxo_internal.h"

#define SQRT2F 1.41421356237309504880f

/* Butterfly constants 1/(2 cos(k pi / 2N)) as stored by the reference
 * (WC4..WC32, src/dct.rs:55-59,658-680,1432-1450). */
static const float WC4[2] = {0.5411961f, 1.306563f};
static const float WC8[4] = {0.5097956f, 0.6013449f, 0.8999762f, 2.5629156f};
static const float WC16[8] = {0.5024193f, 0.5224986f, 0.566944f,  0.6468218f,
                              0.7881546f, 1.0606777f, 1.7224471f, 5.1011486f};
static const float WC32[16] = {0.50060298f, 0.50547096f, 0.5154473f, 0.5310426f,
                               0.5531039f,  0.582935f,   0.6225041f, 0.67480834f,
                               0.7445363f,  0.83934965f, 0.97256825f, 1.16944f,
                               1.4841646f,  2.057781f,   3.4076084f, 10.190008f};
static const float RS16_2[2] = {1.0f, 0.9017642f};                      /* :925 */
static const float RS32_4[4] = {1.0f, 0.9748868f, 0.9017642f, 0.7870549f}; /* :2018 */

const float *jxo_wc(int n) {
  switch (n) {
    case 4: return WC4;
    case 8: return WC8;
    case 16: return WC16;
    case 32: return WC32;
  }
  return 0;
}

/* dct1d_2 / dct1d_4 / dct1d_8 (src/dct.rs:68-127) */
static void dct1d_2(float *buf) {
  float a = buf[0], b = buf[1];
  buf[0] = a + b;
  buf[1] = a - b;
}

static void dct1d_4(float *buf) {
  float tmp[4];
  tmp[0] = buf[0] + buf[3];
  tmp[1] = buf[1] + buf[2];
  dct1d_2(tmp);
  tmp[2] = buf[0] - buf[3];
  tmp[3] = buf[1] - buf[2];
  tmp[2] *= WC4[0];
  tmp[3] *= WC4[1];
  dct1d_2(tmp + 2);
  tmp[2] = jxo_fmla(tmp[2], SQRT2F, tmp[3]);
  buf[0] = tmp[0];
  buf[2] = tmp[1];
  buf[1] = tmp[2];
  buf[3] = tmp[3];
}

static void dct1d_8(float *buf) {
  float tmp[8];
  for (int i = 0; i < 4; i++) tmp[i] = buf[i] + buf[7 - i];
  dct1d_4(tmp);
  for (int i = 0; i < 4; i++) tmp[4 + i] = (buf[i] - buf[7 - i]) * WC8[i];
  dct1d_4(tmp + 4);
  tmp[4] = jxo_fmla(tmp[4], SQRT2F, tmp[5]);
  tmp[5] += tmp[6];
  tmp[6] += tmp[7];
  for (int i = 0; i < 4; i++) {
    buf[2 * i] = tmp[i];
    buf[2 * i + 1] = tmp[4 + i];
  }
}

/* dct1d_16 (src/dct.rs:687-720) */
static void dct1d_16(float *buf) {
  float tmp[16];
  for (int i = 0; i < 8; i++) {
    tmp[i] = buf[i] + buf[15 - i];
    tmp[8 + i] = buf[i] - buf[15 - i];
  }
  dct1d_8(tmp);
  for (int i = 0; i < 8; i++) tmp[8 + i] *= WC16[i];
  dct1d_8(tmp + 8);
  tmp[8] = jxo_fmla(tmp[8], SQRT2F, tmp[9]);
  for (int i = 9; i < 15; i++) tmp[i] += tmp[i + 1];
  for (int i = 0; i < 8; i++) {
    buf[2 * i] = tmp[i];
    buf[2 * i + 1] = tmp[8 + i];
  }
}

/* dct1d_32 (src/dct.rs:1477-1500) */
static void dct1d_32(float *buf) {
  float tmp[32];
  for (int i = 0; i < 16; i++) {
    tmp[i] = buf[i] + buf[31 - i];
    tmp[16 + i] = buf[i] - buf[31 - i];
  }
  dct1d_16(tmp);
  for (int i = 0; i < 16; i++) tmp[16 + i] *= WC32[i];
  dct1d_16(tmp + 16);
  tmp[16] = jxo_fmla(tmp[16], SQRT2F, tmp[17]);
  for (int i = 17; i < 31; i++) tmp[i] += tmp[i + 1];
  for (int i = 0; i < 16; i++) {
    buf[2 * i] = tmp[i];
    buf[2 * i + 1] = tmp[16 + i];
  }
}

/* WC64 (src/dct.rs:2404-2437): the reference's literal butterfly constants
 * (a few differ by one ulp from 1 / (2 cos((2i+1) pi / 128)) rounded to f32,
 * so the literals themselves are used). */
static const float WC64[32] = {
    0.500150621f, 0.501358449f, 0.50378871f,  0.507471144f, 0.51245147f,  0.518792689f,
    0.526577294f, 0.535909832f, 0.546920419f, 0.559769809f, 0.574655175f, 0.591818511f,
    0.611557364f, 0.634238958f, 0.660319805f, 0.690372109f, 0.725120544f, 0.765494168f,
    0.812702119f, 0.868344724f, 0.934583604f, 1.01440823f,  1.11207163f,  1.23383272f,
    1.38929391f,  1.59397233f,  1.87467599f,  2.28204989f,  2.9246285f,   4.08461094f,
    6.79675055f,  20.3738785f};

const float *jxo_wc64(void) { return WC64; }

/* dct1d_64 (src/dct.rs:2529-2546) */
static void dct1d_64(float *buf) {
  float tmp[64];
  for (int i = 0; i < 32; i++) {
    tmp[i] = buf[i] + buf[63 - i];
    tmp[32 + i] = (buf[i] - buf[63 - i]) * WC64[i];
  }
  dct1d_32(tmp);
  dct1d_32(tmp + 32);
  tmp[32] = jxo_fmla(tmp[32], SQRT2F, tmp[33]);
  for (int i = 33; i < 63; i++) tmp[i] += tmp[i + 1];
  for (int i = 0; i < 32; i++) {
    buf[2 * i] = tmp[i];
    buf[2 * i + 1] = tmp[32 + i];
  }
}

/* WC128: 1 / (2 cos((2i+1) pi / 256)) rounded to f32, which equals the
 * reference's literals (src/dct128.rs:76-141) at every index. */
static float WC128_[64];
static pthread_once_t wc128_once = PTHREAD_ONCE_INIT;
static void wc128_init(void) {
  for (int i = 0; i < 64; i++) WC128_[i] = (float)(0.5 / cos((2 * i + 1) * M_PI / 256.0));
}
const float *jxo_wc128(void) {
  pthread_once(&wc128_once, wc128_init);
  return WC128_;
}

/* dct1d_128 (src/dct128.rs:330-352) */
static void dct1d_128(float *buf) {
  const float *wc = jxo_wc128();
  float tmp[128];
  for (int i = 0; i < 64; i++) {
    tmp[i] = buf[i] + buf[127 - i];
    tmp[64 + i] = (buf[i] - buf[127 - i]) * wc[i];
  }
  dct1d_64(tmp);
  dct1d_64(tmp + 64);
  tmp[64] = jxo_fmla(tmp[64], SQRT2F, tmp[65]);
  for (int i = 65; i < 127; i++) tmp[i] += tmp[i + 1];
  for (int i = 0; i < 64; i++) {
    buf[2 * i] = tmp[i];
    buf[2 * i + 1] = tmp[64 + i];
  }
}

/* The 128-family 2-D assemblies: kind 0 DCT128X128 (dct128x128_scalar_input,
 * src/dct128.rs:385-410: columns, then rows scaled 1/16384, out[u*128+v]),
 * kind 1 DCT128X64 (64 wide, 128 tall; :416-441: 128-point columns, 64-point
 * rows, out[u*128+v], 1/8192), kind 2 DCT64X128 (128 wide, 64 tall;
 * :446-468: 128-point rows, 64-point columns, out[v*128+u], 1/8192).  `in`
 * has stride W (the footprint's width). */
void jxo_dct128(int kind, const float *in, float *out) {
  float line[128];
  if (kind == 0 || kind == 1) {
    const int W = kind == 0 ? 128 : 64;
    float *after = (float *)malloc((size_t)128 * W * sizeof(float));
    for (int x = 0; x < W; x++) {
      for (int y = 0; y < 128; y++) line[y] = in[(size_t)y * W + x];
      dct1d_128(line);
      for (int v = 0; v < 128; v++) after[(size_t)v * W + x] = line[v];
    }
    const float scale = kind == 0 ? 1.0f / 16384.0f : 1.0f / 8192.0f;
    for (int v = 0; v < 128; v++) {
      float *row = after + (size_t)v * W;
      if (kind == 0)
        dct1d_128(row);
      else
        dct1d_64(row);
      for (int u = 0; u < W; u++) out[u * 128 + v] = row[u] * scale;
    }
    free(after);
  } else {
    float *after = (float *)malloc((size_t)64 * 128 * sizeof(float));
    for (int y = 0; y < 64; y++) {
      memcpy(after + y * 128, in + (size_t)y * 128, 128 * sizeof(float));
      dct1d_128(after + y * 128);
    }
    const float scale = 1.0f / 8192.0f;
    for (int u = 0; u < 128; u++) {
      float col[64];
      for (int y = 0; y < 64; y++) col[y] = after[y * 128 + u];
      dct1d_64(col);
      for (int v = 0; v < 64; v++) out[v * 128 + u] = col[v] * scale;
    }
    free(after);
  }
}

static void dct1d(float *buf, int n) {
  switch (n) {
    case 8: dct1d_8(buf); break;
    case 16: dct1d_16(buf); break;
    case 32: dct1d_32(buf); break;
    case 64: dct1d_64(buf); break;
    default: abort();
  }
}

/* The 2-D assemblies (src/dct.rs:631-655 8x8, :823-845 8x16, :904-925 16x8,
 * :1010-1030 16x16, :1980-2000 32x32, :2251-2270 32x16, :2333-2352 16x32,
 * :1742-1760 32x8, :1777-1795 8x32).  `in` is a W x H pixel block with
 * stride W.  Two pass orders occur: rows first (8x8, 8x16, 16x32, 8x32) and
 * columns first (all others); the output is the long axis as columns. */
static void dct_rows_first(const float *in, int W, int H, float *out) {
  /* row DCT (length W) of each row, then column DCT (length H) of each
   * column u; output[v * W + u] for wide shapes, output[u * 8 + v] for the
   * square 8x8 (src/dct.rs:648-653). */
  float tmp[64 * 64];
  float row[64], col[64];
  for (int y = 0; y < H; y++) {
    memcpy(row, in + y * W, W * sizeof(float));
    dct1d(row, W);
    memcpy(tmp + y * W, row, W * sizeof(float));
  }
  float scale = 1.0f / (float)(W * H);
  for (int u = 0; u < W; u++) {
    for (int i = 0; i < H; i++) col[i] = tmp[i * W + u];
    dct1d(col, H);
    for (int v = 0; v < H; v++) {
      if (W > H)
        out[v * W + u] = col[v] * scale;
      else
        out[u * H + v] = col[v] * scale; /* square 8x8: transposed storage */
    }
  }
}

static void dct_cols_first(const float *in, int W, int H, float *out) {
  /* column DCT (length H) of each column u into tmp[v*W+u], then row DCT
   * (length W) of each row v; output[u * H + v] when H >= W, else
   * output[v * W + u]. */
  float tmp[64 * 64];
  float col[64];
  for (int u = 0; u < W; u++) {
    for (int i = 0; i < H; i++) col[i] = in[i * W + u];
    dct1d(col, H);
    for (int v = 0; v < H; v++) tmp[v * W + u] = col[v];
  }
  float scale = 1.0f / (float)(W * H);
  for (int v = 0; v < H; v++) {
    float *row = tmp + v * W;
    dct1d(row, W);
    for (int u = 0; u < W; u++) {
      if (H >= W)
        out[u * H + v] = row[u] * scale;
      else
        out[v * W + u] = row[u] * scale;
    }
  }
}

/* 8x8 uses an explicit scale constant 1/64 (src/dct.rs:650) which equals
 * 1/(W*H) in f32; all shapes share one scale convention 1/(W*H). */

/* identity8x8 (src/dct.rs:185-214) */
static void identity8x8(const float *in, float *out) {
  for (int qy = 0; qy < 2; qy++) {
    for (int qx = 0; qx < 2; qx++) {
      float anchor = in[(qy * 4 + 1) * 8 + qx * 4 + 1];
      float mean = 0.0f;
      for (int iy = 0; iy < 4; iy++) {
        const float *row = in + (qy * 4 + iy) * 8;
        for (int ix = 0; ix < 4; ix++) {
          mean += row[qx * 4 + ix];
          if (ix == 1 && iy == 1) continue;
          out[(qy + iy * 2) * 8 + qx + ix * 2] = row[qx * 4 + ix] - anchor;
        }
      }
      out[(qy + 2) * 8 + qx + 2] = out[qy * 8 + qx];
      out[qy * 8 + qx] = mean * (1.0f / 16.0f);
    }
  }
  float b00 = out[0], b01 = out[1], b10 = out[8], b11 = out[9];
  out[0] = (b00 + b01 + b10 + b11) * 0.25f;
  out[1] = (b00 + b01 - b10 - b11) * 0.25f;
  out[8] = (b00 - b01 + b10 - b11) * 0.25f;
  out[9] = (b00 - b01 - b10 + b11) * 0.25f;
}

/* dct2_top_block / dct2x2_8x8 (src/dct.rs:270-306) */
static void dct2_top_block(const float *in, int S, float *out) {
  float temp[64];
  memset(temp, 0, sizeof(temp));
  int half = S / 2;
  for (int y = 0; y < half; y++) {
    const float *row0 = in + (y * 2) * 8;
    const float *row1 = in + (y * 2 + 1) * 8;
    for (int x = 0; x < half; x++) {
      float c00 = row0[x * 2], c01 = row0[x * 2 + 1];
      float c10 = row1[x * 2], c11 = row1[x * 2 + 1];
      temp[y * 8 + x] = (c00 + c01 + c10 + c11) * 0.25f;
      temp[y * 8 + half + x] = (c00 + c01 - c10 - c11) * 0.25f;
      temp[(y + half) * 8 + x] = (c00 - c01 + c10 - c11) * 0.25f;
      temp[(y + half) * 8 + half + x] = (c00 - c01 - c10 + c11) * 0.25f;
    }
  }
  for (int y = 0; y < S; y++) memcpy(out + y * 8, temp + y * 8, S * sizeof(float));
}

static void dct2x2_8x8(const float *in, float *out) {
  float snap[64];
  dct2_top_block(in, 8, out);
  memcpy(snap, out, sizeof(snap));
  dct2_top_block(snap, 4, out);
  memcpy(snap, out, sizeof(snap));
  dct2_top_block(snap, 2, out);
}

/* dct4x4_2d (src/dct.rs:1061-1076): rows, then columns stored transposed
 * (out[hf*4 + vf]), scaled 1/16. */
static void dct4x4_2d(const float *in, float *out) {
  float tmp[16];
  memcpy(tmp, in, sizeof(tmp));
  for (int r = 0; r < 4; r++) dct1d_4(tmp + r * 4);
  for (int x = 0; x < 4; x++) {
    float col[4] = {tmp[x], tmp[4 + x], tmp[8 + x], tmp[12 + x]};
    dct1d_4(col);
    for (int i = 0; i < 4; i++) out[x * 4 + i] = col[i] * (1.0f / 16.0f);
  }
}

/* dct4x8_2d (src/dct.rs:1198-1220): a 4-tall x 8-wide half; columns
 * (4-point) then rows (8-point), out[vf*8 + hf], scaled 1/32. */
static void dct4x8_2d(const float *in, float *out) {
  float tmp[32];
  memcpy(tmp, in, sizeof(tmp));
  for (int c = 0; c < 8; c++) {
    float col[4] = {tmp[c], tmp[8 + c], tmp[16 + c], tmp[24 + c]};
    dct1d_4(col);
    for (int vf = 0; vf < 4; vf++) tmp[vf * 8 + c] = col[vf];
  }
  for (int vf = 0; vf < 4; vf++) {
    dct1d_8(tmp + vf * 8);
    for (int i = 0; i < 8; i++) tmp[vf * 8 + i] *= 1.0f / 32.0f;
  }
  memcpy(out, tmp, sizeof(tmp));
}

/* dct8x4_2d (src/dct.rs:1267-1300): an 8-tall x 4-wide half; columns
 * (8-point) then rows (4-point), stored out[hf*8 + vf], scaled 1/32. */
static void dct8x4_2d(const float *in, float *out) {
  float tmp[32];
  for (int c = 0; c < 4; c++) {
    float col[8];
    for (int i = 0; i < 8; i++) col[i] = in[i * 4 + c];
    dct1d_8(col);
    for (int vf = 0; vf < 8; vf++) tmp[vf * 4 + c] = col[vf];
  }
  for (int vf = 0; vf < 8; vf++) {
    float row[4] = {tmp[vf * 4], tmp[vf * 4 + 1], tmp[vf * 4 + 2], tmp[vf * 4 + 3]};
    dct1d_4(row);
    for (int hf = 0; hf < 4; hf++) out[hf * 8 + vf] = row[hf] * (1.0f / 32.0f);
  }
}

/* dct4x4_scalar_input (src/dct.rs:1150-1180): four quadrant DCT4x4s
 * interleaved (quadrant (qy, qx) at rows qy+2iy, columns qx+2ix), then a
 * 2x2 Hadamard of the four sub-DCs. */
static void dct4x4_8x8(const float *in, float *out) {
  for (int qy = 0; qy < 2; qy++)
    for (int qx = 0; qx < 2; qx++) {
      float blk[16], d[16];
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) blk[r * 4 + c] = in[(qy * 4 + r) * 8 + qx * 4 + c];
      dct4x4_2d(blk, d);
      for (int iy = 0; iy < 4; iy++)
        for (int ix = 0; ix < 4; ix++) out[(qy + iy * 2) * 8 + qx + ix * 2] = d[iy * 4 + ix];
    }
  float b00 = out[0], b01 = out[1], b10 = out[8], b11 = out[9];
  out[0] = (b00 + b01 + b10 + b11) * 0.25f;
  out[1] = (b00 + b01 - b10 - b11) * 0.25f;
  out[8] = (b00 - b01 + b10 - b11) * 0.25f;
  out[9] = (b00 - b01 - b10 + b11) * 0.25f;
}

/* dct4x8_scalar_input (src/dct.rs:1230-1255): the two 4-tall halves
 * interleaved on even / odd rows, then a 2-point Hadamard of their DCs. */
static void dct4x8_8x8(const float *in, float *out) {
  for (int y = 0; y < 2; y++) {
    float d[32];
    dct4x8_2d(in + y * 32, d);
    for (int iy = 0; iy < 4; iy++)
      for (int ix = 0; ix < 8; ix++) out[(y + iy * 2) * 8 + ix] = d[iy * 8 + ix];
  }
  float b0 = out[0], b1 = out[8];
  out[0] = (b0 + b1) * 0.5f;
  out[8] = (b0 - b1) * 0.5f;
}

/* dct8x4_scalar_input (src/dct.rs:1310-1340): the two 4-wide halves. */
static void dct8x4_8x8(const float *in, float *out) {
  for (int x = 0; x < 2; x++) {
    float half[32], d[32];
    for (int r = 0; r < 8; r++)
      for (int c = 0; c < 4; c++) half[r * 4 + c] = in[r * 8 + x * 4 + c];
    dct8x4_2d(half, d);
    for (int iy = 0; iy < 4; iy++)
      for (int ix = 0; ix < 8; ix++) out[(x + iy * 2) * 8 + ix] = d[iy * 8 + ix];
  }
  float b0 = out[0], b1 = out[8];
  out[0] = (b0 + b1) * 0.5f;
  out[8] = (b0 - b1) * 0.5f;
}

/* The JPEG XL AFV basis (the format's 4x4 "AFV DCT" basis; the same values
 * as src/afv.rs:48 AFV_BASIS_TRANSPOSE, f32): row j = pixel j, column i =
 * basis vector i.  Orthonormal; column 0 is the constant 1/4. */
static const float AFV_BASIS[256] = {
    0.25f, 0.876902938f, 0.0f, 0.0f, 0.0f, -0.41053775f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f,
    0.25f, 0.220651805f, 0.0f, 0.0f, -0.707106769f, 0.623548508f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f,
    0.25f, -0.101400502f, 0.40670076f, -0.21255748f, 0.0f, -0.0643507168f, -0.451755673f, -0.304684758f, 0.301792949f, 0.408248305f, 0.174786702f, -0.211056009f, -0.142660841f, -0.138135403f, -0.174376026f, 0.113549873f,
    0.25f, -0.101400502f, 0.444448173f, 0.308549702f, 0.0f, -0.0643507168f, 0.158545032f, 0.511261642f, 0.257923633f, 0.0f, 0.0812611207f, 0.185671806f, -0.341644675f, 0.330228269f, 0.0702790692f, -0.0741750449f,
    0.25f, 0.220651805f, 0.0f, 0.0f, 0.707106769f, 0.623548508f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f,
    0.25f, -0.101400502f, 0.0f, 0.470670223f, 0.0f, -0.0643507168f, -0.0403851531f, 0.0f, 0.162723407f, 0.0f, 0.0f, 0.0f, 0.736749768f, 0.0875511467f, -0.292102665f, 0.194028929f,
    0.25f, -0.101400502f, 0.195743993f, -0.162120521f, 0.0f, -0.0643507168f, 0.00741822645f, -0.290480137f, 0.0952002257f, 0.0f, -0.367539793f, 0.492158592f, 0.246271074f, -0.0794670656f, 0.362381727f, -0.435190499f,
    0.25f, -0.101400502f, 0.29291001f, 0.0f, 0.0f, -0.0643507168f, 0.393510342f, -0.0657870173f, 0.0f, -0.408248305f, -0.30788222f, -0.385250151f, -0.0857401937f, -0.461337477f, 0.0f, 0.219186842f,
    0.25f, -0.101400502f, -0.40670076f, -0.21255748f, 0.0f, -0.0643507168f, -0.451755673f, 0.304684758f, 0.301792949f, -0.408248305f, -0.174786702f, 0.211056009f, -0.142660841f, -0.138135403f, -0.174376026f, 0.113549873f,
    0.25f, -0.101400502f, -0.195743993f, -0.162120521f, 0.0f, -0.0643507168f, 0.00741822645f, 0.290480137f, 0.0952002257f, 0.0f, 0.367539793f, -0.492158592f, 0.246271074f, -0.0794670656f, 0.362381727f, -0.435190499f,
    0.25f, -0.101400502f, 0.0f, -0.470670223f, 0.0f, -0.0643507168f, 0.11074166f, 0.0f, -0.162723407f, 0.0f, 0.0f, 0.0f, 0.14883399f, 0.497246474f, 0.292102665f, 0.555044353f,
    0.25f, -0.101400502f, 0.113790743f, -0.146429181f, 0.0f, -0.0643507168f, 0.0829816312f, -0.238897741f, -0.353123844f, -0.408248305f, 0.482668906f, 0.174194127f, -0.0476868041f, 0.125380591f, -0.432660788f, -0.254682779f,
    0.25f, -0.101400502f, -0.444448173f, 0.308549702f, 0.0f, -0.0643507168f, 0.158545032f, -0.511261642f, 0.257923633f, 0.0f, -0.0812611207f, -0.185671806f, -0.341644675f, 0.330228269f, 0.0702790692f, -0.0741750449f,
    0.25f, -0.101400502f, -0.29291001f, 0.0f, 0.0f, -0.0643507168f, 0.393510342f, 0.0657870173f, 0.0f, 0.408248305f, 0.30788222f, 0.385250151f, -0.0857401937f, -0.461337477f, 0.0f, 0.219186842f,
    0.25f, -0.101400502f, -0.113790743f, -0.146429181f, 0.0f, -0.0643507168f, 0.0829816312f, 0.238897741f, -0.353123844f, 0.408248305f, -0.482668906f, -0.174194127f, -0.0476868041f, 0.125380591f, -0.432660788f, -0.254682779f,
    0.25f, -0.101400502f, 0.0f, 0.425114959f, 0.0f, -0.0643507168f, -0.451755673f, 0.0f, -0.603585899f, 0.0f, 0.0f, 0.0f, -0.142660841f, -0.138135403f, 0.348752052f, 0.113549873f,
};

const float *jxo_afv_basis(void) { return AFV_BASIS; }

/* afv_dct4x4 (src/afv.rs:216-225): coeffs[i] = sum_j pixels[j] * B[j][i],
 * accumulated in pixel order with fmla. */
static void afv_dct4x4(const float *px, float *out) {
  for (int i = 0; i < 16; i++) {
    float acc = 0.0f;
    for (int j = 0; j < 16; j++) acc = jxo_fmla(px[j], AFV_BASIS[j * 16 + i], acc);
    out[i] = acc;
  }
}

/* afv_from_pixels (src/afv.rs:207-256).  kind bit 0: AFV quadrant on the
 * right; bit 1: on the bottom. */
static void afv_8x8(int kind, const float *in, float *out) {
  const int afv_x = kind & 1, afv_y = kind >> 1;
  float quad[16], coeff[16], d44[16], half[32], d48[32];
  for (int i = 0; i < 16; i++) {
    const int dy = i / 4, dx = i % 4;
    const int sy = afv_y ? 3 - dy : dy, sx = afv_x ? 3 - dx : dx;
    quad[i] = in[(sy + 4 * afv_y) * 8 + sx + 4 * afv_x];
  }
  afv_dct4x4(quad, coeff);
  for (int iy = 0; iy < 4; iy++)
    for (int ix = 0; ix < 4; ix++) out[iy * 2 * 8 + ix * 2] = coeff[iy * 4 + ix];
  const int qx = afv_x ? 0 : 4;
  for (int i = 0; i < 16; i++) quad[i] = in[(i / 4 + 4 * afv_y) * 8 + qx + i % 4];
  dct4x4_2d(quad, d44);
  for (int iy = 0; iy < 4; iy++)
    for (int ix = 0; ix < 4; ix++) out[iy * 2 * 8 + ix * 2 + 1] = d44[iy * 4 + ix];
  const int hy = afv_y ? 0 : 4;
  for (int i = 0; i < 32; i++) half[i] = in[(hy + i / 8) * 8 + i % 8];
  dct4x8_2d(half, d48);
  for (int iy = 0; iy < 4; iy++)
    for (int ix = 0; ix < 8; ix++) out[(1 + iy * 2) * 8 + ix] = d48[iy * 8 + ix];
  const float block00 = out[0] * 0.25f, block01 = out[1], block10 = out[8];
  out[0] = (block00 + block01 + 2.0f * block10) * 0.25f;
  out[1] = (block00 - block01) * 0.5f;
  out[8] = (block00 + block01 - 2.0f * block10) * 0.25f;
}

/* ---- strategy geometry (src/dc_group_data.rs:140-170) ---- */
static const uint8_t COV_X[24] = {1, 1, 2, 2, 4, 1, 1, 1, 2, 4, 1, 1,
                                  1, 1, 8, 4, 8, 1, 1, 16, 8, 16, 1, 4};
static const uint8_t COV_Y[24] = {1, 2, 1, 2, 4, 1, 1, 1, 4, 2, 1, 1,
                                  1, 1, 8, 8, 4, 1, 1, 16, 16, 8, 4, 1};
int jxo_covered_x(int s) { return COV_X[s]; }
int jxo_covered_y(int s) { return COV_Y[s]; }

void jxo_strategy_pixel_dims(int s, int *w, int *h) {
  *w = COV_X[s] * 8;
  *h = COV_Y[s] * 8;
}

/* Footprint gather with edge replication (gather_pixels,
 * src/ac_strategy.rs:451-479). */
static void gather(const float *plane, int pw, int ph, int px, int py, int w, int h,
                   float *dst) {
  for (int v = 0; v < h; v++) {
    int sy = py + v < ph ? py + v : ph - 1;
    const float *row = plane + (size_t)sy * pw;
    for (int u = 0; u < w; u++) {
      int sx = px + u < pw ? px + u : pw - 1;
      dst[v * w + u] = row[sx];
    }
  }
}

int jxo_forward_transform(int strategy, const float *plane, int width, int height, int px,
                          int py, float *out) {
  int w, h;
  jxo_strategy_pixel_dims(strategy, &w, &h);
  float pix[64 * 64];
  gather(plane, width, height, px, py, w, h, pix);
  switch (strategy) {
    case JXG_DCT: dct_rows_first(pix, 8, 8, out); break;
    case JXG_DCT8X16: dct_rows_first(pix, 16, 8, out); break;   /* W16 H8 */
    case JXG_DCT16X8: dct_cols_first(pix, 8, 16, out); break;   /* W8 H16 */
    case JXG_DCT16X16: dct_cols_first(pix, 16, 16, out); break;
    case JXG_DCT32X32: dct_cols_first(pix, 32, 32, out); break;
    case JXG_DCT32X16: dct_cols_first(pix, 16, 32, out); break; /* W16 H32 */
    case JXG_DCT16X32: dct_rows_first(pix, 32, 16, out); break; /* W32 H16 */
    case JXG_DCT32X8: dct_cols_first(pix, 8, 32, out); break;   /* W8 H32 */
    case JXG_DCT8X32: dct_rows_first(pix, 32, 8, out); break;   /* W32 H8 */
    case JXG_DCT64X64: dct_cols_first(pix, 64, 64, out); break; /* src/dct.rs:2576 */
    case JXG_DCT64X32: dct_cols_first(pix, 32, 64, out); break; /* W32 H64, :2660 */
    case JXG_DCT32X64: dct_rows_first(pix, 64, 32, out); break; /* W64 H32, :2688 */
    case JXG_IDENTITY: identity8x8(pix, out); break;
    case JXG_DCT2X2: dct2x2_8x8(pix, out); break;
    case JXG_DCT4X4: dct4x4_8x8(pix, out); break;
    case JXG_DCT4X8: dct4x8_8x8(pix, out); break;
    case JXG_DCT8X4: dct8x4_8x8(pix, out); break;
    case JXG_AFV0:
    case JXG_AFV1:
    case JXG_AFV2:
    case JXG_AFV3: afv_8x8(strategy - JXG_AFV0, pix, out); break;
    default: abort();
  }
  return w * h;
}

/* idct1d_4 (src/dct.rs:2144-2168) */
static void idct1d_4(const float v[4], float out[4]) {
  float t[4] = {v[0], v[2], v[1], v[3]};
  float e0 = t[0] + t[1], e1 = t[0] - t[1];
  t[0] = e0;
  t[1] = e1;
  t[3] += t[2];
  t[2] *= SQRT2F;
  float o0 = t[2] + t[3], o1 = t[2] - t[3];
  t[2] = o0;
  t[3] = o1;
  out[0] = jxo_fmla(WC4[0], t[2], t[0]);
  out[3] = jxo_fmla(-WC4[0], t[2], t[0]);
  out[1] = jxo_fmla(WC4[1], t[3], t[1]);
  out[2] = jxo_fmla(-WC4[1], t[3], t[1]);
}

/* inv_dct1d_2 / inv_dct1d_4 / inv_dct1d_8 (src/dct.rs:1500-1543) */
static const float INV_WC4[2] = {1.0f / 0.5411961f, 1.0f / 1.306563f};
static void inv_dct1d_2(float *b) {
  float s = b[0], d = b[1];
  b[0] = (s + d) * 0.5f;
  b[1] = (s - d) * 0.5f;
}
static void inv_dct1d_4(float *b) {
  const float IS2 = 0.70710678118654752440f;
  float t[4] = {b[0], b[2], b[1], b[3]};
  t[2] = (t[2] - t[3]) * IS2;
  inv_dct1d_2(t + 2);
  t[2] *= INV_WC4[0];
  t[3] *= INV_WC4[1];
  inv_dct1d_2(t);
  b[0] = (t[0] + t[2]) * 0.5f;
  b[3] = (t[0] - t[2]) * 0.5f;
  b[1] = (t[1] + t[3]) * 0.5f;
  b[2] = (t[1] - t[3]) * 0.5f;
}
static void inv_dct1d_8(float *b) {
  const float IS2 = 0.70710678118654752440f;
  const float INV_WC8[4] = {1.0f / WC8[0], 1.0f / WC8[1], 1.0f / WC8[2], 1.0f / WC8[3]};
  float t[8];
  for (int i = 0; i < 4; i++) {
    t[i] = b[2 * i];
    t[4 + i] = b[2 * i + 1];
  }
  t[6] -= t[7];
  t[5] -= t[6];
  t[4] = (t[4] - t[5]) * IS2;
  inv_dct1d_4(t + 4);
  for (int i = 0; i < 4; i++) t[4 + i] *= INV_WC8[i];
  inv_dct1d_4(t);
  for (int i = 0; i < 4; i++) {
    b[i] = (t[i] + t[4 + i]) * 0.5f;
    b[7 - i] = (t[i] - t[4 + i]) * 0.5f;
  }
}

/* inv_dct1d_16 / inv_dct1d_32 (src/dct.rs:1547-1627) */
static void inv_dct1d_16(float *b) {
  const float IS2 = 0.70710678118654752440f;
  float inv_wc[8];
  for (int i = 0; i < 8; i++) inv_wc[i] = 1.0f / WC16[i];
  float t[16];
  for (int i = 0; i < 8; i++) {
    t[i] = b[2 * i];
    t[8 + i] = b[2 * i + 1];
  }
  for (int i = 14; i >= 9; i--) t[i] -= t[i + 1];
  t[8] = (t[8] - t[9]) * IS2;
  inv_dct1d_8(t + 8);
  for (int i = 0; i < 8; i++) t[8 + i] *= inv_wc[i];
  inv_dct1d_8(t);
  for (int i = 0; i < 8; i++) {
    b[i] = (t[i] + t[8 + i]) * 0.5f;
    b[15 - i] = (t[i] - t[8 + i]) * 0.5f;
  }
}
static void inv_dct1d_32(float *b) {
  const float IS2 = 0.70710678118654752440f;
  float inv_wc[16];
  for (int i = 0; i < 16; i++) inv_wc[i] = 1.0f / WC32[i];
  float t[32];
  for (int i = 0; i < 16; i++) {
    t[i] = b[2 * i];
    t[16 + i] = b[2 * i + 1];
  }
  for (int i = 30; i >= 17; i--) t[i] -= t[i + 1];
  t[16] = (t[16] - t[17]) * IS2;
  inv_dct1d_16(t + 16);
  for (int i = 0; i < 16; i++) t[16 + i] *= inv_wc[i];
  inv_dct1d_16(t);
  for (int i = 0; i < 16; i++) {
    b[i] = (t[i] + t[16 + i]) * 0.5f;
    b[31 - i] = (t[i] - t[16 + i]) * 0.5f;
  }
}
static void inv_line(float *b, int n) {
  if (n == 8) inv_dct1d_8(b);
  else if (n == 16) inv_dct1d_16(b);
  else inv_dct1d_32(b);
}

/* reconstruct_error (src/inflated_cost.rs:741-788) with the scalar inverse
 * transforms of IdctMethods::scalar (src/dct.rs:1594-1680): inv_dct_square
 * for DCT8 / 16X16 / 32X32 (coefficients transposed into place and scaled
 * by N, columns then rows), inv_dct8x16 (rows of 16 px: 8-point columns,
 * then 16-point rows) and inv_dct16x8 (16 tall: 8-point rows of the
 * transposed storage, then 16-point columns).  out is the pixel footprint,
 * row-major.  Returns 0, or -1 for a strategy these Fast layouts never hold. */
int jxo_inverse_transform(int s, const float *coef, float *out) {
  float col[32], acd[128];
  switch (s) {
    case JXG_DCT:
    case JXG_DCT16X16:
    case JXG_DCT32X32: {
      const int n = s == JXG_DCT ? 8 : (s == JXG_DCT16X16 ? 16 : 32);
      const float scale = (float)(n * n);
      for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) out[b * n + a] = coef[a * n + b] * scale;
      for (int u = 0; u < n; u++) {
        for (int v = 0; v < n; v++) col[v] = out[v * n + u];
        inv_line(col, n);
        for (int v = 0; v < n; v++) out[v * n + u] = col[v];
      }
      for (int v = 0; v < n; v++) inv_line(out + v * n, n);
      return 0;
    }
    case JXG_DCT8X16:
      for (int v = 0; v < 8; v++)
        for (int u = 0; u < 16; u++) out[v * 16 + u] = coef[v * 16 + u] * 128.0f;
      for (int u = 0; u < 16; u++) {
        for (int v = 0; v < 8; v++) col[v] = out[v * 16 + u];
        inv_dct1d_8(col);
        for (int v = 0; v < 8; v++) out[v * 16 + u] = col[v];
      }
      for (int r = 0; r < 8; r++) inv_dct1d_16(out + r * 16);
      return 0;
    case JXG_DCT16X8:
      for (int v = 0; v < 16; v++) {
        float row[8];
        for (int u = 0; u < 8; u++) row[u] = coef[u * 16 + v] * 128.0f;
        inv_dct1d_8(row);
        for (int u = 0; u < 8; u++) acd[v * 8 + u] = row[u];
      }
      for (int u = 0; u < 8; u++) {
        for (int v = 0; v < 16; v++) col[v] = acd[v * 8 + u];
        inv_dct1d_16(col);
        for (int i = 0; i < 16; i++) out[i * 8 + u] = col[i];
      }
      return 0;
  }
  return -1;
}

/* RESAMPLE_SCALE_64_TO_8 (src/dct.rs:2005-2014) */
static const float RS64_8[8] = {1.0f,       0.9936866f, 0.97488683f, 0.94401807f,
                                0.9017642f, 0.8490575f, 0.7870549f,  0.71710813f};

/* dc_from_dct64x64 (src/dct.rs:2602-2611): the 8x8 lowest frequencies,
 * rescaled, through inv_dct8x8 (inv_dct_square, :1594-1621). */
static void dc_from_dct64x64(const float *c, float *dc) {
  float low[64], out[64], col[8];
  for (int y = 0; y < 8; y++)
    for (int x = 0; x < 8; x++) low[y * 8 + x] = c[y * 64 + x] * (RS64_8[x] * RS64_8[y]);
  for (int a = 0; a < 8; a++)
    for (int b = 0; b < 8; b++) out[b * 8 + a] = low[a * 8 + b] * 64.0f;
  for (int u = 0; u < 8; u++) {
    for (int v = 0; v < 8; v++) col[v] = out[v * 8 + u];
    inv_dct1d_8(col);
    for (int v = 0; v < 8; v++) out[v * 8 + u] = col[v];
  }
  for (int v = 0; v < 8; v++) inv_dct1d_8(out + v * 8);
  memcpy(dc, out, sizeof(out));
}

/* dc_from_dct64x32_normalized (src/dct.rs:2767-2785): rows[a] is the
 * unnormalised 8-point inverse of the a-th long-axis frequency row, then a
 * 4-point inverse down each column; out[y][x], y < 4, x < 8. */
static void idct1d_4(const float v[4], float out[4]);
static void dc_64x32_normalized(const float *c, float out[4][8]) {
  float rows[4][8];
  for (int a = 0; a < 4; a++) {
    for (int b = 0; b < 8; b++) rows[a][b] = c[a * 64 + b] * (RS32_4[a] * RS64_8[b]);
    for (int b = 0; b < 8; b++) rows[a][b] *= 8.0f; /* idct1d_8_unnormalized :2756 */
    inv_dct1d_8(rows[a]);
  }
  for (int x = 0; x < 8; x++) {
    const float col_in[4] = {rows[0][x], rows[1][x], rows[2][x], rows[3][x]};
    float col[4];
    idct1d_4(col_in, col);
    for (int y = 0; y < 4; y++) out[y][x] = col[y];
  }
}

/* dc_from_coefficients (src/group.rs:183-260) with dc_from_dct16x8 :930,
 * dc_from_dct16x16 :1033, dc_from_dct32x32 :2170, dc_from_dct32x16 :2356,
 * dc_from_dct16x32 :2381, dc_from_dct8x32 :1864. */
void jxo_dc_from_coefficients(int s, const float *c, float *dc) {
  switch (s) {
    case JXG_DCT:
    case JXG_IDENTITY:
    case JXG_DCT2X2:
    case JXG_DCT4X4:
    case JXG_DCT4X8:
    case JXG_DCT8X4:
    case JXG_AFV0:
    case JXG_AFV1:
    case JXG_AFV2:
    case JXG_AFV3: dc[0] = c[0]; return;
    case JXG_DCT16X8:
    case JXG_DCT8X16: {
      float s0 = c[0] * RS16_2[0];
      float s1 = c[1] * RS16_2[1];
      dc[0] = s0 + s1;
      dc[1] = s0 - s1;
      return;
    }
    case JXG_DCT16X16: {
      float s00 = c[0] * (RS16_2[0] * RS16_2[0]);
      float s01 = c[1] * (RS16_2[0] * RS16_2[1]);
      float s10 = c[16] * (RS16_2[1] * RS16_2[0]);
      float s11 = c[17] * (RS16_2[1] * RS16_2[1]);
      float r00 = s00 + s01, r01 = s00 - s01, r10 = s10 + s11, r11 = s10 - s11;
      dc[0] = r00 + r10;
      dc[1] = r00 - r10;
      dc[2] = r01 + r11;
      dc[3] = r01 - r11;
      return;
    }
    case JXG_DCT32X32: {
      float s4[4][4], rr[4][4];
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) s4[a][b] = c[a * 32 + b] * RS32_4[a] * RS32_4[b];
      for (int a = 0; a < 4; a++) idct1d_4(s4[a], rr[a]);
      for (int bb = 0; bb < 4; bb++) {
        float col_in[4] = {rr[0][bb], rr[1][bb], rr[2][bb], rr[3][bb]}, col[4];
        idct1d_4(col_in, col);
        for (int r = 0; r < 4; r++) dc[bb * 4 + r] = col[r];
      }
      return;
    }
    case JXG_DCT32X16: {
      float rr[2][4];
      for (int a = 0; a < 2; a++) {
        float sv[4];
        for (int b = 0; b < 4; b++) sv[b] = c[a * 32 + b] * RS16_2[a] * RS32_4[b];
        idct1d_4(sv, rr[a]);
      }
      for (int bb = 0; bb < 4; bb++) {
        dc[bb * 2] = rr[0][bb] + rr[1][bb];
        dc[bb * 2 + 1] = rr[0][bb] - rr[1][bb];
      }
      return;
    }
    case JXG_DCT16X32: {
      float rr[2][4];
      for (int a = 0; a < 2; a++) {
        float sv[4];
        for (int b = 0; b < 4; b++) sv[b] = c[a * 32 + b] * RS16_2[a] * RS32_4[b];
        idct1d_4(sv, rr[a]);
      }
      for (int bb = 0; bb < 4; bb++) {
        dc[bb] = rr[0][bb] + rr[1][bb];
        dc[4 + bb] = rr[0][bb] - rr[1][bb];
      }
      return;
    }
    case JXG_DCT32X8:
    case JXG_DCT8X32: {
      float sv[4];
      for (int b = 0; b < 4; b++) sv[b] = c[b] * RS32_4[b];
      idct1d_4(sv, dc);
      return;
    }
    case JXG_DCT64X64: dc_from_dct64x64(c, dc); return;
    case JXG_DCT64X32: { /* 8 rows x 4 columns (src/dct.rs:2789-2796) */
      float n[4][8];
      dc_64x32_normalized(c, n);
      for (int y = 0; y < 8; y++)
        for (int x = 0; x < 4; x++) dc[y * 4 + x] = n[x][y];
      return;
    }
    case JXG_DCT32X64: { /* 4 rows x 8 columns (:2800-2805) */
      float n[4][8];
      dc_64x32_normalized(c, n);
      for (int y = 0; y < 4; y++)
        for (int x = 0; x < 8; x++) dc[y * 8 + x] = n[y][x];
      return;
    }
    default: abort();
  }
}

/* ------------------------------------------------------------------ */
/* Quantization (src/group.rs:394-470, 560-620, 682-810).              */

static float clampf_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

static float x_deadzone_window(float distance) {
  const float a = 0.5f, b = 0.75f, c = 1.5f, d = 2.0f;
  float up = clampf_((distance - a) / (b - a), 0.0f, 1.0f);
  float down = clampf_((d - distance) / (d - c), 0.0f, 1.0f);
  return fminf(up, down);
}

void jxo_ac_thresholds(int c, int xsize, int ysize, float distance, float qm_mult,
                       float thr[4]) {
  float normal[4] = {0.58f, 0.635f, 0.66f, 0.7f};
  if (c == 0) {
    float w = x_deadzone_window(distance);
    normal[0] += w * (0.5f - normal[0]);
    for (int i = 1; i < 4; i++) normal[i] += (1.0f - w) * 0.08f;
  }
  if (c == 2) {
    float bt = qm_mult > 1.0f ? 0.65f : 0.75f;
    for (int i = 1; i < 4; i++) normal[i] = bt;
  }
  if (xsize > 1 || ysize > 1) {
    float delta = clampf_(0.003f * (float)xsize * (float)ysize, 0.0f, c > 0 ? 0.08f : 0.12f);
    for (int i = 0; i < 4; i++) normal[i] -= delta;
  }
  float hq[4];
  if (c == 1) {
    hq[0] = 0.50f;
    hq[1] = 0.51f;
    hq[2] = 0.52f;
    hq[3] = 0.54f;
  } else {
    for (int i = 0; i < 4; i++) hq[i] = normal[i];
  }
  float t = clampf_((distance - 0.1f) / 0.9f, 0.0f, 1.0f);
  for (int i = 0; i < 4; i++) thr[i] = hq[i] + t * (normal[i] - hq[i]);
}

void jxo_quantize_block_ac(const float *in, int c, const float *qm, int32_t quant, float scale,
                           float qm_mult, float distance, int xsize, int ysize, int32_t *out) {
  float thr[4];
  jxo_ac_thresholds(c, xsize, ysize, distance, qm_mult, thr);
  float q_scaled = scale * (float)quant * qm_mult;
  int width = xsize * 8, height = ysize * 8, half = width / 2;
  for (int y = 0; y < height; y++) {
    int yfix = y >= height / 2 ? 2 : 0;
    for (int x = 0; x < width; x++) {
      float threshold = x >= half ? thr[yfix + 1] : thr[yfix];
      int i = y * width + x;
      float q = qm[i] * q_scaled;
      float val = q * in[i];
      out[i] = fabsf(val) >= threshold ? jxo_sat_i32(roundf(val)) : 0;
    }
  }
}

/* DequantBias::level (src/group.rs:770-790); numerator / a is evaluated
 * directly, which is the same correctly rounded quotient the reference
 * tabulates for a < 64. */
float jxo_dequant_level(float one, float numerator, int32_t quant) {
  uint32_t a = quant < 0 ? (uint32_t)(-(int64_t)quant) : (uint32_t)quant;
  if (a <= 1) return quant == 0 ? 0.0f : (quant > 0 ? one : -one);
  float q = (float)quant;
  if (a < 64) {
    float t = numerator / (float)a;
    return q - (quant < 0 ? -t : t);
  }
  return q - numerator / q;
}

static const float INV_DC_QUANT[3] = {4096.0f, 512.0f, 256.0f};
static const float DC_QUANT3[3] = {1.0f / 4096.0f, 1.0f / 512.0f, 1.0f / 256.0f};
static const float B_QM_MULS[6] = {1.0f, 1.25f, 1.5625f, 1.953125f, 2.4414062f, 3.0517578f};

float jxo_b_qm_mul(uint32_t b_qm_scale) {
  uint32_t s = b_qm_scale < 2 ? 2 : (b_qm_scale > 7 ? 7 : b_qm_scale);
  return B_QM_MULS[s - 2];
}

float jxo_x_qm_mul(uint32_t x_qm_scale) { return powf(1.25f, (float)x_qm_scale - 2.0f); }

void jxo_forward_quantize_rows(const float *xyb, int width, int height, const uint8_t *cells,
                               const uint8_t *qf, const int8_t *ytox, const int8_t *ytob,
                               const float *tables, const jxg_frame_params *p, int32_t *levels[3],
                               float *coeffs[3], int16_t *dc[3], int by0, int by1) {
  int xb = (width + 7) / 8;
  int tiles_x = (width + 63) / 64;
  size_t plane = (size_t)width * height;
  float inv_factor[3];
  for (int c = 0; c < 3; c++) inv_factor[c] = INV_DC_QUANT[c] / p->dc_step[c] * p->scale_dc;
  float inv_cf = 1.0f / p->cfl_color_factor;
  /* dc_cfl_factor / dc_cfl_factor_x (src/color_correlation.rs:1598-1610) */
  float cfl_factor_b = INV_DC_QUANT[2] / p->dc_step[2] * (DC_QUANT3[1] * p->dc_step[1]) *
                       (p->cfl_base_b + (float)p->ytob_dc / p->cfl_color_factor);
  float cfl_factor_x =
      INV_DC_QUANT[0] / p->dc_step[0] * (DC_QUANT3[1] * p->dc_step[1]) * p->cfl_base_x;
  float x_qm_mul = jxo_x_qm_mul(p->x_qm_scale);
  float b_qm_mul = jxo_b_qm_mul(p->b_qm_scale);
  static __thread float co[3][64 * 64];
  static __thread int32_t qz[3][64 * 64];
  for (int by = by0; by < by1; by++) {
    for (int bx = 0; bx < xb; bx++) {
      uint8_t cell = cells[(size_t)by * xb + bx];
      if (!(cell & 1)) continue;
      int s = cell >> 1;
      int cov_x = COV_X[s], cov_y = COV_Y[s];
      int cx = cov_x, cy = cov_y;
      if (cov_y > cov_x) {
        cx = cov_y;
        cy = cov_x;
      }
      int size = cx * cy * 64;
      int32_t quant = qf[(size_t)by * xb + bx];
      for (int c = 0; c < 3; c++)
        jxo_forward_transform(s, xyb + c * plane, width, height, bx * 8, by * 8, co[c]);
      int kind = jxo_qt_kind(s);
      uint32_t n = jxo_qt_size(kind), off = jxo_qt_offset(kind);
      const float *qm_y = tables + off + 1 * n;
      const float *inv_qm[3] = {tables + off + 3 * n, tables + off + 4 * n, tables + off + 5 * n};
      float dcv[256];
      jxo_dc_from_coefficients(s, co[1], dcv);
      int covered = cov_x * cov_y;
      int16_t ydcq[256];
      for (int i = 0; i < covered; i++) ydcq[i] = jxo_sat_i16(roundf(dcv[i] * inv_factor[1]));
      /* Y source coefficients (the pass-2 RDOQ input) */
      for (int j = 0; j < cov_y; j++)
        memcpy(coeffs[1] + ((size_t)(by + j) * xb + bx) * 64, co[1] + j * cov_x * 64,
               cov_x * 64 * sizeof(float));
      /* Y quantize / dequantize round trip (quantize_roundtrip_y_block :795) */
      jxo_quantize_block_ac(co[1], 1, inv_qm[1], quant, p->scale, 1.0f, p->distance, cx, cy,
                            qz[1]);
      float inv_qac = 1.0f / (p->scale * (float)quant);
      for (int i = 0; i < size; i++)
        co[1][i] = jxo_dequant_level(p->dequant_one[1], p->dequant_numerator, qz[1][i]) *
                   qm_y[i] * inv_qac;
      /* zero Y's LLF (:1852-1860) */
      int wc = cx * 8;
      for (int r = 0; r < cy; r++)
        for (int k = 0; k < cx; k++) co[1][r * wc + k] = 0.0f;
      /* CfL (apply_cfl_scalar, src/ac_strategy.rs:1076) with the tile's
       * factors (y_to_x_ratio / y_to_b_ratio, src/color_correlation.rs:1117) */
      int tx = bx / 8, ty = by / 8;
      float x_factor = jxo_fmla((float)ytox[ty * tiles_x + tx], inv_cf, p->cfl_base_x);
      float b_factor = jxo_fmla((float)ytob[ty * tiles_x + tx], inv_cf, p->cfl_base_b);
      float ncx = -x_factor, ncb = -b_factor;
      for (int i = 0; i < size; i++) {
        co[0][i] = jxo_fmla(ncx, co[1][i], co[0][i]);
        co[2][i] = jxo_fmla(ncb, co[1][i], co[2][i]);
      }
      float xdc[256], bdc[256];
      jxo_dc_from_coefficients(s, co[0], xdc);
      jxo_dc_from_coefficients(s, co[2], bdc);
      int16_t xdcq[256], bdcq[256];
      for (int i = 0; i < covered; i++) {
        if (cfl_factor_x != 0.0f)
          xdcq[i] = jxo_sat_i16(
              roundf(jxo_fmla(xdc[i], inv_factor[0], -(float)ydcq[i] * cfl_factor_x)));
        else
          xdcq[i] = jxo_sat_i16(roundf(xdc[i] * inv_factor[0]));
        bdcq[i] =
            jxo_sat_i16(roundf(jxo_fmla(bdc[i], inv_factor[2], -(float)ydcq[i] * cfl_factor_b)));
      }
      jxo_quantize_block_ac(co[0], 0, inv_qm[0], quant, p->scale, x_qm_mul, p->distance, cx, cy,
                            qz[0]);
      jxo_quantize_block_ac(co[2], 2, inv_qm[2], quant, p->scale, b_qm_mul, p->distance, cx, cy,
                            qz[2]);
      for (int j = 0; j < cov_y; j++) {
        size_t slot = ((size_t)(by + j) * xb + bx) * 64;
        for (int c = 0; c < 3; c++)
          memcpy(levels[c] + slot, qz[c] + j * cov_x * 64, cov_x * 64 * sizeof(int32_t));
        memcpy(coeffs[0] + slot, co[0] + j * cov_x * 64, cov_x * 64 * sizeof(float));
        memcpy(coeffs[2] + slot, co[2] + j * cov_x * 64, cov_x * 64 * sizeof(float));
        for (int i = 0; i < cov_x; i++) {
          size_t d = (size_t)(by + j) * xb + bx + i;
          dc[0][d] = xdcq[j * cov_x + i];
          dc[1][d] = ydcq[j * cov_x + i];
          dc[2][d] = bdcq[j * cov_x + i];
        }
      }
    }
  }
}

void jxo_forward_quantize(const float *xyb, int width, int height, const uint8_t *cells,
                          const uint8_t *qf, const int8_t *ytox, const int8_t *ytob,
                          const float *tables, const jxg_frame_params *p, int32_t *levels[3],
                          float *coeffs[3], int16_t *dc[3]) {
  jxo_forward_quantize_rows(xyb, width, height, cells, qf, ytox, ytob, tables, p, levels, coeffs,
                            dc, 0, (height + 7) / 8);
}
