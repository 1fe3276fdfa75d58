// jxg_sub8.h - the single-block transforms below 8x8 resolution: IDENTITY,
// DCT2X2, DCT4X4, DCT4X8, DCT8X4 and AFV0-3 (src/dct.rs:185-330,1061-1340,
// src/afv.rs:207-256), each computed by one thread from an 8x8 pixel block
// into the reference's natural 8x8 coefficient order (k = row * 8 + col).
//
// Every loop has compile-time bounds and indices (no dynamic register
// indexing); inputs are read from an LDS copy of the block and outputs go
// straight to the block's LDS slots, which keeps register use low.
#ifndef JXG_SUB8_H_
#define JXG_SUB8_H_

#include "jxg_device.h"

namespace jxg {

// The JPEG XL AFV basis (row j = pixel j, column i = basis vector i); the
// format's table, the same values as src/afv.rs:48, in f32.
static __constant__ float AFV_BASIS[256] = {
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

__host__ __device__ __forceinline__ bool is_sub8(int s) {
  return s == JXG_IDENTITY || s == JXG_DCT2X2 || s == JXG_DCT4X4 || s == JXG_DCT4X8 ||
         s == JXG_DCT8X4 || (s >= JXG_AFV0 && s <= JXG_AFV3);
}

// dct4x4_2d (src/dct.rs:1061-1076): rows, then columns, out[hf*4 + vf] / 16.
template <bool F>
__device__ __forceinline__ void dct4x4_2d(const float *in, float *out) {
  float t[16];
#pragma unroll
  for (int i = 0; i < 16; i++) t[i] = in[i];
#pragma unroll
  for (int r = 0; r < 4; r++) dct4<F>(t + 4 * r);
#pragma unroll
  for (int x = 0; x < 4; x++) {
    float col[4] = {t[x], t[4 + x], t[8 + x], t[12 + x]};
    dct4<F>(col);
#pragma unroll
    for (int i = 0; i < 4; i++) out[x * 4 + i] = col[i] * (1.0f / 16.0f);
  }
}

// dct4x8_2d (src/dct.rs:1198-1220): 4 tall x 8 wide, columns then rows,
// out[vf*8 + hf] / 32.
template <bool F>
__device__ __forceinline__ void dct4x8_2d(const float *in, float *out) {
  float t[32];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    float col[4] = {in[c], in[8 + c], in[16 + c], in[24 + c]};
    dct4<F>(col);
#pragma unroll
    for (int vf = 0; vf < 4; vf++) t[vf * 8 + c] = col[vf];
  }
#pragma unroll
  for (int vf = 0; vf < 4; vf++) {
    dct8<F>(t + vf * 8);
#pragma unroll
    for (int i = 0; i < 8; i++) out[vf * 8 + i] = t[vf * 8 + i] * (1.0f / 32.0f);
  }
}

// dct8x4_2d (src/dct.rs:1267-1300): 8 tall x 4 wide, columns then rows,
// out[hf*8 + vf] / 32.
template <bool F>
__device__ __forceinline__ void dct8x4_2d(const float *in, float *out) {
  float t[32];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    float col[8];
#pragma unroll
    for (int i = 0; i < 8; i++) col[i] = in[i * 4 + c];
    dct8<F>(col);
#pragma unroll
    for (int vf = 0; vf < 8; vf++) t[vf * 4 + c] = col[vf];
  }
#pragma unroll
  for (int vf = 0; vf < 8; vf++) {
    float row[4] = {t[vf * 4], t[vf * 4 + 1], t[vf * 4 + 2], t[vf * 4 + 3]};
    dct4<F>(row);
#pragma unroll
    for (int hf = 0; hf < 4; hf++) out[hf * 8 + vf] = row[hf] * (1.0f / 32.0f);
  }
}

// Output of a sub-8x8 transform: coefficient k goes straight to the block's
// LDS slot (k >> 3, k & 7), so only the 64 input pixels live in registers.
struct Sub8Out {
  float *blk;
  int pitch;
  __device__ __forceinline__ float &operator[](int k) const {
    return blk[(k >> 3) * pitch + (k & 7)];
  }
};

__device__ __forceinline__ void hadamard2x2_dc(const Sub8Out &o) {
  const float b00 = o[0], b01 = o[1], b10 = o[8], b11 = o[9];
  o[0] = (b00 + b01 + b10 + b11) * 0.25f;
  o[1] = (b00 + b01 - b10 - b11) * 0.25f;
  o[8] = (b00 - b01 + b10 - b11) * 0.25f;
  o[9] = (b00 - b01 - b10 + b11) * 0.25f;
}

// identity8x8 (src/dct.rs:185-214).
__device__ __forceinline__ void identity8x8(const float *p, const Sub8Out &o) {
#pragma unroll
  for (int qy = 0; qy < 2; qy++)
#pragma unroll
    for (int qx = 0; qx < 2; qx++) {
      const float anchor = p[(qy * 4 + 1) * 8 + qx * 4 + 1];
      float mean = 0.0f;
#pragma unroll
      for (int iy = 0; iy < 4; iy++)
#pragma unroll
        for (int ix = 0; ix < 4; ix++) {
          const float v = p[(qy * 4 + iy) * 8 + qx * 4 + ix];
          mean += v;
          if (ix == 1 && iy == 1) continue;
          o[(qy + iy * 2) * 8 + qx + ix * 2] = v - anchor;
        }
      // the displaced top-left residual moves to the unused anchor slot
      o[(qy + 2) * 8 + qx + 2] = p[(qy * 4) * 8 + qx * 4] - anchor;
      o[qy * 8 + qx] = mean * (1.0f / 16.0f);
    }
  hadamard2x2_dc(o);
}

// dct2x2_8x8 (src/dct.rs:270-306): three levels of 2x2 Haar steps
// (dct2_top_block at S = 8, 4, 2), each level's averages feeding the next.
__device__ __forceinline__ void dct2x2_8x8(const float *p, const Sub8Out &o) {
  float a4[16], a2[4];
#pragma unroll
  for (int y = 0; y < 4; y++)
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const float c00 = p[(2 * y) * 8 + 2 * x], c01 = p[(2 * y) * 8 + 2 * x + 1];
      const float c10 = p[(2 * y + 1) * 8 + 2 * x], c11 = p[(2 * y + 1) * 8 + 2 * x + 1];
      a4[y * 4 + x] = (c00 + c01 + c10 + c11) * 0.25f;
      o[y * 8 + 4 + x] = (c00 + c01 - c10 - c11) * 0.25f;
      o[(y + 4) * 8 + x] = (c00 - c01 + c10 - c11) * 0.25f;
      o[(y + 4) * 8 + 4 + x] = (c00 - c01 - c10 + c11) * 0.25f;
    }
#pragma unroll
  for (int y = 0; y < 2; y++)
#pragma unroll
    for (int x = 0; x < 2; x++) {
      const float c00 = a4[(2 * y) * 4 + 2 * x], c01 = a4[(2 * y) * 4 + 2 * x + 1];
      const float c10 = a4[(2 * y + 1) * 4 + 2 * x], c11 = a4[(2 * y + 1) * 4 + 2 * x + 1];
      a2[y * 2 + x] = (c00 + c01 + c10 + c11) * 0.25f;
      o[y * 8 + 2 + x] = (c00 + c01 - c10 - c11) * 0.25f;
      o[(y + 2) * 8 + x] = (c00 - c01 + c10 - c11) * 0.25f;
      o[(y + 2) * 8 + 2 + x] = (c00 - c01 - c10 + c11) * 0.25f;
    }
  const float c00 = a2[0], c01 = a2[1], c10 = a2[2], c11 = a2[3];
  o[0] = (c00 + c01 + c10 + c11) * 0.25f;
  o[1] = (c00 + c01 - c10 - c11) * 0.25f;
  o[8] = (c00 - c01 + c10 - c11) * 0.25f;
  o[9] = (c00 - c01 - c10 + c11) * 0.25f;
}

// dct4x4_scalar_input (src/dct.rs:1150-1180).
template <bool F>
__device__ __forceinline__ void dct4x4_8x8(const float *p, const Sub8Out &o) {
#pragma unroll
  for (int qy = 0; qy < 2; qy++)
#pragma unroll
    for (int qx = 0; qx < 2; qx++) {
      float b[16], d[16];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) b[r * 4 + c] = p[(qy * 4 + r) * 8 + qx * 4 + c];
      dct4x4_2d<F>(b, d);
#pragma unroll
      for (int iy = 0; iy < 4; iy++)
#pragma unroll
        for (int ix = 0; ix < 4; ix++) o[(qy + iy * 2) * 8 + qx + ix * 2] = d[iy * 4 + ix];
    }
  hadamard2x2_dc(o);
}

// dct4x8_scalar_input (src/dct.rs:1230-1255).
template <bool F>
__device__ __forceinline__ void dct4x8_8x8(const float *p, const Sub8Out &o) {
#pragma unroll
  for (int y = 0; y < 2; y++) {
    float d[32];
    dct4x8_2d<F>(p + y * 32, d);
#pragma unroll
    for (int iy = 0; iy < 4; iy++)
#pragma unroll
      for (int ix = 0; ix < 8; ix++) o[(y + iy * 2) * 8 + ix] = d[iy * 8 + ix];
  }
  const float b0 = o[0], b1 = o[8];
  o[0] = (b0 + b1) * 0.5f;
  o[8] = (b0 - b1) * 0.5f;
}

// dct8x4_scalar_input (src/dct.rs:1310-1340).
template <bool F>
__device__ __forceinline__ void dct8x4_8x8(const float *p, const Sub8Out &o) {
#pragma unroll
  for (int x = 0; x < 2; x++) {
    float h[32], d[32];
#pragma unroll
    for (int r = 0; r < 8; r++)
#pragma unroll
      for (int c = 0; c < 4; c++) h[r * 4 + c] = p[r * 8 + x * 4 + c];
    dct8x4_2d<F>(h, d);
#pragma unroll
    for (int iy = 0; iy < 4; iy++)
#pragma unroll
      for (int ix = 0; ix < 8; ix++) o[(x + iy * 2) * 8 + ix] = d[iy * 8 + ix];
  }
  const float b0 = o[0], b1 = o[8];
  o[0] = (b0 + b1) * 0.5f;
  o[8] = (b0 - b1) * 0.5f;
}

// afv_from_pixels (src/afv.rs:207-256) for kind AX + 2 AY.
template <bool F, int AX, int AY>
__device__ __forceinline__ void afv_8x8(const float *p, const Sub8Out &o) {
  float q[16], d[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int dy = i / 4, dx = i % 4;
    const int sy = AY ? 3 - dy : dy, sx = AX ? 3 - dx : dx;
    q[i] = p[(sy + 4 * AY) * 8 + sx + 4 * AX];
  }
  // afv_dct4x4 (src/afv.rs:216-225): coefficient i accumulates pixels in order.
#pragma unroll
  for (int i = 0; i < 16; i++) {
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; j++) acc = fmla<F>(q[j], AFV_BASIS[j * 16 + i], acc);
    o[(i / 4) * 16 + (i % 4) * 2] = acc;
  }
  constexpr int QX = AX ? 0 : 4;
#pragma unroll
  for (int i = 0; i < 16; i++) q[i] = p[(i / 4 + 4 * AY) * 8 + QX + i % 4];
  dct4x4_2d<F>(q, d);
#pragma unroll
  for (int iy = 0; iy < 4; iy++)
#pragma unroll
    for (int ix = 0; ix < 4; ix++) o[iy * 16 + ix * 2 + 1] = d[iy * 4 + ix];
  constexpr int HY = AY ? 0 : 4;
  float h[32];
  dct4x8_2d<F>(p + HY * 8, h);
#pragma unroll
  for (int iy = 0; iy < 4; iy++)
#pragma unroll
    for (int ix = 0; ix < 8; ix++) o[(1 + iy * 2) * 8 + ix] = h[iy * 8 + ix];
  const float b00 = o[0] * 0.25f, b01 = o[1], b10 = o[8];
  o[0] = (b00 + b01 + 2.0f * b10) * 0.25f;
  o[1] = (b00 - b01) * 0.5f;
  o[8] = (b00 + b01 - 2.0f * b10) * 0.25f;
}

// Transforms the 8x8 pixel block at `blk` (LDS, row pitch `pitch`) in place:
// afterwards blk[(k >> 3) * pitch + (k & 7)] holds coefficient k.  The
// pixels are first copied to `px` (64 floats of LDS owned by the caller),
// which every sub-transform reads while it writes its outputs into `blk`.
template <bool F>
__device__ __noinline__ void sub8_transform(int s, float *blk, int pitch, float *px) {
#pragma unroll
  for (int i = 0; i < 64; i++) px[i] = blk[(i >> 3) * pitch + (i & 7)];
  const float *p = px;
  const Sub8Out o{blk, pitch};
  switch (s) {
    case JXG_IDENTITY: identity8x8(p, o); break;
    case JXG_DCT2X2: dct2x2_8x8(p, o); break;
    case JXG_DCT4X4: dct4x4_8x8<F>(p, o); break;
    case JXG_DCT4X8: dct4x8_8x8<F>(p, o); break;
    case JXG_DCT8X4: dct8x4_8x8<F>(p, o); break;
    case JXG_AFV0: afv_8x8<F, 0, 0>(p, o); break;
    case JXG_AFV1: afv_8x8<F, 1, 0>(p, o); break;
    case JXG_AFV2: afv_8x8<F, 0, 1>(p, o); break;
    default: afv_8x8<F, 1, 1>(p, o); break;
  }
}

}  // namespace jxg

#endif  // JXG_SUB8_H_
