// jxg_kernels.hip - gfx950 kernels of the VarDCT encoder front end and the
// C ABI declared in include/jxg.h.
//
// Stage kernels (one launch per stage per image):
//   k_xyb          u8 sRGB -> LUT linear -> opsin mix -> cube root -> XYB.
//                  One thread per output pixel of the padded plane; padding
//                  is edge replication, so later stages need no bounds checks.
//   k_aq_pre       per 4x4 pixel cell: luma masking pre-pass, 4x4 average.
//   k_aq_block     per 8x8 block: fuzzy erosion of the pre map (clamped to the
//                  block's 2048-px DC group) and the per-block modulations.
//   k_select       one wavefront per 32x32-px region: DCT8 / DCT16X16 /
//                  DCT32X32 candidate costs (transform, CfL, quantize, SSE and
//                  rate in LDS, sums in the reference's sequential order), the
//                  square-merge decisions and the quant-field aggregation.
//   k_quantize     one 256-thread workgroup per 64x64-px tile: every transform
//                  whose footprint lies in the tile is transformed in LDS and
//                  quantized; transforms that cross the tile are queued.
//   k_quantize_xt  persistent pass over the queued tile-crossing transforms.
//
// All arithmetic follows the reference's scalar operation order; see
// jxg_device.h.  No float atomics are used, so results do not depend on the
// launch shape.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "jxg_batch.h"
#include "jxg_ctx.h"
#include "jxg_device.h"
#include "jxg_internal.h"
#include "jxg_sub8.h"

using namespace jxg;

// ---------------------------------------------------------------------------
// errors

static thread_local char g_err[512] = "";

int set_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

// ---------------------------------------------------------------------------
// XYB

// A workgroup covers XYB_PX_WG x XYB_ROWS pixels: each thread four
// consecutive pixels of XYB_ROWS rows, read as ch aligned dwords and stored
// as 16-byte vectors.
#define XYB_PIX 4
#define XYB_PX_WG (256 * XYB_PIX)
#ifndef XYB_ROWS
#define XYB_ROWS 1
#endif

// One pixel's linear r, g, b (LUT values): X, Y, B.
template <bool F, int ORDER, bool GRAY>
__device__ __forceinline__ void xyb_pixel(float r, float g, float b, const float *m,
                                          float &Xo, float &Yo, float &Bo) {
  const float bias = 0.0037930734f, neg = -0.1559542f;
  float tm[3];
#pragma unroll
  for (int row = 0; row < 3; row++) {
    const float m0 = m[3 * row], m1 = m[3 * row + 1], m2 = m[3 * row + 2];
    if constexpr (ORDER == JXG_XYB_AVX2) {
      // src/avx/xyb.rs:234-262 (SCALAR_ORDER = false, FUSED = true)
      float mixed = __builtin_fmaf(b, m2, bias);
      mixed = __builtin_fmaf(g, m1, mixed);
      mixed = __builtin_fmaf(r, m0, mixed);
      tm[row] = cbrt_avx2_positive(fmaxf(mixed, 0.0f)) + neg;
    } else {
      // src/xyb.rs:172-185
      float mixed = fmla<F>(m0, r, fmla<F>(m1, g, fmla<F>(m2, b, bias)));
      tm[row] = cbrt_scalar<F>(fmaxf(mixed, 0.0f)) + neg;
    }
  }
  float X, Y;
  if constexpr (ORDER == JXG_XYB_AVX2) {
    X = (tm[0] - tm[1]) * 0.5f;
    Y = (tm[0] + tm[1]) * 0.5f;
  } else {
    X = 0.5f * (tm[0] - tm[1]);
    Y = 0.5f * (tm[0] + tm[1]);
  }
  // snap_achromatic_xyb (src/frame/encoder.rs:347-357): X = 0, B = Y.
  Xo = GRAY ? 0.0f : X;
  Yo = Y;
  Bo = GRAY ? Y : tm[2];
}

// The XYB_PIX pixels of one thread on row y, CH bytes per pixel.
template <bool F, int ORDER, int CH>
__device__ __forceinline__ void xyb_run(const uint8_t *__restrict__ rgb, size_t row_bytes,
                                        int W, int H, const float *lut, const float *m,
                                        float *__restrict__ xyb, size_t plane, int stride,
                                        int x0, int y, bool vec) {
  // Gray (+ alpha) input decodes one sample to all three channels
  // (linearize_gray, src/encode_image.rs:1571-1599).
  constexpr bool GRAY = CH <= 2;
  const uint8_t *row = rgb + (size_t)min(y, H - 1) * row_bytes;
  const uint8_t *p = row + (size_t)x0 * CH;
  uint8_t px[XYB_PIX * CH];
  if (x0 + XYB_PIX <= W && ((uintptr_t)p & 3) == 0) {
    // XYB_PIX * CH bytes = CH dwords
    uint32_t w[CH];
#pragma unroll
    for (int i = 0; i < CH; i++) w[i] = reinterpret_cast<const uint32_t *>(p)[i];
#pragma unroll
    for (int i = 0; i < XYB_PIX * CH; i++) px[i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
  } else {
#pragma unroll
    for (int k = 0; k < XYB_PIX; k++) {
      const uint8_t *q = row + (size_t)min(x0 + k, W - 1) * CH;
#pragma unroll
      for (int c = 0; c < CH; c++) px[k * CH + c] = q[c];
    }
  }
  float Xo[XYB_PIX], Yo[XYB_PIX], Bo[XYB_PIX];
#pragma unroll
  for (int k = 0; k < XYB_PIX; k++) {
    const float r = lut[px[k * CH]];
    const float g = GRAY ? r : lut[px[k * CH + 1]], b = GRAY ? r : lut[px[k * CH + 2]];
    xyb_pixel<F, ORDER, GRAY>(r, g, b, m, Xo[k], Yo[k], Bo[k]);
  }
  float *o = xyb + (size_t)y * stride + x0;
  if (vec) {
    *reinterpret_cast<float4 *>(o) = make_float4(Xo[0], Xo[1], Xo[2], Xo[3]);
    *reinterpret_cast<float4 *>(o + plane) = make_float4(Yo[0], Yo[1], Yo[2], Yo[3]);
    *reinterpret_cast<float4 *>(o + 2 * plane) = make_float4(Bo[0], Bo[1], Bo[2], Bo[3]);
  } else {
#pragma unroll
    for (int k = 0; k < XYB_PIX; k++) {
      o[k] = Xo[k];
      o[plane + k] = Yo[k];
      o[2 * plane + k] = Bo[k];
    }
  }
}

template <bool F, int ORDER>
__device__ __forceinline__ void xyb_body(const uint8_t *__restrict__ rgb, size_t row_bytes,
                                             int ch, int W, int H,
                                             const float *__restrict__ mlut,
                                             float *__restrict__ xyb, int stride, int padded_h, int bix, int biy) {
  __shared__ float lut[256];
  __shared__ float m[9];
  const int t = threadIdx.x;
  lut[t] = mlut[9 + t];
  if (t < 9) m[t] = mlut[t];
  __syncthreads();
  const int x0 = bix * XYB_PX_WG + t * XYB_PIX;  // stride is a multiple of 64
  if (x0 >= stride) return;
  const size_t plane = (size_t)stride * padded_h;
  const bool vec = ((uintptr_t)xyb & 15) == 0;  // planes a caller passes may not be
  for (int r = 0; r < XYB_ROWS; r++) {
    const int y = biy * XYB_ROWS + r;  // padded_h is a multiple of 64
    if (y >= padded_h) break;
    switch (ch) {
      case 1: xyb_run<F, ORDER, 1>(rgb, row_bytes, W, H, lut, m, xyb, plane, stride, x0, y, vec); break;
      case 2: xyb_run<F, ORDER, 2>(rgb, row_bytes, W, H, lut, m, xyb, plane, stride, x0, y, vec); break;
      case 3: xyb_run<F, ORDER, 3>(rgb, row_bytes, W, H, lut, m, xyb, plane, stride, x0, y, vec); break;
      default: xyb_run<F, ORDER, 4>(rgb, row_bytes, W, H, lut, m, xyb, plane, stride, x0, y, vec); break;
    }
  }
}

template <bool F, int ORDER>
__global__ __launch_bounds__(256) void k_xyb(const uint8_t *__restrict__ rgb, size_t row_bytes,
                                             int ch, int W, int H,
                                             const float *__restrict__ mlut,
                                             float *__restrict__ xyb, int stride, int padded_h) {
  xyb_body<F, ORDER>(rgb, row_bytes, ch, W, H, mlut, xyb, stride, padded_h, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Adaptive quantization (src/adaptive_quant.rs:41-206,410-470,547-729)

template <bool INVERT>
__device__ __forceinline__ float ratio_gamma(float v) {
  const float k_sg_mul = 226.77216153508914f;
  const float k_sg_mul2 = 1.0f / 73.377132366608819f;
  const float k_log2 = 0.693147181f;
  const float k_ret_mul = k_sg_mul2 * 18.6580932135f * k_log2;
  const float k_eps = 1e-2f;
  if (v < 0.0f) v = 0.0f;
  const float k_num_mul = k_ret_mul * 3.0f * k_sg_mul;
  const float k_v_offset = 7.7825991679894591f * k_log2 + k_eps;
  const float k_den_mul = k_log2 * k_sg_mul;
  float v2 = v * v;
  float num = k_num_mul * v2 + k_eps;
  float den = (k_den_mul * v) * v2 + k_v_offset;
  return INVERT ? num / den : den / num;
}

template <bool F>
__device__ __forceinline__ float masking_sqrt(float v) {
  const float mul_v = 211.66567973503678f * 1e8f;
  return 0.25f * sqrtf(fmla<F>(v, sqrtf(mul_v), 27.505837037000106f));
}

template <bool F>
__device__ __forceinline__ void aq_pre_body(const float *__restrict__ yp, int stride, int W,
                                                int H, int pre_w, int pre_h,
                                                float *__restrict__ pre, int bix, int biy) {
  // Lane = pixel column gx, so a wavefront's loads are consecutive floats;
  // the four lanes of output i = gx / 4 (which exit together) each run one
  // column's sum over the four rows, and lane 4 i combines them in the
  // order (((c0 + c1) + c2) + c3).
  const int gx = bix * 256 + threadIdx.x;
  const int i = gx >> 2, j = biy;
  if (i >= pre_w || j >= pre_h) return;
  const int gxc = min(gx, W - 1), gx1 = min(max(gx - 1, 0), W - 1), gx2 = min(gx + 1, W - 1);
  float col[6];  // column gxc, rows 4 j - 1 .. 4 j + 4 clamped into the image
#pragma unroll
  for (int k = 0; k < 6; k++) col[k] = yp[(size_t)min(max(4 * j - 1 + k, 0), H - 1) * stride + gxc];
  float lft[4], rgt[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const float *row = yp + (size_t)min(4 * j + r, H - 1) * stride;
    lft[r] = row[gx1];
    rgt[r] = row[gx2];
  }
  float acc = 0.0f;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    // rows gy - 1, gy, gy + 1 (clamped) are col[r], col[r + 1], col[r + 2]
    const float in_y = col[r + 1];
    const float base = 0.25f * (col[r + 2] + col[r] + lft[r] + rgt[r]);
    const float gammac = ratio_gamma<false>(in_y + 0.019f);
    float diff = gammac * (in_y - base);
    diff *= diff;
    if (diff >= 0.2f) diff = 0.2f;
    diff = masking_sqrt<F>(diff);
    acc = r == 0 ? diff : acc + diff;
  }
  const float c1 = __shfl_down(acc, 1, 64), c2 = __shfl_down(acc, 2, 64), c3 = __shfl_down(acc, 3, 64);
  if ((gx & 3) == 0) pre[(size_t)j * pre_w + i] = (((acc + c1) + c2) + c3) * 0.25f;
}

template <bool F>
__global__ __launch_bounds__(256) void k_aq_pre(const float *__restrict__ yp, int stride, int W,
                                                int H, int pre_w, int pre_h,
                                                float *__restrict__ pre) {
  aq_pre_body<F>(yp, stride, W, H, pre_w, pre_h, pre, blockIdx.x, blockIdx.y);
}

__device__ __forceinline__ void sort4(float *a) {
#define JXG_CSWAP(i, j)    \
  if (a[i] > a[j]) {       \
    float t_ = a[i];       \
    a[i] = a[j];           \
    a[j] = t_;             \
  }
  JXG_CSWAP(0, 1);
  JXG_CSWAP(2, 3);
  JXG_CSWAP(0, 2);
  JXG_CSWAP(1, 3);
  JXG_CSWAP(1, 2);
#undef JXG_CSWAP
}

// Inserts v into the ascending m[0..3], keeping the four smallest (v goes
// before the first element it is strictly below).  Selects on the same
// comparisons: for an ascending m, v < m[i] implies v < m[i + 1].
__device__ __forceinline__ void store_min4(float v, float *m) {
  const bool c0 = v < m[0], c1 = v < m[1], c2 = v < m[2], c3 = v < m[3];
  const float n3 = c3 ? (c2 ? m[2] : v) : m[3];
  const float n2 = c2 ? (c1 ? m[1] : v) : m[2];
  const float n1 = c1 ? (c0 ? m[0] : v) : m[1];
  const float n0 = c0 ? v : m[0];
  m[0] = n0;
  m[1] = n1;
  m[2] = n2;
  m[3] = n3;
}

// Per-block modulations (fill_quant_field_scalar stage 3,
// src/adaptive_quant.rs:692-729) for a strip of AQ_STRIP blocks of one block
// row.  Map, then fold: every pixel's gamma / HF / blue term is computed in
// parallel (rows across the strip, so loads coalesce), then one lane per
// (block, modulation) folds the terms in the reference's loop order, and one
// lane per block finishes the block.  The per-pixel terms and every sum keep
// the reference's f32 operation order.
#define AQ_STRIP 32

template <bool F, int STRIP>
__device__ __forceinline__ void aq_block_body(const float *__restrict__ xyb, int stride,
                                                  int padded_h, int W, int H,
                                                  const float *__restrict__ pre, int pre_w,
                                                  float distance, float inv_scale,
                                                  const float *__restrict__ exp2_table,
                                                  uint8_t *__restrict__ qf, int xb, int yb, int bix, int biy) {
  __shared__ float tg[8][STRIP * 8];  // gamma pair sum, [dy][block * 8 + dx]
  __shared__ float thr_[8][STRIP * 8];  // HF term to the right
  __shared__ float thd[8][STRIP * 8];   // HF term downwards
  __shared__ float tbl[8][STRIP * 8];   // blue term
  __shared__ float sums[3][STRIP];
  const int tid = threadIdx.x;
  const int bx0 = bix * STRIP, by = biy;
  const size_t plane = (size_t)stride * padded_h;
  const float *px_ = xyb, *py_ = xyb + plane, *pb_ = xyb + 2 * plane;
  const int y = min(by * 8, max(H - 8, 0));
  // ---- map: thread = (block, dx) of the strip, one row dy at a time ----
  {
    const int b = tid >> 3, dx = tid & 7, bx = bx0 + b;
    if (bx < xb) {
      const int x = min(bx * 8, max(W - 8, 0));
      const int cx = min(x + dx, W - 1), cr = min(x + dx + 1, W - 1);
      const float k_limit = 0.010474084867598155f;
#pragma unroll 2
      for (int dy = 0; dy < 8; dy++) {
        const int ry = min(y + dy, H - 1);
        const int rn = dy == 7 ? ry : min(y + dy + 1, H - 1);
        const size_t i = (size_t)ry * stride + cx;
        const float vy = py_[i], vx = px_[i], vb = pb_[i];
        // gamma_modulation (src/adaptive_quant.rs:410-435)
        const float iny = vy + 0.16f, inx = vx;
        tg[dy][tid] = ratio_gamma<true>(iny - inx) + ratio_gamma<true>(iny + inx);
        // hf_modulation (:127-160)
        thr_[dy][tid] = fminf(0.0206f, fabsf(vy - py_[(size_t)ry * stride + cr]));
        thd[dy][tid] = fminf(0.0206f, fabsf(vy - py_[(size_t)rn * stride + cx]));
        // blue_modulation (:163-206)
        const float p_y = vy + 0.0031994768654636393f + fabsf(vx);
        tbl[dy][tid] = vb > p_y ? fminf(vb - p_y, k_limit) : 0.0f;
      }
    }
  }
  __syncthreads();
  // ---- fold: lane (which, block) runs one sum in the reference's order ----
  if (tid < 3 * STRIP) {
    const int which = tid / STRIP, b = tid - which * STRIP;
    if (bx0 + b < xb) {
      float acc = 0.0f;
      if (which == 0) {
        for (int dy = 0; dy < 8; dy++)
#pragma unroll
          for (int dx = 0; dx < 8; dx++) acc = fmla<F>(0.5f, tg[dy][b * 8 + dx], acc);
      } else if (which == 1) {
        for (int dy = 0; dy < 8; dy++)
#pragma unroll
          for (int dx = 0; dx < 8; dx++) {
            if (dx < 7) acc += thr_[dy][b * 8 + dx];
            acc += thd[dy][b * 8 + dx];
          }
      } else {
        for (int dy = 0; dy < 8; dy++)
#pragma unroll
          for (int dx = 0; dx < 8; dx++) acc += tbl[dy][b * 8 + dx];
      }
      sums[which][b] = acc;
    }
  }
  __syncthreads();
  if (tid >= STRIP) return;
  const int bx = bx0 + tid;
  if (bx >= xb || by >= yb) return;
  const float d = distance;
  // DC group of this block and its pre-map extent.
  const int gx = bx >> 8, gy = by >> 8;
  const int gxb = min(256, xb - gx * 256), gyb = min(256, yb - gy * 256);
  const int pw = 2 * gxb, ph = 2 * gyb;
  const int ppx0 = gx * 512, ppy0 = gy * 512;
  // FuzzyErosion (src/adaptive_quant.rs:437-470, 640-690)
  const float fe_mul = d < 2.0f ? (2.0f - d) * 0.5f : 0.0f;
  const float fe_base[4] = {0.125f, 0.1f, 0.09f, 0.06f};
  const float fe_add[4] = {0.0f, -0.1f, -0.09f, -0.06f};
  float kmul[4];
  float norm = 0.0f;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    kmul[i] = fe_base[i] + fe_mul * fe_add[i];
    norm += kmul[i];
  }
  const float k_total = 0.29959705784054957f;
#pragma unroll
  for (int i = 0; i < 4; i++) kmul[i] *= k_total / norm;
  // The block's four quadrants (fx, fy) = (fx0 + qx, fy0 + qy) and their
  // neighbours, clamped into the group's pre-map, are exactly the 4x4 window
  // rows fy0 - 1 .. fy0 + 2, columns fx0 - 1 .. fx0 + 2 clamped the same way
  // (fx0, fy0 are even and below pw - 1, ph - 1): all sixteen loads are
  // issued together.
  const int fy0 = 2 * (by - gy * 256), fx0 = 2 * (bx - gx * 256);
  float win[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int yy = min(max(fy0 - 1 + r, 0), ph - 1);
    const float *row = pre + (size_t)(ppy0 + yy) * pre_w + ppx0;
#pragma unroll
    for (int c = 0; c < 4; c++) win[r][c] = row[min(max(fx0 - 1 + c, 0), pw - 1)];
  }
  float aq = 0.0f;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int R = 1 + (q >> 1), C = 1 + (q & 1);  // the quadrant's window position
    float mins[4] = {win[R][C], win[R][C - 1], win[R][C + 1], win[R - 1][C - 1]};
    sort4(mins);
    store_min4(win[R - 1][C], mins);
    store_min4(win[R - 1][C + 1], mins);
    store_min4(win[R + 1][C - 1], mins);
    store_min4(win[R + 1][C], mins);
    store_min4(win[R + 1][C + 1], mins);
    const float v = kmul[0] * mins[0] + kmul[1] * mins[1] + kmul[2] * mins[2] + kmul[3] * mins[3];
    aq = q == 0 ? v : aq + v;
  }
  // compute_mask (:67-84)
  float mask;
  {
    const float k_offset3 = 3.7179635626140772f;
    const float v1 = fmaxf(aq * 1.07267f, 1e-3f);
    const float v2 = 1.0f / (v1 + 302.59587815579727f);
    const float v3 = 1.0f / (v1 * v1 + k_offset3);
    const float v4 = 1.0f / (v1 * v1 + 0.25f * k_offset3);
    mask = fmla<F>(9.4708735624378946f, v4,
                   fmla<F>(17.35036561631863f, v2, fmla<F>(6.7943250517376494f, v3, -0.699765f)));
  }
  // gamma_modulation
  {
    float overall = sums[0][tid];
    overall *= 1.0f / 64.0f;
    mask = fmla<F>(0.119146f, dirty_log2f<F>(overall), mask);
  }
  // hf_modulation
  float out_val;
  {
    const float strength = fminf(fmaxf(1.06836f * d, 0.1f), 0.890398f);
    out_val = fmla<F>(sums[1][tid], -0.506626f * strength, 0.655866f) + mask;
  }
  // blue_modulation
  {
    const float k_limit = 0.010474084867598155f;
    float s = sums[2][tid];
    if (s >= 32.0f * k_limit) s = 64.0f * k_limit - s;
    const float k_max = 15.463398341612438f;
    if (s >= k_max * k_limit) s = k_max * k_limit;
    s *= 0.127191f;
    out_val = fminf(out_val, s + mask);
  }
  // quant field (fill_quant_field_scalar stage 3)
  const float scale = 0.8294f / d;
  // aq_schedule of the base level and strength
  const float knots[7] = {0.5f, 1.0f, 1.4f, 1.75f, 3.0f, 5.0f, 8.0f};
  const float strength_t[7] = {0.26565f, 0.275459f, 0.058287f, 0.0f, 0.0f, 0.0f, 0.0f};
  const float base_t[7] = {0.636493f, 0.648351f, 0.606759f, 0.48f, 0.48f, 0.48f, 0.48f};
  float base_level = base_t[6], dampen = strength_t[6];
  if (d <= knots[0]) {
    base_level = base_t[0];
    dampen = strength_t[0];
  } else {
    for (int i = 1; i < 7; i++)
      if (d <= knots[i]) {
        const float f = (d - knots[i - 1]) / (knots[i] - knots[i - 1]);
        base_level = base_t[i - 1] + f * (base_t[i] - base_t[i - 1]);
        dampen = strength_t[i - 1] + f * (strength_t[i] - strength_t[i - 1]);
        break;
      }
  }
  dampen = fminf(fmaxf(dampen, 0.0f), 1.0f);
  base_level = base_level * scale;
  const float mul = scale * dampen;
  const float add = (1.0f - dampen) * base_level;
  const float qv = fast_exp2<F>(out_val * 1.442695041f, exp2_table) * mul + add;
  int32_t qi = sat_i32(fmla<F>(qv, inv_scale, 0.5f));
  qi = qi < 1 ? 1 : (qi > 255 ? 255 : qi);
  qf[(size_t)by * xb + bx] = (uint8_t)qi;
}

template <bool F>
__global__ __launch_bounds__(256) void k_aq_block(const float *__restrict__ xyb, int stride,
                                                  int padded_h, int W, int H,
                                                  const float *__restrict__ pre, int pre_w,
                                                  float distance, float inv_scale,
                                                  const float *__restrict__ exp2_table,
                                                  uint8_t *__restrict__ qf, int xb, int yb) {
  aq_block_body<F, AQ_STRIP>(xyb, stride, padded_h, W, H, pre, pre_w, distance, inv_scale, exp2_table, qf, xb, yb, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Candidate costs and square-merge selection

// LDS row pitch of the 32x32 region: odd, so the 1-D line passes are
// conflict-free, and 37 rather than 33 because the transposed scan-order
// coefficient reads of the quantize step then spread over the banks (33
// puts up to ten lanes of a scan-order wavefront on one bank).
#define SEL_PITCH 37
#define SEL_T 256     // threads per region
#define SELQ_N (64 + 256 + 1024)  // scan-ordered inverse weights per channel
#define RD_LAMBDA 0.08086717f
#define META_R 4.0f

struct SelLds {
  float co[3][32 * SEL_PITCH];  // coefficients (spatial), then d^2 terms (storage order)
  float mg[3][1024 + 4];        // magnitude-rate terms, storage order (rows padded: the
                                // six 32x32 sums read their arrays on distinct banks)
  float quot[3][64];            // per channel: 0, 1 - one, numerator / a
  float rate[256];              // log2(1 + i)
  float sse[16][3], mag[16][3];
  uint32_t nzp[3][16], msp[3][16];  // per slot 4 j + wave: nonzeros, 1 + last nonzero scan position
  float cost8[16], cost16[4], cost32;
  double red1[4][5], red2[4][5];
  double site_s1[5], site_s2[5];
  uint32_t site_cnt[5];
  double erased[4];
  float qs[3][16];  // per channel, per site: quantizer * selection scale * qm multiplier
  float sbias[5], sgate[5];  // site biases and risk gates: four 16x16 sites, the region
  float child[4], base[4];   // per 16x16 site: cost of its choice, cost of its four DCT8s
  int pick16[4], merge32;
};

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
  return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off, 64));
  return v;
}

// DequantBias::level_f32 (src/group.rs:748-768) with the quotient table the
// reference keeps for integral levels below 64.
__device__ __forceinline__ float level_f32_tab(float one, float num, float q, const float *quot) {
  const float aq = fabsf(q);
  if (aq < 1.125f) return q == 0.0f ? 0.0f : copysignf(one, q);
  const uint32_t a = (uint32_t)aq;
  if (a < 64 && (float)a == aq) {
    const float t = quot[a];
    return q - (q < 0.0f ? -t : t);
  }
  return q - num / q;
}

__device__ __forceinline__ float rate_log2_big(const jxg_derived *D, float aq) {
  return jxg_rate_log2(D, aq);
}

__device__ __noinline__ float rate_log2_slow(const jxg_derived *D, float aq) { return rate_log2_big(D, aq); }

// Candidate evaluation over the whole region for square size n (8, 16, 32):
// all (32/n)^2 sites at once.  qac_in[site] is the site's quantizer.  The
// region's pixels must already be in L.co (spatial, pitch SEL_PITCH).
template <bool F, int LG>
__device__ void eval_squares(SelLds &L, const float *qac_in, float cmx, float cmb,
                             const float *__restrict__ tables, const uint32_t *__restrict__ scan,
                             const float *__restrict__ selq, const jxg_derived *__restrict__ D,
                             int W, int H, int px0, int py0, float *cost_out) {
  constexpr int n = 1 << LG;
  constexpr int sites_1d = 32 / n, nsites = sites_1d * sites_1d, size = n * n;
  constexpr int kind = n == 8 ? JXG_QT_DCT8 : (n == 16 ? JXG_QT_DCT16X16 : JXG_QT_DCT32X32);
  constexpr int cls = n == 8 ? 0 : (n == 16 ? 2 : 4);
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t qoff = D->qt_off[kind], qn = D->qt_n[kind];
  // 2-D transform in place: 8x8 rows first, 16x16 / 32x32 columns first
  // (src/dct.rs:631-655,1010-1030,1980-2000).
  const float scale = 1.0f / (float)(n * n);
  constexpr int nlines = 3 * nsites * n;
#pragma unroll
  for (int pass = 0; pass < 2; pass++) {
    const bool along_rows = (n == 8) ? (pass == 0) : (pass == 1);
    for (int e = tid; e < nlines; e += SEL_T) {
      const int c = e / (nsites * n);
      const int rem = e - c * nsites * n;
      const int site = rem >> LG, line = rem & (n - 1);
      const int sy = site / sites_1d, sx = site - sy * sites_1d;
      float *base = &L.co[c][(sy * n) * SEL_PITCH + sx * n];
      if (along_rows)
        dct_line<F, n>(base + line * SEL_PITCH, 1, scale, pass == 1);
      else
        dct_line<F, n>(base + line, SEL_PITCH, scale, pass == 1);
    }
    __syncthreads();
  }
  // This thread's four coefficients (element e = tid + 256 j, scan position
  // p = e mod size): storage index and the three channels' inverse weights,
  // loaded here so they arrive while CfL and the quantizer scale run.
  int kq[4];
  float ipq[3][4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int p = (tid + SEL_T * j) & (size - 1);
    kq[j] = (int)scan[p];
#pragma unroll
    for (int c = 0; c < 3; c++) ipq[c][j] = selq[c * SELQ_N + p];
  }
  // CfL against the unquantized luma (apply_cfl_scalar, src/ac_strategy.rs:1076)
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int e = tid + SEL_T * j;
    const int a = (e >> 5) * SEL_PITCH + (e & 31);
    const float yv = L.co[1][a];
    L.co[0][a] = fmla<F>(-cmx, yv, L.co[0][a]);
    L.co[2][a] = fmla<F>(-cmb, yv, L.co[2][a]);
  }
  // selection_quantizer::scale of the 32-family
  // (src/ac_strategy/selection_quantizer.rs:82-120): erased luma energy in f64.
  float qscale = 1.0f;
  if (n == 32 && D->selection_quantizer) {
    const float *inv_y = tables + qoff + 4 * qn;
    const float qac = qac_in[0];
    double energy = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int k = tid + SEL_T * j;
      const int y = k >> 5, x = k & 31;  // storage row / column (= u, v)
      if (x < 8 && y < 8) continue;
      const float l = L.co[1][x * SEL_PITCH + y];
      if (fabsf(l * inv_y[k] * qac) < 0.5f) energy += (double)l * (double)l;
    }
    energy = wave_sum_f64(energy);
    if (lane == 0) L.erased[wave] = energy;
    __syncthreads();
    const float erased = (float)(((L.erased[0] + L.erased[1]) + L.erased[2]) + L.erased[3]);
    const float dd = D->distance;
    const float t = fminf(fmaxf(erased / (2e-4f * dd * dd) - 1.0f, 0.0f), 1.0f);
    qscale = 1.0f + 0.6f * t;
  }
  __syncthreads();
  // Per-coefficient terms (sse_and_rate_scalar_impl::<true, false>,
  // src/inflated_cost.rs:435-500).  Element e = tid + 256 j lies in site
  // e / size (one site per wavefront and j); within the site it is storage
  // index k.
  constexpr int half = n / 2, cblk = n / 8;
  if (tid < 3 * nsites) {
    const int c = tid / nsites, site = tid - c * nsites;
    L.qs[c][site] = qac_in[site] * qscale * D->qm_mult[c];
  }
  __syncthreads();
  // Lanes take coefficients in scan order: within a site, e's low bits are
  // the natural-scan position p and order[p] the storage index, so a
  // wavefront's 64 lanes hold 64 consecutive scan positions of one site and
  // the nonzero count and last nonzero position come from one ballot.  All
  // three channels are read before one barrier and their d^2 terms written
  // over the coefficients after it.
  // A channel's four coefficients are read together and every table lookup
  // is clamped into its table, so the LDS reads of a channel are in flight
  // together; the two out-of-table cases (|q| >= 64: the quotient, |q| >=
  // 256: the rate) are patched afterwards, only when a lane has one.
  float dd_r[3][4];
  int kk_r[4], off[4], tsel[4];
  bool ac[4];
  float qsv[3][4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int e = tid + SEL_T * j;
    const int site = __builtin_amdgcn_readfirstlane(e >> (2 * LG));
    const int k = kq[j];  // storage index of scan position e mod size
    const int sy = site / sites_1d, sx = site - sy * sites_1d;
    const int u = k >> LG, v = k & (n - 1);  // square storage is u-major
    const int row = u, col = v;              // storage row / column
    off[j] = (sy * n + v) * SEL_PITCH + sx * n + u;
    kk_r[j] = site * size + k;
    ac[j] = !(col < cblk && row < cblk);
    tsel[j] = (row >= half ? 2 : 0) + (col >= half ? 1 : 0);
#pragma unroll
    for (int c = 0; c < 3; c++)
      qsv[c][j] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(L.qs[c][site])));
  }
  uint32_t vn = 0, vm = 0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const float thr0 = D->thr[c][cls][0], thr1 = D->thr[c][cls][1];
    const float thr2 = D->thr[c][cls][2], thr3 = D->thr[c][cls][3];
    float a[4], q[4], qt[4], rt[4], lev[4];
    bool slow = false;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float coef = L.co[c][off[j]];
      const float threshold = tsel[j] >= 2 ? (tsel[j] == 3 ? thr3 : thr2) : (tsel[j] == 1 ? thr1 : thr0);
      a[j] = ipq[c][j] * qsv[c][j] * coef;
      q[j] = (ac[j] && fabsf(a[j]) >= threshold) ? round_away(a[j]) : 0.0f;
      const float aq = fabsf(q[j]);
      qt[j] = L.quot[c][(int)fminf(aq, 63.0f)];
      rt[j] = L.rate[(int)fminf(aq, 255.0f)];
      slow |= aq >= 64.0f;
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
      // level_f32 with the extended quotient table (quot[0] = 0,
      // quot[1] = 1 - one): bit-identical to src/group.rs:748-768.
      lev[j] = q[j] - copysignf(qt[j], q[j]);
    if (__builtin_expect(__ballot(slow) != 0, 0)) {
      const float num = D->dequant_numerator;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const float aq = fabsf(q[j]);
        if (aq >= 64.0f) lev[j] = q[j] - num / q[j];
        if (aq >= 256.0f) rt[j] = rate_log2_big(D, aq);
      }
    }
    // Per wavefront and j: the nonzero count and 1 + the last nonzero's
    // scan position of its 64 scan positions, one slot (4 j + wave) each.
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float diff = a[j] - lev[j];
      dd_r[c][j] = ac[j] ? diff * diff : 0.0f;
      const bool nzq = q[j] != 0.0f;
      L.mg[c][kk_r[j]] = nzq ? rt[j] : 0.0f;
      const uint64_t m = __ballot(nzq);
      const uint32_t p0 = (uint32_t)((wave * 64 + SEL_T * j) & (size - 1));  // lane 0's scan position
      const bool mine = lane == 4 * c + j;
      vn = mine ? (uint32_t)__popcll(m) : vn;
      vm = mine ? (m ? p0 + 64u - (uint32_t)__clzll(m) : 0u) : vm;
    }
  }
  if (lane < 12) {
    L.nzp[lane >> 2][4 * (lane & 3) + wave] = vn;
    L.msp[lane >> 2][4 * (lane & 3) + wave] = vm;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int j = 0; j < 4; j++) (&L.co[c][0])[kk_r[j]] = dd_r[c][j];
  __syncthreads();
  // Sums in the reference's sequential row-major storage order (f32).
  for (int wk = tid; wk < nsites * 6; wk += SEL_T) {
    const int site = wk / 6, rest = wk - site * 6, c = rest >> 1, which = rest & 1;
    const float4 *src = reinterpret_cast<const float4 *>(
        which == 0 ? &L.co[c][0] + site * size : &L.mg[c][0] + site * size);
    // Eight float4 loads per step keep the LDS reads ahead of the
    // (necessarily sequential) f32 accumulation.
    float acc = 0.0f;
    for (int k0 = 0; k0 < (size >> 2); k0 += 8) {
      float4 v4[8];
#pragma unroll
      for (int i = 0; i < 8; i++) v4[i] = src[k0 + i];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        acc += v4[i].x;
        acc += v4[i].y;
        acc += v4[i].z;
        acc += v4[i].w;
      }
    }
    if (which == 0)
      L.sse[site][c] = acc;
    else
      L.mag[site][c] = acc;
  }
  __syncthreads();
  if (tid < nsites) {
    const int site = tid;
    const int sy = site / sites_1d, sx = site - sy * sites_1d;
    float d_total = 0.0f, r_total = 0.0f;
    for (int c = 0; c < 3; c++) {
      d_total += D->channel_weight[c] * L.sse[site][c];
      // site s of nsites holds the 16 / nsites slots from s * 16 / nsites
      constexpr int per = 16 / nsites;
      uint32_t nzc = 0, mscan = 0;
#pragma unroll
      for (int i = 0; i < per; i++) {
        nzc += L.nzp[c][site * per + i];
        mscan = max(mscan, L.msp[c][site * per + i]);
      }
      const float nzf = (float)nzc;
      // L.rate holds rate_lut[0..255]: the common counts read LDS
      const float header =
          0.4f * (nzc < 256u ? L.rate[nzc] : jxg_rate_log2(D, nzf));
      float visited = 0.0f;
      if (nzc != 0) {
        // maxscan holds 1 + the last nonzero's scan position
        const int64_t span = (int64_t)mscan - (int64_t)(cblk * cblk + nzc);
        visited = span > 0 ? (float)span : 0.0f;
      }
      r_total += nzf * 1.6f + 1.0f * L.mag[site][c] + header + 0.5f * visited;
    }
    // Footprint::visible_fraction (src/inflated_cost.rs:650-660)
    const int fx = px0 + sx * n, fy = py0 + sy * n;
    const int vw = min(n, max(W - fx, 0)), vh = min(n, max(H - fy, 0));
    const float visible = (vw == n && vh == n) ? 1.0f : (float)(vw * vh) / (float)(n * n);
    cost_out[site] = fmla<F>(RD_LAMBDA, r_total + META_R, d_total * visible);
  }
  __syncthreads();
}

// Loads the region's three 32x32 pixel planes into L.co.
__device__ __forceinline__ void load_region_nosync(SelLds &L, const float *__restrict__ xyb,
                                                   size_t plane, int stride, int px0, int py0) {
  const int tid = threadIdx.x;
  for (int e = tid; e < 3 * 32 * 8; e += SEL_T) {
    const int c = e >> 8, r = (e >> 3) & 31, q = e & 7;
    const float4 v = *reinterpret_cast<const float4 *>(xyb + c * plane + (size_t)(py0 + r) * stride +
                                                       px0 + 4 * q);
    float *dst = &L.co[c][r * SEL_PITCH + 4 * q];
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
    dst[3] = v.w;
  }
}

__device__ __forceinline__ void load_region(SelLds &L, const float *__restrict__ xyb, size_t plane,
                                            int stride, int px0, int py0) {
  load_region_nosync(L, xyb, plane, stride, px0, py0);
  __syncthreads();
}

__device__ __forceinline__ float mean_max_mixer(float d) {
  return fmaxf(1.0f - fmaxf(d - 1.54138f, 0.0f) * 0.56391f, 0.0f);
}

__device__ __forceinline__ uint8_t aggregate_quant(uint8_t max_q, uint32_t sum, int covered,
                                                   float d) {
  return jxg_aggregate_quant(max_q, sum, covered, d);
}

template <bool F>
__device__ __forceinline__ float risk_gated(float k, float accept, float q_min, float q_max,
                                            float area, const float *table) {
  if (q_min <= 0.0f || k <= 0.0f) return accept;
  const float spread = q_max / q_min - 1.0f;
  return accept * fast_exp2<F>(-k * area * spread, table);
}

template <bool F>
__device__ __forceinline__ float site_bias(const jxg_derived *D, double s1, double s2,
                                           uint32_t cnt) {
  if (D->site_k == 0.0f) return 1.0f;
  float var = 0.0f;
  if (cnt) {
    const double mean = s1 / (double)cnt;
    const double v = s2 / (double)cnt - mean * mean;
    var = (float)(v > 0.0 ? v : 0.0);
  }
  var = fmaxf(var, 1e-9f);
  const float s = fminf(fmaxf(dirty_log2f<F>(var / 0.003f) / 8.0f, -1.0f), 1.0f);
  return fast_exp<F>(D->site_k * s);
}

template <bool F>
__device__ __forceinline__ void select_body(const float *__restrict__ xyb, int stride,
                                                  int padded_h, int W, int H,
                                                  uint8_t *__restrict__ qf,
                                                  const int8_t *__restrict__ ytox,
                                                  const int8_t *__restrict__ ytob,
                                                  const float *__restrict__ tables,
                                                  const uint32_t *__restrict__ scans,
                                                  const float *__restrict__ selq,
                                                  const jxg_derived *__restrict__ D,
                                                  uint8_t *__restrict__ cells,
                                                  float *__restrict__ c8out,
                                                  float *__restrict__ c16out,
                                                  float *__restrict__ c32out, float *__restrict__ c32raw,
                                                  float *__restrict__ c8raw, int xb, int yb, int bix,
                                                  int biy) {
  __shared__ SelLds L;
  __shared__ uint8_t q_in[16];
  __shared__ uint8_t cell_out[16];
  __shared__ uint8_t q_out[16];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int rbx = bix * 4, rby = biy * 4;  // region's first block (global)
  const int gx = rbx >> 8, gy = rby >> 8;
  const int xsize = min(256, xb - gx * 256), ysize = min(256, yb - gy * 256);
  const int bx = rbx - gx * 256, by = rby - gy * 256;  // group-relative
  if (tid < 16) {
    const int ix = tid & 3, iy = tid >> 2;
    const bool in = rbx + ix < xb && rby + iy < yb;
    const uint8_t q = in ? qf[(size_t)(rby + iy) * xb + rbx + ix] : 1;
    q_in[tid] = q;
    q_out[tid] = q;
    cell_out[tid] = JXG_CELL(JXG_DCT, 1);
  }
  {
    const float num = D->dequant_numerator;
    if (tid < 192) {
      const int c = tid >> 6, a = tid & 63;
      L.quot[c][a] = a >= 2 ? num / (float)a : (a == 1 ? 1.0f - D->dequant_one[c] : 0.0f);
    }
    L.rate[tid] = D->rate_lut[tid];
  }
  const float scale = D->scale, dist = D->distance;
  const bool four_row = by + 4 <= ysize;
  const bool four_col = bx + 4 <= xsize;
  const bool any_row = four_row || by + 1 < ysize;
  const bool any_col = bx + 1 < xsize;
  const bool work = D->run_selection && any_row && any_col;
  const size_t plane = (size_t)stride * padded_h;
  const int px0 = rbx * 8, py0 = rby * 8;
  // cmap_factors (src/ac_strategy.rs:1838-1855): this region's 64x64 tile,
  // clamped to the DC group's tile grid; read ahead of the barrier.
  int8_t ytox_t = 0, ytob_t = 0;
  if (work) {
    const int tiles_x = (W + 63) / 64;
    const int ntx = (xsize + 7) / 8, nty = (ysize + 7) / 8;
    const int tx = min(bx / 8, ntx - 1), ty = min(by / 8, nty - 1);
    const size_t ti = (size_t)(gy * 32 + ty) * tiles_x + gx * 32 + tx;
    ytox_t = ytox[ti];
    ytob_t = ytob[ti];
  }
  // the region's pixels, loaded under the table set-up (one barrier for both)
  if (work) load_region_nosync(L, xyb, plane, stride, px0, py0);
  __syncthreads();
  if (work) {
    const float cmx = fmla<F>((float)ytox_t, D->cfl_inv_color_factor, D->cfl_base_x);
    const float cmb = fmla<F>((float)ytob_t, D->cfl_inv_color_factor, D->cfl_base_b);
    // Site luma statistics for MergeTuning::site_bias (site_variance_rect,
    // src/ac_strategy.rs:224-245): f64 sums of the in-image pixels of the
    // four 16x16 sites and the 32x32 region, reduced over the workgroup.
    {
      double s1[5] = {0, 0, 0, 0, 0}, s2[5] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int e = tid + SEL_T * j, r = e >> 5, col = e & 31;
        if (py0 + r < H && px0 + col < W) {
          const double v = (double)L.co[1][r * SEL_PITCH + col];
          const int q = (r >> 4) * 2 + (col >> 4);
#pragma unroll
          for (int s = 0; s < 4; s++)
            if (s == q) {
              s1[s] += v;
              s2[s] += v * v;
            }
          s1[4] += v;
          s2[4] += v * v;
        }
      }
      // the ten wavefront sums level by level (each keeps wave_sum_f64's
      // butterfly), so their shuffles are in flight together
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        double t1[5], t2[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
          t1[q] = __shfl_xor(s1[q], off, 64);
          t2[q] = __shfl_xor(s2[q], off, 64);
        }
#pragma unroll
        for (int q = 0; q < 5; q++) {
          s1[q] += t1[q];
          s2[q] += t2[q];
        }
      }
      if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 5; q++) {
          L.red1[wave][q] = s1[q];
          L.red2[wave][q] = s2[q];
        }
      }
      __syncthreads();
      if (tid < 5) {
        L.site_s1[tid] = ((L.red1[0][tid] + L.red1[1][tid]) + L.red1[2][tid]) + L.red1[3][tid];
        L.site_s2[tid] = ((L.red2[0][tid] + L.red2[1][tid]) + L.red2[2][tid]) + L.red2[3][tid];
        // in-image pixels of the site: 16x16 quadrant tid (0..3) or the region
        const int oy = tid < 4 ? 16 * (tid >> 1) : 0, ox = tid < 4 ? 16 * (tid & 1) : 0;
        const int ext = tid < 4 ? 16 : 32;
        const int ny = min(max(H - py0 - oy, 0), ext), nx = min(max(W - px0 - ox, 0), ext);
        L.site_cnt[tid] = (uint32_t)(ny * nx);
      }
    }
    float qac8[16], qac16[4], qac32[1];
#pragma unroll
    for (int i = 0; i < 16; i++) qac8[i] = scale * (float)q_in[i];
    // aggregate_qac_2x2 (src/ac_strategy/selection.rs:488-496)
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int sx = (s & 1) * 2, sy = (s >> 1) * 2;
      const float a = qac8[sy * 4 + sx], b = qac8[sy * 4 + sx + 1];
      const float c = qac8[(sy + 1) * 4 + sx], d = qac8[(sy + 1) * 4 + sx + 1];
      const float max_q = fmaxf(fmaxf(fmaxf(a, b), c), d);
      const float mixer = mean_max_mixer(dist);
      const float mean = (a + b + c + d) * 0.25f;
      const float mixed = (max_q * mixer + mean * (1.0f - mixer)) / scale;
      qac16[s] = scale * floorf(fminf(fmaxf(mixed + 0.5f, 1.0f), 255.0f));
    }
    {
      // region_qac (src/ac_strategy/selection.rs:898-927)
      uint8_t max_q = 1;
      uint32_t sum = 0;
      for (int i = 0; i < 16; i++) {
        max_q = q_in[i] > max_q ? q_in[i] : max_q;
        sum += q_in[i];
      }
      qac32[0] = scale * (float)aggregate_quant(max_q, sum, 16, dist);
    }
    eval_squares<F, 3>(L, qac8, cmx, cmb, tables, scans, selq, D, W, H, px0, py0, L.cost8);
    load_region(L, xyb, plane, stride, px0, py0);
    eval_squares<F, 4>(L, qac16, cmx, cmb, tables, scans + 64, selq + 64, D, W, H, px0, py0,
                       L.cost16);
    if (four_row && four_col) {
      load_region(L, xyb, plane, stride, px0, py0);
      eval_squares<F, 5>(L, qac32, cmx, cmb, tables, scans + 320, selq + 320, D, W, H, px0, py0,
                         &L.cost32);
    }
    // The five site biases and risk gates (four 16x16 sites, then the
    // region), one thread each, ahead of the decisions.
    if (tid < 5) {
      L.sbias[tid] = site_bias<F>(D, L.site_s1[tid], L.site_s2[tid], L.site_cnt[tid]);
      if (tid < 4) {
        const int b0 = (tid >> 1) * 8 + (tid & 1) * 2;
        float q_min = INFINITY, q_max = 0.0f;
        // the qac8 values of the site's blocks, recomputed from q_in (no
        // thread-dependent index into the register array)
        const float qs[4] = {scale * (float)q_in[b0], scale * (float)q_in[b0 + 1],
                             scale * (float)q_in[b0 + 4], scale * (float)q_in[b0 + 5]};
        for (int i = 0; i < 4; i++) {
          q_min = fminf(q_min, qs[i]);
          q_max = fmaxf(q_max, qs[i]);
        }
        L.sgate[tid] = risk_gated<F>(D->risk_k, D->accept_16, q_min, q_max, 1.0f, D->exp2_table);
      } else {
        uint8_t qmin = 255, qmax = 0;
        for (int i = 0; i < 16; i++) {
          qmin = q_in[i] < qmin ? q_in[i] : qmin;
          qmax = q_in[i] > qmax ? q_in[i] : qmax;
        }
        L.sgate[4] =
            risk_gated<F>(D->risk_k, D->accept_32, (float)qmin, (float)qmax, 2.0f, D->exp2_table);
      }
    }
    __syncthreads();
    // select_super_block at square scope (src/ac_strategy/selection.rs:238-455),
    // one thread per 16x16 super block.
    if (tid < 4) {
      const int sx = tid & 1, sy = tid >> 1;
      float child = 0.0f, base = 0.0f;
      int pick = 0;
      if ((four_row || sy == 0) && bx + sx * 2 + 1 < xsize) {
        const int b0 = (sy * 2) * 4 + sx * 2;
        const float c00 = L.cost8[b0], c01 = L.cost8[b0 + 1];
        const float c10 = L.cost8[b0 + 4], c11 = L.cost8[b0 + 5];
        const float left = c00 + c10, right = c01 + c11, top = c00 + c01, bottom = c10 + c11;
        const float total = left + right;
        const float c16 = D->bias_16x16 * L.sbias[tid] * L.cost16[tid];
        const float cost_16x8 = left + right, cost_8x16 = top + bottom;
        const float best_rect = fminf(cost_16x8, cost_8x16);
        // can_place_strategy(DCT16X16): aligned super blocks always fit
        // their 64-px tile and 256-px group once inside the image.
        const bool fits = bx + sx * 2 + 2 <= xsize && by + sy * 2 + 2 <= ysize;
        pick = fits && c16 < best_rect && c16 < total * L.sgate[tid];
        child = pick ? c16 : (cost_16x8 <= cost_8x16 ? cost_16x8 : cost_8x16);
        base = total;
      }
      L.child[tid] = child;
      L.base[tid] = base;
      L.pick16[tid] = pick;
    }
    __syncthreads();
    // The 32x32 merge (choose_merge_partition, src/ac_strategy/selection.rs:
    // 115-150 and 1040-1150, with rectangles priced +inf: only the square
    // merge can pass).
    if (tid == 0) {
      int m32 = 0;
      if (four_row && four_col) {
        const float merged = D->bias_32x32 * L.sbias[4] * L.cost32;
        const float child = L.child[0] + L.child[1] + L.child[2] + L.child[3];
        const float base = L.base[0] + L.base[1] + L.base[2] + L.base[3];
        m32 = merged < child && merged < base * L.sgate[4];
      }
      L.merge32 = m32;
    }
    __syncthreads();
    // Each block's cell, and adjust_quant_field (src/ac_strategy/selection.rs:
    // 503-527): the aggregate quant value over its transform's blocks.
    if (tid < 16) {
      const int ix = tid & 3, iy = tid >> 2;
      const int site = (iy >> 1) * 2 + (ix >> 1);
      int first = tid, cv = 1;
      uint8_t cell = JXG_CELL(JXG_DCT, 1);
      if (L.merge32) {
        first = 0;
        cv = 4;
        cell = JXG_CELL(JXG_DCT32X32, tid == 0);
      } else if (L.pick16[site]) {
        first = (iy >> 1) * 8 + (ix >> 1) * 2;
        cv = 2;
        cell = JXG_CELL(JXG_DCT16X16, tid == first);
      }
      cell_out[tid] = cell;
      // With the re-tiling after the search, the field is adjusted once the
      // layout is final (jxg_retile.hip).
      if (cv > 1 && !D->retile) {
        uint8_t max_q = 0;
        uint32_t sum = 0;
        for (int jy = 0; jy < cv; jy++)
          for (int jx = 0; jx < cv; jx++) {
            const uint8_t q = q_in[first + jy * 4 + jx];
            max_q = q > max_q ? q : max_q;
            sum += q;
          }
        q_out[tid] = aggregate_quant(max_q, sum, cv * cv, dist);
      }
    }
    __syncthreads();
    if (c32raw && tid == 0) c32raw[(size_t)rby * xb + rbx] = (four_row && four_col) ? L.cost32 : NAN;
    if (c8raw && tid < 16) {
      // the DCT8 costs for the re-tiling (NaN where the search did not price one)
      const int ix = tid & 3, iy = tid >> 2;
      const bool eval = (four_row ? true : iy < 2) && (bx + (ix >> 1) * 2 + 1 < xsize);
      if (rbx + ix < xb && rby + iy < yb) c8raw[(size_t)(rby + iy) * xb + rbx + ix] = eval ? L.cost8[tid] : NAN;
    }
    if (c8out && tid < 16) {
      const int ix = tid & 3, iy = tid >> 2;
      const int sx = ix >> 1, sy = iy >> 1;
      const bool eval = (four_row ? true : iy < 2) && (bx + sx * 2 + 1 < xsize);
      if (rbx + ix < xb && rby + iy < yb) {
        const size_t o = (size_t)(rby + iy) * xb + rbx + ix;
        c8out[o] = eval ? L.cost8[tid] : NAN;
        const bool s16 = (ix & 1) == 0 && (iy & 1) == 0 && eval;
        c16out[o] = s16 ? L.cost16[sy * 2 + sx] : NAN;
        c32out[o] = (tid == 0 && four_row && four_col) ? L.cost32 : NAN;
      }
    }
  } else if (c8raw && tid < 16) {
    const int ix = tid & 3, iy = tid >> 2;
    if (rbx + ix < xb && rby + iy < yb) c8raw[(size_t)(rby + iy) * xb + rbx + ix] = NAN;
  }
  if (!work && c8out && tid < 16) {
    const int ix = tid & 3, iy = tid >> 2;
    if (rbx + ix < xb && rby + iy < yb) {
      const size_t o = (size_t)(rby + iy) * xb + rbx + ix;
      c8out[o] = NAN;
      c16out[o] = NAN;
      c32out[o] = NAN;
    }
  }
  if (tid < 16) {
    const int ix = tid & 3, iy = tid >> 2;
    if (rbx + ix < xb && rby + iy < yb) {
      const size_t o = (size_t)(rby + iy) * xb + rbx + ix;
      if (cells) cells[o] = cell_out[tid];
      qf[o] = q_out[tid];
    }
  }
}

template <bool F>
__global__ __launch_bounds__(SEL_T, 5) void k_select(const float *__restrict__ xyb, int stride,
                                                  int padded_h, int W, int H,
                                                  uint8_t *__restrict__ qf,
                                                  const int8_t *__restrict__ ytox,
                                                  const int8_t *__restrict__ ytob,
                                                  const float *__restrict__ tables,
                                                  const uint32_t *__restrict__ scans,
                                                  const float *__restrict__ selq,
                                                  const jxg_derived *__restrict__ D,
                                                  uint8_t *__restrict__ cells,
                                                  float *__restrict__ c8out,
                                                  float *__restrict__ c16out,
                                                  float *__restrict__ c32out,
                                                  float *__restrict__ c32raw,
                                                  float *__restrict__ c8raw, int xb, int yb) {
  select_body<F>(xyb, stride, padded_h, W, H, qf, ytox, ytob, tables, scans, selq, D, cells, c8out,
                 c16out, c32out, c32raw, c8raw, xb, yb, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Forward transform + quantization (src/group.rs:1259-2240, first pass)

#define QT_TILE 64
#define QT_PITCH 65
#define QT_MAXT 64

#define QT_SUB8_SLOTS 12

// Sized so that four workgroups fit the 160 KB of LDS of a CU.
// The derived parameters are never written by a kernel: read through the
// constant address space, uniform reads of them are scalar loads even in
// loops that store to global memory.
typedef __attribute__((address_space(4))) const jxg_derived jxg_derived_c;
__device__ __forceinline__ const jxg_derived_c *derived_c(const jxg_derived *p) {
  return (const jxg_derived_c *)p;
}

struct QLds {
  float y2[QT_TILE * QT_PITCH];   // dequantized luma with zeroed LLF (CfL source)
  float cur[QT_TILE * QT_PITCH];  // the channel being transformed
  uint4 cdesc[64];                 // 64-coefficient chunk -> packed descriptor (chunk_desc)
  int qnext[3];                    // per channel: the next chunk a wavefront takes
  int n;
  uint8_t s[QT_MAXT];
  uint16_t bx[QT_MAXT], by[QT_MAXT];  // global first block
  uint8_t lx[QT_MAXT], ly[QT_MAXT];   // LDS origin (pixels)
  uint16_t pre1[QT_MAXT + 1];         // prefix of pass-1 lines (one channel)
  uint16_t pre2[QT_MAXT + 1];         // prefix of pass-2 lines
  uint16_t prec[QT_MAXT + 1];         // prefix of 64-coefficient chunks
  int16_t ydc[64];                    // quantized Y DC per 8x8 block of the window
  float xf[QT_MAXT], bf[QT_MAXT];
  uint8_t quant[QT_MAXT];
  float inv_qac[QT_MAXT];  // 1 / (scale * quant), src/group.rs:806
  float quot[64];          // 0, 1 - one, numerator / a for a >= 2 (see dequant_level_tab)
  uint8_t line1_t[512], line2_t[512];  // line -> transform, per pass
  union {
    float sub8_px[QT_SUB8_SLOTS][64];  // pixel copies for sub-8x8 transforms
    struct {
      float dcs[2][64];    // DCT64 DC: inverse-transform scratch
      float dc64v[2][64];  // DCT64 DC values of the current channel
    };
  };
  int n64;  // DCT64-family transforms in the list (<= 2)
  int t64[2];
  int16_t ydc64[2][64];  // quantized Y DC of the DCT64 transforms
};

// Index of the 8x8 block of the window holding DC value i of transform t.
__device__ __forceinline__ int win_block(const QLds &L, int t, int i, int cvx) {
  return ((L.ly[t] >> 3) + i / cvx) * 8 + (L.lx[t] >> 3) + i % cvx;
}

__device__ __forceinline__ int find_item(const int *pre, int n, int e) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// dc_from_coefficients (src/group.rs:183-260; src/dct.rs:930,1033,1864,2170,
// 2356,2381) on a transform held spatially in LDS.
template <bool F>
__device__ void dc_from(int s, const float *co, int pitch, int w, int h, float *dc) {
  const float r2[2] = {1.0f, 0.9017642f};
  const float r4[4] = {1.0f, 0.9748868f, 0.9017642f, 0.7870549f};
  auto at = [&](int k) {
    int v, u;
    storage_to_spatial(s, w, h, k, &v, &u);
    return co[v * pitch + u];
  };
  switch (s) {
    case JXG_DCT16X8:
    case JXG_DCT8X16: {
      const float s0 = at(0) * r2[0], s1 = at(1) * r2[1];
      dc[0] = s0 + s1;
      dc[1] = s0 - s1;
      return;
    }
    case JXG_DCT16X16: {
      const float s00 = at(0) * (r2[0] * r2[0]);
      const float s01 = at(1) * (r2[0] * r2[1]);
      const float s10 = at(16) * (r2[1] * r2[0]);
      const float s11 = at(17) * (r2[1] * r2[1]);
      const float r00 = s00 + s01, r01 = s00 - s01, r10 = s10 + s11, r11 = s10 - s11;
      dc[0] = r00 + r10;
      dc[1] = r00 - r10;
      dc[2] = r01 + r11;
      dc[3] = r01 - r11;
      return;
    }
    case JXG_DCT32X32: {
      float rr[4][4];
      for (int a = 0; a < 4; a++) {
        float sv[4];
        for (int b = 0; b < 4; b++) sv[b] = at(a * 32 + b) * r4[a] * r4[b];
        idct4<F>(sv, rr[a]);
      }
      for (int bb = 0; bb < 4; bb++) {
        const float cin[4] = {rr[0][bb], rr[1][bb], rr[2][bb], rr[3][bb]};
        float col[4];
        idct4<F>(cin, col);
        for (int r = 0; r < 4; r++) dc[bb * 4 + r] = col[r];
      }
      return;
    }
    case JXG_DCT32X16:
    case JXG_DCT16X32: {
      float rr[2][4];
      for (int a = 0; a < 2; a++) {
        float sv[4];
        for (int b = 0; b < 4; b++) sv[b] = at(a * 32 + b) * r2[a] * r4[b];
        idct4<F>(sv, rr[a]);
      }
      for (int bb = 0; bb < 4; bb++) {
        if (s == JXG_DCT32X16) {
          dc[bb * 2] = rr[0][bb] + rr[1][bb];
          dc[bb * 2 + 1] = rr[0][bb] - rr[1][bb];
        } else {
          dc[bb] = rr[0][bb] + rr[1][bb];
          dc[4 + bb] = rr[0][bb] - rr[1][bb];
        }
      }
      return;
    }
    case JXG_DCT32X8:
    case JXG_DCT8X32: {
      float sv[4];
      for (int b = 0; b < 4; b++) sv[b] = at(b) * r4[b];
      idct4<F>(sv, dc);
      return;
    }
    default: dc[0] = at(0); return;
  }
}

// Raw strategies the quantize stage transforms.
__host__ __device__ __forceinline__ bool quantize_supported(int s) {
  return s == JXG_DCT || s == JXG_DCT16X8 || s == JXG_DCT8X16 || s == JXG_DCT16X16 ||
         s == JXG_DCT32X32 || s == JXG_DCT32X16 || s == JXG_DCT16X32 || s == JXG_DCT32X8 ||
         s == JXG_DCT8X32 || is_sub8(s) || is_dct64(s);
}

struct QArgs {
  const float *xyb;
  int stride, padded_h, W, H, xb, yb;
  const uint8_t *cells, *qf;
  const int8_t *ytox, *ytob;
  const float *tables;
  const jxg_derived *D;
  int32_t *lv0, *lv1, *lv2;
  float *co0, *co1, *co2;
  int16_t *dc0, *dc1, *dc2;
  uint32_t *xt_count, *xt_list;
  uint32_t *status;  // [0..1] copy of {count, bad} for the caller, [2] finish ticket
  uint32_t xt_cap;
  uint32_t *bad;  // set to 1 when an unsupported strategy cell is seen
  uint32_t *xt_img = nullptr;  // batch: the image of each queued transform
  int img = 0;                 // batch: this image's index
};

__device__ __forceinline__ int ctz32(int v) { return __builtin_ctz((unsigned)v); }

// DequantBias::level (src/group.rs:770-790) with an extended quotient
// table: quot[0] = 0, quot[1] = 1 - one, quot[a] = numerator / a.  For
// |q| = 1 the result q - copysign(1 - one, q) is exactly +-one (1 - one is
// exact for one in [0.5, 1]), so the table form is bit-identical.
__device__ __forceinline__ float dequant_level_tab(float num, int32_t q, const float *quot) {
  const uint32_t a = q < 0 ? (uint32_t)(-(int64_t)q) : (uint32_t)q;
  const float qf = (float)q;
  if (a >= 64) return qf - num / qf;
  return qf - copysignf(quot[a], qf);
}

// Geometry of one coefficient of transform t: LDS address, storage row /
// column and output index (all sizes are powers of two).
struct CoefGeom {
  int a, row, col, hc, wc;
  size_t out;
};

__device__ __forceinline__ CoefGeom coef_geom(const QLds &L, int t, int s, int k, int xb) {
  const int cvx = cov_x(s), cvy = cov_y(s);
  const int lw = 3 + ctz32(cvx), lh = 3 + ctz32(cvy);
  int v, u;
  if (s == JXG_DCT) {
    u = k >> 3;
    v = k & 7;
  } else if (rows_first(s) || is_sub8(s)) {
    v = k >> lw;
    u = k & ((1 << lw) - 1);
  } else {
    u = k >> lh;
    v = k & ((1 << lh) - 1);
  }
  // 64-long axes keep dct64_line's even / odd split order in LDS
  if (cvy == 8) v = dct64_pos(v);
  if (cvx == 8) u = dct64_pos(u);
  const int lwc = lw > lh ? lw : lh;
  CoefGeom g;
  g.a = (L.ly[t] + v) * QT_PITCH + L.lx[t] + u;
  g.row = k >> lwc;
  g.col = k & ((1 << lwc) - 1);
  g.wc = 1 << lwc;
  g.hc = 1 << (lw > lh ? lh : lw);
  const int lch = 6 + ctz32(cvx);
  g.out = ((size_t)(L.by[t] + (k >> lch)) * xb + L.bx[t]) * 64 + (k & ((1 << lch) - 1));
  return g;
}

// coef_geom split for a chunk of 64 consecutive storage indices kc ..
// kc + 63 of transform t: everything but the lane's own position is uniform
// (scalar registers), and the chunk's outputs are contiguous (chunks start
// at multiples of 64 and an output row holds cov_x * 64 >= 64 values).
struct ChunkGeom {
  // storage index k splits into k >> sh and k mod 2^sh: (u, v) in that order,
  // (v, u) when swap (8x8: sh = 3; v-major: sh = lw, swap; u-major: sh = lh)
  int kc, lw, lh, lwc, sh, origin;
  bool swap, d64x, d64y;
  size_t out;  // output index of storage index kc
};

// A chunk's uniform data, packed once per tile into four words:
//   d0: LDS origin (12 bits) | (sh - 3) << 12 | (lw - 3) << 14 | (lh - 3) << 16 |
//       d64x << 18 | d64y << 19 | area class << 20 | kc mod 2^lch << 23
//   d1: block row of kc (by + kc >> lch) | block column << 16
//   d2: kc | transform << 12 | quant-table size << 18 | swap << 31
//   d3: quant-table offset + kc
// (lch = 6 + log2 cov_x: an output row of blocks holds 2^lch values; block
// coordinates are below 2^15 and the LDS origin below 2^12.)
// Writes the descriptors of transform t's chunks prec[t] .. prec[t + 1] - 1
// (the per-transform fields once, then the per-chunk ones).
__device__ __forceinline__ void chunk_pack(QLds &L, int t, uint32_t qoff, uint32_t qn) {
  const int s = L.s[t], cvx = cov_x(s), cvy = cov_y(s);
  const int lw = 3 + ctz32(cvx), lh = 3 + ctz32(cvy), lch = 6 + ctz32(cvx);
  const bool swap = s != JXG_DCT && (rows_first(s) || is_sub8(s));  // v-major storage
  const int sh = s == JXG_DCT ? 3 : (swap ? lw : lh);
  const int cls = jxg_area_class(max(cvx, cvy), min(cvx, cvy));
  const uint32_t x0 = (uint32_t)(L.ly[t] * QT_PITCH + L.lx[t]) | (uint32_t)(sh - 3) << 12 |
                      (uint32_t)(lw - 3) << 14 | (uint32_t)(lh - 3) << 16 | (uint32_t)(cvx == 8) << 18 |
                      (uint32_t)(cvy == 8) << 19 | (uint32_t)cls << 20;
  const uint32_t by = L.by[t], bx16 = (uint32_t)L.bx[t] << 16;
  const uint32_t z0 = (uint32_t)t << 12 | qn << 18 | (uint32_t)swap << 31;
  const int e0 = L.prec[t], e1 = L.prec[t + 1];
  for (int e = e0; e < e1; e++) {
    const uint32_t kc = (uint32_t)(e - e0) << 6;
    L.cdesc[e] = make_uint4(x0 | (kc & ((1u << lch) - 1)) << 23, (by + (kc >> lch)) | bx16, kc | z0,
                            qoff + kc);
  }
}

struct ChunkDesc {
  ChunkGeom G;
  int t, cls, qn;
  uint32_t toff;  // quant-table offset of storage index kc
};

// Chunk ch's descriptor in scalar registers.
__device__ __forceinline__ ChunkDesc chunk_desc(const QLds &L, int ch, int xb) {
  const uint4 d = L.cdesc[ch];
  const uint32_t d0 = __builtin_amdgcn_readfirstlane(d.x), d1 = __builtin_amdgcn_readfirstlane(d.y);
  const uint32_t d2 = __builtin_amdgcn_readfirstlane(d.z), d3 = __builtin_amdgcn_readfirstlane(d.w);
  ChunkDesc C;
  C.G.origin = (int)(d0 & 0xfffu);
  C.G.sh = 3 + (int)((d0 >> 12) & 3u);
  C.G.swap = d2 >> 31;
  C.G.lw = 3 + (int)((d0 >> 14) & 3u);
  C.G.lh = 3 + (int)((d0 >> 16) & 3u);
  C.G.lwc = C.G.lw > C.G.lh ? C.G.lw : C.G.lh;
  C.G.d64x = (d0 >> 18) & 1u;
  C.G.d64y = (d0 >> 19) & 1u;
  C.cls = (int)((d0 >> 20) & 7u);
  C.G.kc = (int)(d2 & 0xfffu);
  C.t = (int)((d2 >> 12) & 63u);
  C.qn = (int)((d2 >> 18) & 0x1fffu);
  C.toff = d3;
  C.G.out = (size_t)((d1 & 0xffffu) * (uint32_t)xb + (d1 >> 16)) * 64 + (d0 >> 23);
  return C;
}

// The lane's LDS address, storage row and column (as coef_geom).
template <bool BIG>
__device__ __forceinline__ void lane_geom(const ChunkGeom &G, int lane, int &a, int &row, int &col) {
  const int k = G.kc + lane;
  const int hi = k >> G.sh, lo = k & ((1 << G.sh) - 1);
  int u = G.swap ? lo : hi, v = G.swap ? hi : lo;
  // only the DCT64 family has 64-long axes
  if (BIG && G.d64y) v = dct64_pos(v);
  if (BIG && G.d64x) u = dct64_pos(u);
  a = G.origin + v * QT_PITCH + u;
  row = k >> G.lwc;
  col = k & ((1 << G.lwc) - 1);
}

// Copies channel c of the w x h pixel rectangle at (px, py) of the padded
// planes into L.cur at LDS origin (0, 0).  w is a multiple of 8.
__device__ __forceinline__ void load_plane(QLds &L, const QArgs &A, int c, int px, int py, int w,
                                           int h) {
  const size_t plane = (size_t)A.stride * A.padded_h;
  const float *src = A.xyb + c * plane;
  const int q4 = w >> 2;
  for (int e = threadIdx.x; e < h * q4; e += blockDim.x) {
    const int r = e / q4, q = e - r * q4;
    const float4 v = *reinterpret_cast<const float4 *>(src + (size_t)(py + r) * A.stride + px + 4 * q);
    float *dst = &L.cur[r * QT_PITCH + 4 * q];
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
    dst[3] = v.w;
  }
}

// Every transform of the list on L.cur (one channel): sub-8x8 transforms
// one block per thread, the DCTs as two passes of 1-D lines.
template <bool F, bool BIG>
__device__ __forceinline__ void transform_plane(QLds &L, int n) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  // A few threads, each owning one pixel-copy slot, take the sub-8x8 blocks.
  if (tid < QT_SUB8_SLOTS) {
    for (int t = tid; t < n; t += QT_SUB8_SLOTS) {
      const int s = L.s[t];
      if (is_sub8(s))
        sub8_transform<F>(s, &L.cur[L.ly[t] * QT_PITCH + L.lx[t]], QT_PITCH, L.sub8_px[tid]);
    }
  }
  for (int e = tid; e < L.pre1[n]; e += nthr) {
    const int t = L.line1_t[e];
    const int s = L.s[t];
    const int w = cov_x(s) * 8, h = cov_y(s) * 8;
    const int li = e - L.pre1[t];
    float *base = &L.cur[L.ly[t] * QT_PITCH + L.lx[t]];
    if (rows_first(s))
      dct_line_n<F, BIG>(w, base + li * QT_PITCH, 1, 0.0f, false);
    else
      dct_line_n<F, BIG>(h, base + li, QT_PITCH, 0.0f, false);
  }
  __syncthreads();
  for (int e = tid; e < L.pre2[n]; e += nthr) {
    const int t = L.line2_t[e];
    const int s = L.s[t];
    const int w = cov_x(s) * 8, h = cov_y(s) * 8;
    const int li = e - L.pre2[t];
    const float scale = 1.0f / (float)(w * h);
    float *base = &L.cur[L.ly[t] * QT_PITCH + L.lx[t]];
    if (rows_first(s))
      dct_line_n<F, BIG>(h, base + li, QT_PITCH, scale, true);
    else
      dct_line_n<F, BIG>(w, base + li * QT_PITCH, 1, scale, true);
  }
  __syncthreads();
}

// DC values of the DCT64-family transforms of the list from the channel's
// coefficients in L.cur (src/dct.rs:2602-2611 dc_from_dct64x64,
// :2767-2805 dc_from_dct64x32 / dc_from_dct32x64), into L.dc64v in the
// covered-block order iy * cov_x + ix.  Eight lanes per transform follow the
// reference's two inverse passes; called by every thread of the workgroup.
template <bool F>
__device__ void dc64_values(QLds &L) {
  const int tid = threadIdx.x;
  const int j = tid >> 3, i = tid & 7;
  const float r4[4] = {1.0f, 0.9748868f, 0.9017642f, 0.7870549f};
  if (tid < 16 && j < L.n64) {
    const int t = L.t64[j], s = L.s[t];
    const float *co = &L.cur[L.ly[t] * QT_PITCH + L.lx[t]];
    const int W = cov_x(s) * 8, H = cov_y(s) * 8;
    auto C = [&](int vf, int hf) {
      return co[(H == 64 ? dct64_pos(vf) : vf) * QT_PITCH + (W == 64 ? dct64_pos(hf) : hf)];
    };
    if (s == JXG_DCT64X64) {
      // inv_dct8x8 of low * 64: first the column of horizontal frequency i
      float col[8];
#pragma unroll
      for (int v = 0; v < 8; v++) col[v] = (C(v, i) * (RS64_8_DEV[v] * RS64_8_DEV[i])) * 64.0f;
      inv_dct8(col);
#pragma unroll
      for (int v = 0; v < 8; v++) L.dcs[j][v * 8 + i] = col[v];
    } else if (i < 4) {
      // rows[a] = idct1d_8_unnormalized over the long axis, a = short-axis frequency
      const int a = i;
      float r[8];
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const float c = s == JXG_DCT64X32 ? C(b, a) : C(a, b);
        r[b] = c * (r4[a] * RS64_8_DEV[b]);
      }
#pragma unroll
      for (int b = 0; b < 8; b++) r[b] *= 8.0f;
      inv_dct8(r);
#pragma unroll
      for (int b = 0; b < 8; b++) L.dcs[j][a * 8 + b] = r[b];
    }
  }
  __syncthreads();
  if (tid < 16 && j < L.n64) {
    const int s = L.s[L.t64[j]];
    if (s == JXG_DCT64X64) {
      float row[8];
#pragma unroll
      for (int u = 0; u < 8; u++) row[u] = L.dcs[j][i * 8 + u];
      inv_dct8(row);
#pragma unroll
      for (int u = 0; u < 8; u++) L.dc64v[j][i * 8 + u] = row[u];
    } else {
      const float cin[4] = {L.dcs[j][i], L.dcs[j][8 + i], L.dcs[j][16 + i], L.dcs[j][24 + i]};
      float col[4];
      idct4<F>(cin, col);
#pragma unroll
      for (int y = 0; y < 4; y++) {
        if (s == JXG_DCT64X32)
          L.dc64v[j][i * 4 + y] = col[y];  // 8 rows x 4 columns
        else
          L.dc64v[j][y * 8 + i] = col[y];  // 4 rows x 8 columns
      }
    }
  }
  __syncthreads();
}

// X / B DC with the luma term (src/group.rs:2003,2148 quantize_dc_cfl).
template <bool F>
__device__ __forceinline__ int16_t dc_cfl_q(int c, float xd, float yq, const jxg_derived *D) {
  if (c == 0) {
    if (D->cfl_factor_x != 0.0f)
      return sat_i16(round_away(fmla<F>(xd, D->inv_factor[0], -yq * D->cfl_factor_x)));
    return sat_i16(round_away(xd * D->inv_factor[0]));
  }
  return sat_i16(round_away(fmla<F>(xd, D->inv_factor[2], -yq * D->cfl_factor_b)));
}

// Transform + quantize the L.n transforms listed in L, one channel at a
// time: Y first (its dequantized form stays in L.y2 for CfL), then X and B.
// (px0, py0, win_w, win_h) is the pixel rectangle that L.cur holds.
// BIG: the list may hold DCT64-family transforms (the crossing / large
// transform pass); the tile pass is compiled without that code.
template <bool F, bool BIG>
__device__ void quantize_list(QLds &L, const QArgs &A, int px0, int py0, int win_w, int win_h) {
  const int tid = threadIdx.x;
  // the wavefront index is uniform: scalar registers, scalar branches
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nwaves = blockDim.x >> 6;
  const jxg_derived_c *D = derived_c(A.D);
  const int n = L.n;
  if (n == 0) return;
  // The list in order of line length (no output depends on the list order,
  // which the slot atomics leave arbitrary), so the lanes of a wavefront in
  // the line passes mostly run one length.  Wavefront 0 ranks the
  // transforms by (length, index) with ballots.
  if (n > 1) {
    int s_ = 0, bx_ = 0, by_ = 0, lx_ = 0, ly_ = 0, key = 4;
    if (tid < n) {
      s_ = L.s[tid];
      bx_ = L.bx[tid];
      by_ = L.by[tid];
      lx_ = L.lx[tid];
      ly_ = L.ly[tid];
      key = is_sub8(s_) ? 0 : ctz32(max(cov_x(s_), cov_y(s_)));
    }
    __syncthreads();
    if (tid < 64) {
      const uint64_t lt = (1ull << lane) - 1;
      int pos = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint64_t m = __ballot(key == k);
        pos += k < key ? __popcll(m) : (k == key ? __popcll(m & lt) : 0);
      }
      if (tid < n) {
        L.s[pos] = (uint8_t)s_;
        L.bx[pos] = (uint16_t)bx_;
        L.by[pos] = (uint16_t)by_;
        L.lx[pos] = (uint8_t)lx_;
        L.ly[pos] = (uint8_t)ly_;
      }
    }
    __syncthreads();
  }
  // Prefix sums of the transforms' pass-1 lines, pass-2 lines and
  // 64-coefficient chunks (n <= 64: one wavefront, lane = transform), and the
  // first two DCT64-family transforms in list order.
  if (tid < 64) {
    const int t = tid;
    int l1 = 0, l2 = 0, cc = 0;
    bool d64 = false;
    if (t < n) {
      const int s = L.s[t];
      d64 = is_dct64(s);
      const int w = cov_x(s) * 8, h = cov_y(s) * 8;
      if (!is_sub8(s)) {  // sub-8x8 transforms have no 1-D lines
        l1 = rows_first(s) ? h : w;
        l2 = rows_first(s) ? w : h;
      }
      cc = cov_x(s) * cov_y(s);
    }
    int i1 = l1, i2 = l2, ic = cc;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int a1 = __shfl_up(i1, o, 64), a2 = __shfl_up(i2, o, 64), ac = __shfl_up(ic, o, 64);
      if (t >= o) {
        i1 += a1;
        i2 += a2;
        ic += ac;
      }
    }
    if (t < n) {
      L.pre1[t] = i1 - l1;
      L.pre2[t] = i2 - l2;
      L.prec[t] = ic - cc;
    }
    if (t == n - 1) {
      L.pre1[n] = i1;
      L.pre2[n] = i2;
      L.prec[n] = ic;
    }
    const uint64_t m = __ballot(d64);
    if (t == 0) {
      const int n64 = __popcll(m) < 2 ? __popcll(m) : 2;
      if (n64 > 0) L.t64[0] = __builtin_ctzll(m);
      if (n64 > 1) L.t64[1] = __builtin_ctzll(m & (m - 1));
      L.n64 = n64;
    }
  }
  uint32_t t_qoff = 0, t_qn = 0;  // this thread's transform: quant-table offset and size
  if (tid < n) {
    const int t = tid;
    const int tiles_x = (A.W + 63) / 64;
    const size_t ti = (size_t)(L.by[t] / 8) * tiles_x + L.bx[t] / 8;
    L.xf[t] = fmla<F>((float)A.ytox[ti], D->cfl_inv_color_factor, D->cfl_base_x);
    L.bf[t] = fmla<F>((float)A.ytob[ti], D->cfl_inv_color_factor, D->cfl_base_b);
    const int quant = A.qf[(size_t)L.by[t] * A.xb + L.bx[t]];
    L.quant[t] = quant;
    L.inv_qac[t] = 1.0f / (D->scale * (float)quant);
    const int kind = qt_kind(L.s[t]);
    t_qoff = D->qt_off[kind];
    t_qn = D->qt_n[kind];
  }
  if (tid < 64)
    L.quot[tid] = tid >= 2 ? D->dequant_numerator / (float)tid
                           : (tid == 1 ? 1.0f - D->dequant_one[1] : 0.0f);
  if (tid < 3) L.qnext[tid] = 0;
  __syncthreads();
  // line / chunk -> transform tables
  if (tid < n) {
    const int a1 = L.pre1[tid], b1 = L.pre1[tid + 1], a2 = L.pre2[tid], b2 = L.pre2[tid + 1];
    for (int e = a1; e < b1; e++) L.line1_t[e] = (uint8_t)tid;
    for (int e = a2; e < b2; e++) L.line2_t[e] = (uint8_t)tid;
    chunk_pack(L, tid, t_qoff, t_qn);
  }
  // ---- Y ----
  load_plane(L, A, 1, px0, py0, win_w, win_h);
  __syncthreads();
  transform_plane<F, BIG>(L, n);
  // Y DC (dc_from_coefficients, quantize_dc_scalar src/group.rs:290)
  if constexpr (BIG) dc64_values<F>(L);
  if (BIG && tid < 128) {
    const int j = tid >> 6, i = tid & 63;
    if (j < L.n64) {
      const int t = L.t64[j], cvx = cov_x(L.s[t]);
      if (i < cvx * cov_y(L.s[t])) {
        const int16_t q = sat_i16(round_away(L.dc64v[j][i] * D->inv_factor[1]));
        L.ydc64[j][i] = q;
        A.dc1[(size_t)(L.by[t] + i / cvx) * A.xb + L.bx[t] + i % cvx] = q;
      }
    }
  }
  if (tid < n && !is_dct64(L.s[tid])) {
    const int t = tid, s = L.s[t];
    float dcv[16];
    dc_from<F>(s, &L.cur[L.ly[t] * QT_PITCH + L.lx[t]], QT_PITCH, cov_x(s) * 8, cov_y(s) * 8, dcv);
    const int cvx = cov_x(s), covered = cvx * cov_y(s);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      if (i < covered) {
        const int16_t q = sat_i16(round_away(dcv[i] * D->inv_factor[1]));
        L.ydc[win_block(L, t, i, cvx)] = q;
        A.dc1[(size_t)(L.by[t] + i / cvx) * A.xb + L.bx[t] + i % cvx] = q;
      }
    }
  }
  // Y quantize / dequantize round trip (quantize_roundtrip_y_block,
  // src/group.rs:795) and Y LLF zeroing (:1852).  A chunk is 64 consecutive
  // coefficients of one transform; a wavefront takes four chunks at a time
  // from the channel's queue (so the wavefront that ran the DC above takes
  // fewer) and issues their table loads together.
  const int nch = L.prec[n];
  const float num = D->dequant_numerator;
  (void)nwaves;
  for (;;) {
    int base = 0;
    if (lane == 0) base = atomicAdd(&L.qnext[1], 4);
    base = __builtin_amdgcn_readfirstlane(base);
    if (base >= nch) break;
    float tq[4], td[4], yv[4];
    ChunkDesc C[4];
    int aa[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int ch = base + u;
      if (ch < nch) {
        C[u] = chunk_desc(L, ch, A.xb);
        int row, col;
        lane_geom<BIG>(C[u].G, lane, aa[u], row, col);
        const float *tb = A.tables + C[u].toff + lane;
        tq[u] = tb[4 * C[u].qn];
        td[u] = tb[1 * C[u].qn];
        yv[u] = L.cur[aa[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int ch = base + u;
      if (ch < nch) {
        const ChunkGeom &G = C[u].G;
        const int t = C[u].t;
        int a, row, col;
        lane_geom<BIG>(G, lane, a, row, col);
        const int lmin = G.lw > G.lh ? G.lh : G.lw;
        const int cx = 1 << (G.lwc - 3), cy = 1 << (lmin - 3);
        A.co1[G.out + lane] = yv[u];
        const auto *thr = D->thr[1][C[u].cls];
        const float thr0 = thr[0], thr1 = thr[1], thr2 = thr[2], thr3 = thr[3];
        const int hc = 1 << lmin, wc = 1 << G.lwc;
        const float threshold = row >= hc / 2 ? (col >= wc / 2 ? thr3 : thr2)
                                              : (col >= wc / 2 ? thr1 : thr0);
        const float q_scaled = D->scale * (float)L.quant[t] * 1.0f;
        const float val = (tq[u] * q_scaled) * yv[u];
        const int32_t q = fabsf(val) >= threshold ? sat_i32(round_away(val)) : 0;
        A.lv1[G.out + lane] = q;
        // dequant_level_tab with the table lookup clamped and |q| >= 64
        // patched only when a lane of the chunk has one
        const uint32_t aq = q < 0 ? (uint32_t)(-(int64_t)q) : (uint32_t)q;
        const float qf = (float)q;
        float lev = qf - copysignf(L.quot[aq < 63u ? aq : 63u], qf);
        if (__builtin_expect(__ballot(aq >= 64u) != 0, 0))
          if (aq >= 64u) lev = qf - num / qf;
        float y2 = lev * td[u] * L.inv_qac[t];
        if (row < cy && col < cx) y2 = 0.0f;
        L.y2[a] = y2;
      }
    }
  }
  __syncthreads();
  // ---- X (c = 0) and B (c = 2): CfL (:1863), DC with CfL (:1871-2003,
  // 2128-2150) and AC quantization (quantize_block_ac_scalar :560) ----
  for (int c = 0; c <= 2; c += 2) {
    load_plane(L, A, c, px0, py0, win_w, win_h);
    __syncthreads();
    transform_plane<F, BIG>(L, n);
    // DC with CfL from the low-frequency coefficients.  The CfL step of the
    // AC loop subtracts f * y2 and y2 is 0 there, so they hold the same
    // values before and after it (up to the sign of a zero, which the
    // integer DC does not see): the DC runs first, beside the other
    // wavefronts' chunks.
    if constexpr (BIG) dc64_values<F>(L);
    if (BIG && tid < 128) {
      const int j = tid >> 6, i = tid & 63;
      if (j < L.n64) {
        const int t = L.t64[j], cvx = cov_x(L.s[t]);
        if (i < cvx * cov_y(L.s[t])) {
          int16_t *dc = c == 0 ? A.dc0 : A.dc2;
          dc[(size_t)(L.by[t] + i / cvx) * A.xb + L.bx[t] + i % cvx] =
              dc_cfl_q<F>(c, L.dc64v[j][i], (float)L.ydc64[j][i], A.D);
        }
      }
    }
    if (tid < n && !is_dct64(L.s[tid])) {
      const int t = tid, s = L.s[t];
      float xd[16];
      dc_from<F>(s, &L.cur[L.ly[t] * QT_PITCH + L.lx[t]], QT_PITCH, cov_x(s) * 8, cov_y(s) * 8, xd);
      const int cvx = cov_x(s), covered = cvx * cov_y(s);
      int16_t *dc = c == 0 ? A.dc0 : A.dc2;
#pragma unroll
      for (int i = 0; i < 16; i++) {
        if (i < covered) {
          const float yq = (float)L.ydc[win_block(L, t, i, cvx)];
          int16_t qd;
          if (c == 0) {
            if (D->cfl_factor_x != 0.0f)
              qd = sat_i16(round_away(fmla<F>(xd[i], D->inv_factor[0], -yq * D->cfl_factor_x)));
            else
              qd = sat_i16(round_away(xd[i] * D->inv_factor[0]));
          } else {
            qd = sat_i16(round_away(fmla<F>(xd[i], D->inv_factor[2], -yq * D->cfl_factor_b)));
          }
          dc[(size_t)(L.by[t] + i / cvx) * A.xb + L.bx[t] + i % cvx] = qd;
        }
      }
    }
    int32_t *lv = c == 0 ? A.lv0 : A.lv2;
    float *co = c == 0 ? A.co0 : A.co2;
    const float qm = D->qm_mult[c];
    for (;;) {
      int base = 0;
      if (lane == 0) base = atomicAdd(&L.qnext[c], 4);
      base = __builtin_amdgcn_readfirstlane(base);
      if (base >= nch) break;
      float tq[4], yq[4], cv[4];
      ChunkDesc C[4];
      int aa[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int ch = base + u;
        if (ch < nch) {
          C[u] = chunk_desc(L, ch, A.xb);
          int row, col;
          lane_geom<BIG>(C[u].G, lane, aa[u], row, col);
          tq[u] = A.tables[C[u].toff + (3 + c) * C[u].qn + lane];
          yq[u] = L.y2[aa[u]];
          cv[u] = L.cur[aa[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int ch = base + u;
        if (ch < nch) {
          const ChunkGeom &G = C[u].G;
          const int t = C[u].t;
          int a, row, col;
          lane_geom<BIG>(G, lane, a, row, col);
          const float f = c == 0 ? L.xf[t] : L.bf[t];
          const float v = fmla<F>(-f, yq[u], cv[u]);
          co[G.out + lane] = v;
          const auto *thr = D->thr[c][C[u].cls];
          const float thr0 = thr[0], thr1 = thr[1], thr2 = thr[2], thr3 = thr[3];
          const int hc = 1 << (G.lw > G.lh ? G.lh : G.lw), wc = 1 << G.lwc;
          const float threshold = row >= hc / 2 ? (col >= wc / 2 ? thr3 : thr2)
                                                : (col >= wc / 2 ? thr1 : thr0);
          const float q_scaled = D->scale * (float)L.quant[t] * qm;
          const float val = (tq[u] * q_scaled) * v;
          lv[G.out + lane] = fabsf(val) >= threshold ? sat_i32(round_away(val)) : 0;
        }
      }
    }
    __syncthreads();
  }
}

template <bool F>
__device__ __forceinline__ void quantize_body(QArgs A, int bix, int biy) {
  __shared__ QLds L;
  const int tid = threadIdx.x;
  const int bx0 = bix * 8, by0 = biy * 8;
  const int nbx = min(8, A.xb - bx0), nby = min(8, A.yb - by0);
  if (tid == 0) L.n = 0;
  __syncthreads();
  if (tid < 64) {
    const int lx = tid & 7, ly = tid >> 3;
    if (lx < nbx && ly < nby) {
      const uint8_t cell = A.cells[(size_t)(by0 + ly) * A.xb + bx0 + lx];
      if (cell & 1) {
        const int s = cell >> 1;
        if (s >= JXG_NUM_STRATEGIES || !quantize_supported(s) || bx0 + lx + cov_x(s) > A.xb ||
            by0 + ly + cov_y(s) > A.yb) {
          atomicOr(A.bad, 1u);
        } else if (lx + cov_x(s) <= 8 && ly + cov_y(s) <= 8 && !is_dct64(s)) {
          const int slot = atomicAdd(&L.n, 1);
          L.s[slot] = s;
          L.bx[slot] = bx0 + lx;
          L.by[slot] = by0 + ly;
          L.lx[slot] = lx * 8;
          L.ly[slot] = ly * 8;
        } else {
          const uint32_t slot = atomicAdd(A.xt_count, 1u);
          if (slot < A.xt_cap) {
            A.xt_list[slot] = (uint32_t)((by0 + ly) * A.xb + bx0 + lx);
            if (A.xt_img) A.xt_img[slot] = (uint32_t)A.img;
          }
        }
      }
    }
  }
  __syncthreads();
  quantize_list<F, false>(L, A, bx0 * 8, by0 * 8, 64, 64);
}

template <bool F>
__global__ __launch_bounds__(256, 4) void k_quantize(QArgs A) {
  quantize_body<F>(A, blockIdx.x, blockIdx.y);
}

// Transforms whose footprint crosses their 64x64 tile (the layout DP's
// crossing placements, src/dc_group_data.rs:207-217) and the DCT64-family
// transforms: one workgroup per queued transform, footprint gathered
// straight from the padded planes.
template <bool F>
__global__ __launch_bounds__(256) void k_quantize_xt(QArgs A) {
  __shared__ QLds L;
  const int tid = threadIdx.x;
  const uint32_t queued = *A.xt_count;
  if (queued == 0) {
    // Nothing crossed (the common case): every workgroup read the same 0,
    // written by the finished quantize launch, so workgroup 0 alone
    // publishes the status; the queue counter is already clear.
    if (blockIdx.x == 0 && tid == 0) {
      A.status[0] = 0u;
      A.status[1] = atomicExch(A.bad, 0u);
      __threadfence();
    }
    return;
  }
  const uint32_t count = min(queued, A.xt_cap);
  for (uint32_t item = blockIdx.x; item < count; item += gridDim.x) {
    const uint32_t blk = A.xt_list[item];
    const int bx = (int)(blk % (uint32_t)A.xb), by = (int)(blk / (uint32_t)A.xb);
    const int s = A.cells[blk] >> 1;
    if (tid == 0) {
      L.n = 1;
      L.s[0] = s;
      L.bx[0] = bx;
      L.by[0] = by;
      L.lx[0] = 0;
      L.ly[0] = 0;
    }
    __syncthreads();
    quantize_list<F, true>(L, A, bx * 8, by * 8, cov_x(s) * 8, cov_y(s) * 8);
    __syncthreads();
  }
  // The last workgroup to finish publishes the counters and clears them for
  // the next call (every workgroup has read them by then), so no separate
  // clearing launch is needed.
  if (tid == 0) {
    __threadfence();
    const uint32_t ticket = atomicAdd(&A.status[2], 1u);
    if (ticket == gridDim.x - 1) {
      A.status[0] = atomicExch(A.xt_count, 0u);
      A.status[1] = atomicExch(A.bad, 0u);
      A.status[2] = 0u;
      __threadfence();
    }
  }
}

// ---------------------------------------------------------------------------
// Batched stages: one launch per stage over every image of a batch (see
// jxg_batch.h).  Each workgroup maps its flat index to (image, local 2-D
// block) and runs the single-image stage body.

#define BATCH_BLOCK(stage)                                   \
  const int bi_ = batch_image(pfx, n, blockIdx.x);           \
  const BatchImg &I = img[bi_];                              \
  const uint32_t bl_ = blockIdx.x - pfx[bi_];                \
  const int bx_ = (int)(bl_ % (uint32_t)I.gw[stage]);        \
  const int by_ = (int)(bl_ / (uint32_t)I.gw[stage]);

template <bool F, int ORDER>
__global__ __launch_bounds__(256) void k_xyb_b(const BatchImg *__restrict__ img,
                                               const uint32_t *__restrict__ pfx, int n,
                                               const float *__restrict__ mlut) {
  BATCH_BLOCK(BS_XYB)
  xyb_body<F, ORDER>(I.rgb, (size_t)I.row_bytes, I.ch, I.W, I.H, mlut, I.xyb, I.stride, I.ph,
                     bx_, by_);
}

template <bool F>
__global__ __launch_bounds__(256) void k_aq_pre_b(const BatchImg *__restrict__ img,
                                                  const uint32_t *__restrict__ pfx, int n) {
  BATCH_BLOCK(BS_AQPRE)
  aq_pre_body<F>(I.xyb + (size_t)I.stride * I.ph, I.stride, I.W, I.H, I.pre_w, I.pre_h, I.pre, bx_,
                 by_);
}

// STRIP blocks per workgroup (STRIP * 8 threads): a tile shape the batch
// autotuner chooses.
template <bool F, int STRIP>
__global__ __launch_bounds__(STRIP * 8) void k_aq_block_b(const BatchImg *__restrict__ img,
                                                          const uint32_t *__restrict__ pfx, int n,
                                                          const float *__restrict__ exp2_table) {
  BATCH_BLOCK(BS_AQBLK)
  aq_block_body<F, STRIP>(I.xyb, I.stride, I.ph, I.W, I.H, I.pre, I.pre_w, I.distance, I.inv_scale,
                          exp2_table, I.qf, I.xb, I.yb, bx_, by_);
}

template <bool F>
__global__ __launch_bounds__(SEL_T, 5) void k_select_b(const BatchImg *__restrict__ img,
                                                    const uint32_t *__restrict__ pfx, int n,
                                                    const uint32_t *__restrict__ scans) {
  BATCH_BLOCK(BS_SEL)
  select_body<F>(I.xyb, I.stride, I.ph, I.W, I.H, I.qf, I.ytox, I.ytob, I.tables, scans, I.selq, I.D,
                 I.cells, nullptr, nullptr, nullptr, I.c32, I.D->retile ? I.pc : nullptr, I.xb,
                 I.yb, bx_, by_);
}

struct BatchQueue {
  uint32_t *xt_count, *xt_list, *xt_img, *status, *bad;
  uint32_t xt_cap;
};

__device__ __forceinline__ QArgs batch_qargs(const BatchImg &I, const BatchQueue &Q, int i) {
  QArgs A;
  A.xyb = I.xyb;
  A.stride = I.stride;
  A.padded_h = I.ph;
  A.W = I.W;
  A.H = I.H;
  A.xb = I.xb;
  A.yb = I.yb;
  A.cells = I.cells;
  A.qf = I.qf;
  A.ytox = I.ytox;
  A.ytob = I.ytob;
  A.tables = I.tables;
  A.D = I.D;
  A.lv0 = I.lv[0];
  A.lv1 = I.lv[1];
  A.lv2 = I.lv[2];
  A.co0 = I.co[0];
  A.co1 = I.co[1];
  A.co2 = I.co[2];
  A.dc0 = I.dc[0];
  A.dc1 = I.dc[1];
  A.dc2 = I.dc[2];
  A.xt_count = Q.xt_count;
  A.xt_list = Q.xt_list;
  A.xt_img = Q.xt_img;
  A.status = Q.status;
  A.xt_cap = Q.xt_cap;
  A.bad = Q.bad;
  A.img = i;
  return A;
}

template <bool F>
__global__ __launch_bounds__(256, 4) void k_quantize_b(const BatchImg *__restrict__ img,
                                                       const uint32_t *__restrict__ pfx, int n,
                                                       BatchQueue Q) {
  BATCH_BLOCK(BS_QUANT)
  quantize_body<F>(batch_qargs(I, Q, bi_), bx_, by_);
}

// The crossing / DCT64 queue of the whole batch; the last workgroup
// publishes and clears the counters as k_quantize_xt does.
template <bool F>
__global__ __launch_bounds__(256) void k_quantize_xt_b(const BatchImg *__restrict__ img,
                                                       BatchQueue Q) {
  __shared__ QLds L;
  const int tid = threadIdx.x;
  const uint32_t queued = *Q.xt_count;
  if (queued == 0) {  // as in k_quantize_xt: workgroup 0 alone publishes
    if (blockIdx.x == 0 && tid == 0) {
      Q.status[0] = 0u;
      Q.status[1] = atomicExch(Q.bad, 0u);
      __threadfence();
    }
    return;
  }
  const uint32_t count = min(queued, Q.xt_cap);
  for (uint32_t item = blockIdx.x; item < count; item += gridDim.x) {
    const int i = (int)Q.xt_img[item];
    const BatchImg &I = img[i];
    const QArgs A = batch_qargs(I, Q, i);
    const uint32_t blk = Q.xt_list[item];
    const int bx = (int)(blk % (uint32_t)A.xb), by = (int)(blk / (uint32_t)A.xb);
    const int s = A.cells[blk] >> 1;
    if (tid == 0) {
      L.n = 1;
      L.s[0] = s;
      L.bx[0] = bx;
      L.by[0] = by;
      L.lx[0] = 0;
      L.ly[0] = 0;
    }
    __syncthreads();
    quantize_list<F, true>(L, A, bx * 8, by * 8, cov_x(s) * 8, cov_y(s) * 8);
    __syncthreads();
  }
  if (tid == 0) {
    __threadfence();
    const uint32_t ticket = atomicAdd(&Q.status[2], 1u);
    if (ticket == gridDim.x - 1) {
      Q.status[0] = atomicExch(Q.xt_count, 0u);
      Q.status[1] = atomicExch(Q.bad, 0u);
      Q.status[2] = 0u;
      __threadfence();
    }
  }
}

// Stage launches of a batch (the CfL stages live in jxg_cfl.hip).  total[s]
// is the number of workgroups of stage s; pfx holds BS_N prefixes of n + 1.
template <bool F>
static void launch_pre(hipStream_t st, const BatchImg *d_img, const uint32_t *p, int n,
                       const uint32_t total[BS_N], const float *mlut, const float *exp2,
                       int aq_strip, int order) {
  const int m = n + 1;
  if (order == JXG_XYB_AVX2)
    k_xyb_b<F, JXG_XYB_AVX2><<<total[BS_XYB], 256, 0, st>>>(d_img, p + BS_XYB * m, n, mlut);
  else
    k_xyb_b<F, JXG_XYB_SCALAR><<<total[BS_XYB], 256, 0, st>>>(d_img, p + BS_XYB * m, n, mlut);
  k_aq_pre_b<F><<<total[BS_AQPRE], 256, 0, st>>>(d_img, p + BS_AQPRE * m, n);
  if (aq_strip == 16)
    k_aq_block_b<F, 16><<<total[BS_AQBLK], 128, 0, st>>>(d_img, p + BS_AQBLK * m, n, exp2);
  else if (aq_strip == 64)
    k_aq_block_b<F, 64><<<total[BS_AQBLK], 512, 0, st>>>(d_img, p + BS_AQBLK * m, n, exp2);
  else
    k_aq_block_b<F, 32><<<total[BS_AQBLK], 256, 0, st>>>(d_img, p + BS_AQBLK * m, n, exp2);
}

JXG_HIDDEN int batch_stages_pre(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                                const uint32_t *d_pfx, int n, const uint32_t total[BS_N],
                                int fused, const float *mlut, int aq_strip, int xyb_order) {
  if (fused)
    launch_pre<true>(st, d_img, d_pfx, n, total, mlut, ctx->d_exp2, aq_strip, xyb_order);
  else
    launch_pre<false>(st, d_img, d_pfx, n, total, mlut, ctx->d_exp2, aq_strip, xyb_order);
  LAUNCH_CHECK("batch xyb / aq");
  return JXG_OK;
}

JXG_HIDDEN int batch_stages_post(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                                 const uint32_t *d_pfx, int n, const uint32_t total[BS_N],
                                 int fused, uint32_t *xt_list, uint32_t *xt_img, uint32_t xt_cap,
                                 uint32_t *counters, int xt_grid) {
  const uint32_t *p = d_pfx;
  const int m = n + 1;
  BatchQueue Q;
  Q.xt_count = counters;
  Q.bad = counters + 1;
  Q.status = counters + 4;
  Q.xt_list = xt_list;
  Q.xt_img = xt_img;
  Q.xt_cap = xt_cap;
  if (fused)
    k_select_b<true><<<total[BS_SEL], SEL_T, 0, st>>>(d_img, p + BS_SEL * m, n, ctx->d_scans);
  else
    k_select_b<false><<<total[BS_SEL], SEL_T, 0, st>>>(d_img, p + BS_SEL * m, n, ctx->d_scans);
  LAUNCH_CHECK("batch select");
  int rc = batch_retile(ctx, st, d_img, d_pfx, n, total, fused);
  if (rc) return rc;
  if ((rc = batch_refine(ctx, st, d_img, d_pfx, n, total, fused))) return rc;
  if (fused) {
    k_quantize_b<true><<<total[BS_QUANT], 256, 0, st>>>(d_img, p + BS_QUANT * m, n, Q);
    k_quantize_xt_b<true><<<xt_grid, 256, 0, st>>>(d_img, Q);
  } else {
    k_quantize_b<false><<<total[BS_QUANT], 256, 0, st>>>(d_img, p + BS_QUANT * m, n, Q);
    k_quantize_xt_b<false><<<xt_grid, 256, 0, st>>>(d_img, Q);
  }
  LAUNCH_CHECK("batch quantize");
  return JXG_OK;
}

// Workgroup grids of the single-image stages for a W x H image (the batch
// uses the same shapes).
JXG_HIDDEN void batch_grids(uint32_t W, uint32_t H, uint32_t stride, uint32_t ph, int aq_strip,
                            int cm_rows, uint32_t gw[BS_N], uint32_t gh[BS_N]) {
  const uint32_t xb = (W + 7) / 8, yb = (H + 7) / 8, tx = (W + 63) / 64, ty = (H + 63) / 64;
  gw[BS_XYB] = (stride + XYB_PX_WG - 1) / XYB_PX_WG;
  gh[BS_XYB] = (ph + XYB_ROWS - 1) / XYB_ROWS;
  gw[BS_AQPRE] = (8 * xb + 255) / 256;  // four lanes per pre-map value
  gh[BS_AQPRE] = 2 * yb;
  gw[BS_AQBLK] = (xb + aq_strip - 1) / aq_strip;
  gh[BS_AQBLK] = yb;
  gw[BS_CMAPB] = tx;
  gh[BS_CMAPB] = (64 / cm_rows) * ty;
  gw[BS_CMAPS] = tx * ty;
  gh[BS_CMAPS] = 1;
  gw[BS_SEL] = (xb + 3) / 4;
  gh[BS_SEL] = (yb + 3) / 4;
  gw[BS_QUANT] = (xb + 7) / 8;
  gh[BS_QUANT] = (yb + 7) / 8;
  gw[BS_PIECE] = (xb + 3) / 4;
  gh[BS_PIECE] = (yb + 3) / 4;
  gw[BS_RETILE] = 32 * ((xb + 255) / 256);  // 32 tile slots per DC group and axis
  gh[BS_RETILE] = 32 * ((yb + 255) / 256);
  gw[BS_ADJUST] = (xb * yb + 255) / 256;
  gh[BS_ADJUST] = 1;
  refine_grids(W, H, &gw[BS_RLIST], &gw[BS_RCOST_S], &gw[BS_RCOST_B], &gw[BS_TRELLIS]);
  gh[BS_RLIST] = gh[BS_RCOST_S] = gh[BS_RCOST_B] = gh[BS_TRELLIS] = 1;
  rerank_grids(W, H, &gw[BS_RRLIST], &gw[BS_RRCOST_S], &gw[BS_RRCOST_B], &gw[BS_RRDECIDE]);
  gh[BS_RRLIST] = gh[BS_RRCOST_S] = gh[BS_RRCOST_B] = gh[BS_RRDECIDE] = 1;
}

// ---------------------------------------------------------------------------
// Candidate costs at listed sites: strategy_cost with the coefficient model
// (src/ac_strategy.rs:724 -> strategy_cost_impl :1293-1377) for any strategy
// up to 64x64, the DCT64 family in strategy_cost64's final form (:646-720),
// sub-8x8 strategies with the arithmetic of sub8_strategy_costs' evaluate
// (:1768-1795, identical to strategy_cost_impl for them).  One workgroup per
// site: the three channels' footprints are transformed in LDS (the quantize
// stage's transforms), CfL is subtracted from X and B, every coefficient's
// squared error and rate term is stored, and six lanes fold the six sums in
// the reference's sequential storage order.

struct SiteArgs {
  const float *xyb;
  int stride, padded_h, W, H, tiles_x, tiles_y;
  const int8_t *ytox, *ytob;
  const float *tables;
  const jxg_derived *D;
  const uint16_t *scanpos;  // scan position of each storage index, per shape
  int s;
  uint32_t n;
  const uint32_t *sites;  // by << 16 | bx
  const float *qac;
  float meta_r;
  float *cost;
};

// Offset of shape (cx >= cy) in the scan-position table: (1,1) (2,1) (2,2)
// (4,1) (4,2) (4,4) (8,4) (8,8).
__host__ __device__ __forceinline__ int scanpos_offset(int cx, int cy) {
  const int a = cx * cy;
  switch (a) {
    case 1: return 0;
    case 2: return 64;
    case 4: return cx == 2 ? 192 : 448;
    case 8: return 704;
    case 16: return 1216;
    case 32: return 2240;
    default: return 4288;
  }
}
#define SCANPOS_TOTAL (4288 + 4096)

template <bool F, bool BIG>
__global__ __launch_bounds__(256) void k_site_cost(SiteArgs S) {
  constexpr int NMAX = BIG ? 4096 : 1024;
  __shared__ QLds L;
  __shared__ float terms[3][2][NMAX];  // per channel: squared errors, rate terms
  __shared__ double erased[4];
  __shared__ float sums[3][2];
  __shared__ uint32_t nz[3], mx[3];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const jxg_derived *D = S.D;
  const int s = S.s;
  const uint32_t site = S.sites[blockIdx.x];
  const int bx = (int)(site & 0xffffu), by = (int)(site >> 16);
  const int cvx = cov_x(s), cvy = cov_y(s), w = cvx * 8, h = cvy * 8;
  const int cx = max(cvx, cvy), cy = min(cvx, cvy), size = cx * cy * 64;
  if ((bx + cvx) * 8 > S.stride || (by + cvy) * 8 > S.padded_h) {
    if (tid == 0) S.cost[blockIdx.x] = __int_as_float(0x7fc00000);  // NaN: site off the planes
    return;
  }
  QArgs A;
  A.xyb = S.xyb;
  A.stride = S.stride;
  A.padded_h = S.padded_h;
  if (tid == 0) {
    L.n = 1;
    L.s[0] = (uint8_t)s;
    L.bx[0] = (uint16_t)bx;
    L.by[0] = (uint16_t)by;
    L.lx[0] = 0;
    L.ly[0] = 0;
    L.pre1[0] = 0;
    L.pre2[0] = 0;
    L.pre1[1] = is_sub8(s) ? 0 : (rows_first(s) ? h : w);
    L.pre2[1] = is_sub8(s) ? 0 : (rows_first(s) ? w : h);
  }
  if (tid < 3) {
    nz[tid] = 0;
    mx[tid] = 0;
  }
  if (tid < 64) {
    L.line1_t[tid] = 0;
    L.line2_t[tid] = 0;
  }
  const int kind = qt_kind(s);
  const float *tb = S.tables + D->qt_off[kind];
  const int qn = D->qt_n[kind];
  const uint16_t *scanpos = S.scanpos + scanpos_offset(cx, cy);
  // cmap_factors (src/ac_strategy.rs:1838-1855)
  const int tx = min(bx / 8, S.tiles_x - 1), ty = min(by / 8, S.tiles_y - 1);
  const float xf = fmla<F>((float)S.ytox[ty * S.tiles_x + tx], D->cfl_inv_color_factor, D->cfl_base_x);
  const float bf = fmla<F>((float)S.ytob[ty * S.tiles_x + tx], D->cfl_inv_color_factor, D->cfl_base_b);
  // ---- Y: transform, keep the coefficients in L.y2 ----
  load_plane(L, A, 1, bx * 8, by * 8, w, h);
  __syncthreads();
  transform_plane<F, BIG>(L, 1);
  for (int e = tid; e < h * w; e += 256) {
    const int r = e / w, c = e - r * w;
    L.y2[r * QT_PITCH + c] = L.cur[r * QT_PITCH + c];
  }
  __syncthreads();
  // selection_quantizer::scale (src/ac_strategy/selection_quantizer.rs:82-120)
  float qac = S.qac[blockIdx.x];
  if ((s == JXG_DCT32X32 || s == JXG_DCT32X16 || s == JXG_DCT16X32) && S.meta_r != 0.0f &&
      D->selection_quantizer) {
    const float *inv_y = tb + 4 * qn;
    double energy = 0.0;
    for (int k = tid; k < size; k += 256) {
      const CoefGeom g = coef_geom(L, 0, s, k, 0);
      if (g.col < cx && g.row < cy) continue;
      const float l = L.y2[g.a];
      if (fabsf(l * inv_y[k] * qac) < 0.5f) energy += (double)l * (double)l;
    }
    energy = wave_sum_f64(energy);
    if (lane == 0) erased[wave] = energy;
    __syncthreads();
    const float e32 = (float)(((erased[0] + erased[1]) + erased[2]) + erased[3]);
    const float dd = D->distance;
    const float t = fminf(fmaxf(e32 / (2e-4f * dd * dd) - 1.0f, 0.0f), 1.0f);
    qac = qac * (1.0f + 0.6f * t);
  }
  const int cls = jxg_area_class(cx, cy);
  // Per-coefficient terms of channel c (sse_and_rate_scalar_impl,
  // src/inflated_cost.rs:435-500); src = L.cur (X / B, CfL applied) or
  // L.y2 (Y).
  auto channel_terms = [&](int c) {
    const float *inv = tb + (3 + c) * qn;
    const float qs = qac * D->qm_mult[c];
    const float one = D->dequant_one[c], num = D->dequant_numerator;
    const float *thr = D->thr[c][cls];
    const float f = c == 0 ? xf : bf;
    uint32_t cnt = 0, best = 0;
    for (int k0 = 0; k0 < size; k0 += 256) {
      const int k = k0 + tid;
      bool nzq = false;
      uint32_t pos1 = 0;
      if (k < size) {
        const CoefGeom g = coef_geom(L, 0, s, k, 0);
        float coef = c == 1 ? L.y2[g.a] : fmla<F>(-f, L.y2[g.a], L.cur[g.a]);
        float dd = 0.0f, mg = 0.0f;
        if (!(g.col < cx && g.row < cy)) {
          const int yfix = g.row >= cy * 4 ? 2 : 0;
          const float threshold = g.col >= cx * 4 ? thr[yfix + 1] : thr[yfix];
          const float a = inv[k] * qs * coef;
          const float q = fabsf(a) >= threshold ? round_away(a) : 0.0f;
          const float diff = a - level_f32(one, num, q);
          dd = diff * diff;
          if (q != 0.0f) {
            mg = rate_log2_slow(D, fabsf(q));
            nzq = true;
            pos1 = (uint32_t)scanpos[k] + 1u;
          }
        }
        terms[c][0][k] = dd;
        terms[c][1][k] = mg;
      }
      cnt += (uint32_t)__popcll(__ballot(nzq));
      best = max(best, pos1);
    }
    for (int o = 32; o > 0; o >>= 1) best = max(best, (uint32_t)__shfl_xor((int)best, o));
    if (lane == 0) {
      atomicAdd(&nz[c], cnt);
      atomicMax(&mx[c], best);
    }
  };
  channel_terms(1);
  // ---- X, then B ----
  for (int c = 0; c <= 2; c += 2) {
    __syncthreads();
    load_plane(L, A, c, bx * 8, by * 8, w, h);
    __syncthreads();
    transform_plane<F, BIG>(L, 1);
    channel_terms(c);
  }
  __syncthreads();
  // six sequential f32 folds in storage order
  {
    const int which = tid >> 5;  // lanes 0 and 32 of each of the first three waves
    if ((tid & 31) == 0 && which < 6) {
      const int c = which >> 1, part = which & 1;
      const float4 *src = reinterpret_cast<const float4 *>(&terms[c][part][0]);
      float acc = 0.0f;
      for (int k4 = 0; k4 < (size >> 2); k4 += 2) {
        const float4 a0 = src[k4], a1 = src[k4 + 1];
        acc += a0.x;
        acc += a0.y;
        acc += a0.z;
        acc += a0.w;
        acc += a1.x;
        acc += a1.y;
        acc += a1.z;
        acc += a1.w;
      }
      sums[c][part] = acc;
    }
  }
  __syncthreads();
  if (tid == 0) {
    float d_total = 0.0f, r_total = 0.0f;
    for (int c = 0; c < 3; c++) {
      d_total += D->channel_weight[c] * sums[c][0];
      const uint32_t nzc = nz[c];
      const float nzf = (float)nzc;
      const float header = 0.4f * jxg_rate_log2(D, nzf);
      float visited = 0.0f;
      if (nzc != 0) {  // mx holds 1 + the last nonzero's scan position
        const int64_t span = (int64_t)mx[c] - (int64_t)(cx * cy + nzc);
        visited = span > 0 ? (float)span : 0.0f;
      }
      r_total += nzf * 1.6f + 1.0f * sums[c][1] + header + 0.5f * visited;
    }
    // Footprint::visible_fraction (src/inflated_cost.rs:650-660)
    const int fx = bx * 8, fy = by * 8;
    const int vw = min(w, max(S.W - fx, 0)), vh = min(h, max(S.H - fy, 0));
    const float visible = (vw == w && vh == h) ? 1.0f : (float)(vw * vh) / (float)(w * h);
    float cost;
    if (is_dct64(s))
      cost = d_total * visible + RD_LAMBDA * (r_total + S.meta_r);  // strategy_cost64
    else
      cost = fmla<F>(RD_LAMBDA, r_total + S.meta_r, d_total * visible);
    S.cost[blockIdx.x] = cost;
  }
}

// Byte fill, used instead of hipMemsetAsync so that streams captured into a
// graph hold kernel nodes only.
__global__ void k_fill_u8(uint8_t *p, size_t n, uint8_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

JXG_HIDDEN int fill_u8(jxg_ctx *ctx, uint8_t *p, size_t n, uint8_t v) {
  if (n == 0) return JXG_OK;
  const size_t blocks = (n + 255) / 256;
  k_fill_u8<<<(unsigned)(blocks < 4096 ? blocks : 4096), 256, 0, ctx->stream>>>(p, n, v);
  LAUNCH_CHECK("k_fill_u8");
  return JXG_OK;
}

// ---------------------------------------------------------------------------
// Context and C ABI

// Device-resident AQ pre-map of jxg_fill_quant_field_noalloc_dev: one copy
// per GPU, held until the process ends.  Host code never passes the array's
// name to a kernel; jxg_ctx_create resolves its device address once per
// context (hipGetSymbolAddress, with the context's device current) into
// ctx->d_noalloc_pre, and only that address is used.
static __device__ __align__(256) float scratch[JXG_NOALLOC_AQ_PRE_FLOATS];

// Uploads a quant-table blob and the scan-ordered copy of the DCT8 /
// DCT16X16 / DCT32X32 inverse weights the strategy search reads.
static int upload_tables(jxg_ctx *ctx, const float *blob, float *d_blob, float *d_selq) {
  HIP_TRY(hipMemcpy(d_blob, blob, JXG_QT_BLOB_FLOATS * sizeof(float), hipMemcpyHostToDevice));
  static const int kinds[3] = {JXG_QT_DCT8, JXG_QT_DCT16X16, JXG_QT_DCT32X32};
  static const int offs[3] = {0, 64, 320};
  float selq[3 * SELQ_N];
  for (int i = 0; i < 3; i++) {
    const uint32_t off = jxg_qt_offset(kinds[i]), n = jxg_qt_size(kinds[i]);
    for (int c = 0; c < 3; c++)
      for (uint32_t p = 0; p < n; p++)
        selq[c * SELQ_N + offs[i] + p] = blob[off + (3 + c) * n + ctx->h_order[offs[i] + p]];
  }
  HIP_TRY(hipMemcpy(d_selq, selq, sizeof(selq), hipMemcpyHostToDevice));
  return JXG_OK;
}

JXG_HIDDEN int tables_for(jxg_ctx *ctx, const jxg_frame_params *p, const float **tables,
                          const float **selq) {
  if (ctx->tables_set) {
    *tables = ctx->d_tables;
    *selq = ctx->d_selq;
    return JXG_OK;
  }
  if (!p) return set_err(JXG_ERR_INVALID_ARG, "null params");
  const int slot = 2 * jxg_quant_tier(p->distance, p->arbiter) + (p->fused ? 1 : 0);
  float *d_blob = ctx->d_tier_tables + (size_t)slot * JXG_QT_BLOB_FLOATS;
  float *d_sq = ctx->d_tier_selq + (size_t)slot * 3 * SELQ_N;
  if (!ctx->tier_valid[slot]) {
    float *blob = new float[JXG_QT_BLOB_FLOATS];
    int rc = jxg_default_quant_tables(p->distance, p->arbiter, p->fused, blob);
    if (rc == JXG_OK) rc = upload_tables(ctx, blob, d_blob, d_sq);
    delete[] blob;
    if (rc != JXG_OK) return rc == JXG_ERR_INVALID_DISTANCE || rc == JXG_ERR_INVALID_ARG
                                 ? set_err(rc, "invalid table parameters (distance %g, arbiter %d)",
                                           (double)p->distance, p->arbiter)
                                 : rc;
    ctx->tier_valid[slot] = true;
  }
  *tables = d_blob;
  *selq = d_sq;
  return JXG_OK;
}

// scan_pos_lut (src/coeff_order.rs:358-415) of the square shapes from the
// natural order (natural_scan_generic, src/coeff_order.rs:206-248).
static void natural_scan(int cx, int cy, uint32_t *out) {
  const int side = cx * 8, xs = cx / cy, xsm = xs - 1;
  int xss = 0;
  while ((1 << xss) < xs) xss++;
  int cur = cx * cy;
  for (int i = 0; i < side; i++)
    for (int j = 0; j <= i; j++) {
      int x = j, y = i - j;
      if (i & 1) {
        const int t = x;
        x = y;
        y = t;
      }
      if (y & xsm) continue;
      y >>= xss;
      const int val = (x < cx && y < cy) ? y * cx + x : cur++;
      out[val] = (uint32_t)(y * side + x);
    }
  for (int ip = side - 1; ip >= 1; ip--) {
    const int i = ip - 1;
    for (int j = 0; j <= i; j++) {
      int x = side - 1 - (i - j), y = side - 1 - j;
      if (i & 1) {
        const int t = x;
        x = y;
        y = t;
      }
      if (y & xsm) continue;
      y >>= xss;
      out[cur++] = (uint32_t)(y * side + x);
    }
  }
}

int check_image(uint32_t w, uint32_t h) {
  if (w == 0 || h == 0) return set_err(JXG_ERR_EMPTY_IMAGE, "empty image %ux%u", w, h);
  if (w > JXG_MAX_DIM || h > JXG_MAX_DIM)
    return set_err(JXG_ERR_TOO_LARGE, "image %ux%u exceeds %u", w, h, JXG_MAX_DIM);
  return JXG_OK;
}

int check_speed(const jxg_frame_params *p) {
  if (!p) return set_err(JXG_ERR_INVALID_ARG, "null params");
  if (p->speed > JXG_SPEED_FAST)
    return set_err(JXG_ERR_UNSUPPORTED,
                   "speed %d: the transform-layout decisions of MEDIUM and slower tiers are not "
                   "built (ULTRAFAST, FASTEST and FAST are)", p->speed);
  return JXG_OK;
}

int check_planes(uint32_t w, uint32_t h, uint32_t stride, uint32_t padded_h) {
  uint32_t s, ph;
  jxg_plane_geometry(w, h, &s, &ph);
  if (stride < s || padded_h < ph || (stride & 3))
    return set_err(JXG_ERR_INVALID_ARG, "plane geometry %ux%u too small for %ux%u (need %ux%u)",
                   stride, padded_h, w, h, s, ph);
  return JXG_OK;
}

int upload_derived(jxg_ctx *ctx, const jxg_frame_params *p) {
  if (!p) return set_err(JXG_ERR_INVALID_ARG, "null params");
  if (ctx->derived_valid && memcmp(&ctx->last_params, p, sizeof(*p)) == 0) return JXG_OK;
  jxg_derived tmp;
  const int rc = jxg_derive(p, &tmp);
  if (rc != JXG_OK) return set_err(rc, "invalid frame parameters (distance %g)", (double)p->distance);
  HIP_TRY(hipStreamSynchronize(ctx->stream));  // staging buffer may still be in flight
  tmp.rate_big = ctx->d_rate_big;
  *ctx->h_derived = tmp;
  HIP_TRY(hipMemcpyAsync(ctx->d_derived, ctx->h_derived, sizeof(jxg_derived), hipMemcpyHostToDevice,
                         ctx->stream));
  ctx->last_params = *p;
  ctx->derived_valid = true;
  return JXG_OK;
}

int zero_cmaps(jxg_ctx *ctx, uint32_t w, uint32_t h, const int8_t **x, const int8_t **b) {
  if (*x && *b) return JXG_OK;
  const size_t n = (size_t)((w + 63) / 64) * ((h + 63) / 64);
  if (ctx->zero_cap < n) {
    int rc = ensure(&ctx->d_zero_cmap, &ctx->zero_cap, n);
    if (rc) return rc;
    rc = fill_u8(ctx, (uint8_t *)ctx->d_zero_cmap, n, 0);
    if (rc) return rc;
  }
  if (!*x) *x = ctx->d_zero_cmap;
  if (!*b) *b = ctx->d_zero_cmap;
  return JXG_OK;
}

extern "C" {

const char *jxg_last_error(void) { return g_err; }

int jxg_ctx_create(int device, void *stream, jxg_ctx **out) {
  if (!out) return set_err(JXG_ERR_INVALID_ARG, "null out");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return set_err(JXG_ERR_NO_DEVICE, "no HIP device visible");
  if (device < 0 || device >= count)
    return set_err(JXG_ERR_INVALID_ARG, "device %d out of range (%d devices)", device, count);
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_err(JXG_ERR_NO_DEVICE, "device %d is %s, this library is built for gfx950", device,
                   prop.gcnArchName);
  HIP_TRY(hipSetDevice(device));
  jxg_ctx *ctx = new jxg_ctx();
  ctx->device = device;
  ctx->stream = (hipStream_t)stream;
  auto fail = [&](int rc) {
    jxg_ctx_destroy(ctx);
    return rc;
  };
  {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(set_err(JXG_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e)));
    ctx->d_noalloc_pre = (float *)p;
  }
  if (hipMalloc((void **)&ctx->d_tables, JXG_QT_BLOB_FLOATS * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&ctx->d_tier_tables,
                (size_t)2 * JXG_QUANT_TIERS * JXG_QT_BLOB_FLOATS * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&ctx->d_tier_selq,
                (size_t)2 * JXG_QUANT_TIERS * 3 * SELQ_N * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&ctx->d_derived, sizeof(jxg_derived)) != hipSuccess ||
      hipHostMalloc((void **)&ctx->h_derived, sizeof(jxg_derived), hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void **)&ctx->d_scans, (64 + 256 + 1024) * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void **)&ctx->d_selq, 3 * SELQ_N * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&ctx->d_mlut, (9 + 256) * sizeof(float)) != hipSuccess ||
      hipHostMalloc((void **)&ctx->h_mlut, (9 + 256) * sizeof(float), hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void **)&ctx->d_misc, 16 * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void **)&ctx->d_exp2, 64 * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&ctx->d_rate_big, JXG_RATE_BIG_N * sizeof(float)) != hipSuccess)
    return fail(set_err(JXG_ERR_ALLOC, "context allocation failed"));
  {
    std::vector<float> big(JXG_RATE_BIG_N);
    for (int i = 0; i < JXG_RATE_BIG_N; i++) big[i] = log2f(1.0f + (float)i);
    if (hipMemcpy(ctx->d_rate_big, big.data(), big.size() * sizeof(float), hipMemcpyHostToDevice) !=
        hipSuccess)
      return fail(set_err(JXG_ERR_DEVICE, "rate table upload failed"));
  }
  if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess)
    return fail(set_err(JXG_ERR_DEVICE, "context stream / event creation failed"));
  // scan position tables of the square shapes, computed once
  uint32_t scans[64 + 256 + 1024];
  {
    uint32_t order[1024];
    const int sides[3] = {1, 2, 4};
    int off = 0;
    for (int i = 0; i < 3; i++) {
      const int cx = sides[i], n = cx * cx * 64;
      natural_scan(cx, cx, order);
      for (int k = 0; k < n; k++) scans[off + k] = order[k];  // scan position -> storage index
      for (int k = 0; k < n; k++) ctx->h_order[off + k] = order[k];
      off += n;
    }
  }
  if (hipMemcpy(ctx->d_scans, scans, sizeof(scans), hipMemcpyHostToDevice) != hipSuccess)
    return fail(set_err(JXG_ERR_DEVICE, "scan table upload failed"));
  {
    // scan position of every storage index, for each coefficient shape
    static const int shapes[8][2] = {{1, 1}, {2, 1}, {2, 2}, {4, 1}, {4, 2}, {4, 4}, {8, 4}, {8, 8}};
    uint16_t *pos = (uint16_t *)malloc(SCANPOS_TOTAL * sizeof(uint16_t));
    uint32_t *order = (uint32_t *)malloc(4096 * sizeof(uint32_t));
    if (!pos || !order) {
      free(pos);
      free(order);
      return fail(set_err(JXG_ERR_ALLOC, "context allocation failed"));
    }
    for (int i = 0; i < 8; i++) {
      const int cx = shapes[i][0], cy = shapes[i][1], n = cx * cy * 64;
      natural_scan(cx, cy, order);
      for (int p = 0; p < n; p++) pos[scanpos_offset(cx, cy) + order[p]] = (uint16_t)p;
    }
    const bool ok = hipMalloc((void **)&ctx->d_scanpos, SCANPOS_TOTAL * sizeof(uint16_t)) == hipSuccess &&
                    hipMemcpy(ctx->d_scanpos, pos, SCANPOS_TOTAL * sizeof(uint16_t),
                              hipMemcpyHostToDevice) == hipSuccess;
    free(pos);
    free(order);
    if (!ok) return fail(set_err(JXG_ERR_DEVICE, "scan position upload failed"));
  }
  {
    float e2[64];
    for (int i = 0; i < 64; i++) e2[i] = (float)exp2((double)(i - 32) / 64.0);
    if (hipMemcpy(ctx->d_exp2, e2, sizeof(e2), hipMemcpyHostToDevice) != hipSuccess)
      return fail(set_err(JXG_ERR_DEVICE, "exp2 table upload failed"));
  }
  {
    const uint32_t zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (hipMemcpy(ctx->d_misc, zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess)
      return fail(set_err(JXG_ERR_DEVICE, "counter initialisation failed"));
  }
  {
    // The default tables of every distance tier and fmla flavour, built now
    // so that no later stage call uploads (one representative distance per
    // tier of jxg_quant_tier).
    static const float tier_distance[JXG_QUANT_TIERS] = {0.1f, 1.0f, 2.0f, 3.0f, 4.0f};
    for (int tier = 0; tier < JXG_QUANT_TIERS; tier++)
      for (int f = 0; f < 2; f++) {
        jxg_frame_params p;
        memset(&p, 0, sizeof(p));
        p.distance = tier_distance[tier];
        p.arbiter = JXG_ARBITER_SSIMULACRA2;
        p.fused = f;
        const float *t, *q;
        const int trc = tables_for(ctx, &p, &t, &q);
        if (trc) return fail(trc);
      }
  }
  *out = ctx;
  return JXG_OK;
}

void jxg_ctx_destroy(jxg_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  else (void)hipDeviceSynchronize();
  if (ctx->side) {
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamDestroy(ctx->side);
  }
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  (void)hipFree(ctx->d_tables);
  (void)hipFree(ctx->d_tier_tables);
  (void)hipFree(ctx->d_tier_selq);
  (void)hipFree(ctx->d_derived);
  (void)hipHostFree(ctx->h_derived);
  (void)hipFree(ctx->d_scans);
  (void)hipFree(ctx->d_selq);
  (void)hipFree(ctx->d_mlut);
  (void)hipHostFree(ctx->h_mlut);
  (void)hipFree(ctx->d_pre);
  (void)hipFree(ctx->d_xyb);
  (void)hipFree(ctx->d_misc);
  (void)hipFree(ctx->d_exp2);
  (void)hipFree(ctx->d_xt);
  (void)hipFree(ctx->d_scanpos);
  (void)hipFree(ctx->d_c128);
  (void)hipFree(ctx->d_wc128);
  (void)hipFree(ctx->d_astat);
  (void)hipFree(ctx->d_alpha);
  (void)hipFree(ctx->d_alpha2);
  (void)hipFree(ctx->d_vis);
  (void)hipFree(ctx->d_fxchg);
  (void)hipFree(ctx->d_afac);
  (void)hipFree(ctx->d_pc);
  (void)hipFree(ctx->d_rf);
  (void)hipFree(ctx->d_c32);
  (void)hipFree(ctx->d_zero_cmap);
  (void)hipFree(ctx->d_rgb);
  (void)hipFree(ctx->d_tmp);
  (void)hipFree(ctx->d_cmap);
  (void)hipFree(ctx->d_c64);
  (void)hipFree(ctx->d_c64x6);
  (void)hipFree(ctx->d_rate_big);
  if (ctx->prof_ev) {
    for (int i = 0; i < JXG_PROF_N * 2 * JXG_PROF_CAP; i++)
      if (ctx->prof_ev[i]) (void)hipEventDestroy(ctx->prof_ev[i]);
    delete[] ctx->prof_ev;
  }
  delete ctx;
}

int jxg_ctx_profile(jxg_ctx *ctx, int enable) {
  if (!ctx) return set_err(JXG_ERR_INVALID_ARG, "null ctx");
  HIP_TRY(hipSetDevice(ctx->device));
  if (enable && !ctx->prof_ev) {
    const int n = JXG_PROF_N * 2 * JXG_PROF_CAP;
    ctx->prof_ev = new hipEvent_t[n]();
    for (int i = 0; i < n; i++) HIP_TRY(hipEventCreate(&ctx->prof_ev[i]));
  }
  for (int k = 0; k < JXG_PROF_N; k++) ctx->prof_n[k] = 0;
  ctx->prof_on = enable != 0;
  return JXG_OK;
}

int jxg_ctx_profile_read(jxg_ctx *ctx, double total_ms[JXG_PROF_N], uint32_t count[JXG_PROF_N]) {
  if (!ctx || !total_ms || !count) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipDeviceSynchronize());
  for (int k = 0; k < JXG_PROF_N; k++) {
    total_ms[k] = 0.0;
    count[k] = ctx->prof_ev ? ctx->prof_n[k] : 0;
    for (uint32_t i = 0; i < count[k]; i++) {
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, ctx->prof_ev[(k * JXG_PROF_CAP + i) * 2],
                                  ctx->prof_ev[(k * JXG_PROF_CAP + i) * 2 + 1]));
      total_ms[k] += ms;
    }
  }
  return JXG_OK;
}

int jxg_ctx_synchronize(jxg_ctx *ctx) {
  if (!ctx) return set_err(JXG_ERR_INVALID_ARG, "null ctx");
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return JXG_OK;
}

int jxg_ctx_set_tables(jxg_ctx *ctx, const float *blob) {
  if (!ctx) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));  // kernels in flight may read the tables
  if (!blob) {
    ctx->tables_set = false;
    return JXG_OK;
  }
  ctx->tables_set = false;
  const int rc = upload_tables(ctx, blob, ctx->d_tables, ctx->d_selq);
  if (rc == JXG_OK) ctx->tables_set = true;
  return rc;
}

int jxg_xyb_dev(jxg_ctx *ctx, const uint8_t *d_rgb, size_t row_bytes, int channels,
                uint32_t width, uint32_t height, const float fwd[9], const float lut[256],
                int order, int fused, float *d_xyb, uint32_t stride, uint32_t padded_h) {
  if (!ctx || !d_rgb || !fwd || !lut || !d_xyb) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  if (channels < 1 || channels > 4) return set_err(JXG_ERR_INVALID_ARG, "channels %d", channels);
  if (row_bytes < (size_t)width * channels) return set_err(JXG_ERR_INVALID_ARG, "row_bytes too small");
  if (order != JXG_XYB_SCALAR && order != JXG_XYB_AVX2) return set_err(JXG_ERR_INVALID_ARG, "order %d", order);
  if (!ctx->mlut_valid || memcmp(ctx->h_mlut, fwd, 9 * sizeof(float)) != 0 ||
      memcmp(ctx->h_mlut + 9, lut, 256 * sizeof(float)) != 0) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));  // staging buffer may be in flight
    memcpy(ctx->h_mlut, fwd, 9 * sizeof(float));
    memcpy(ctx->h_mlut + 9, lut, 256 * sizeof(float));
    HIP_TRY(hipMemcpyAsync(ctx->d_mlut, ctx->h_mlut, 265 * sizeof(float), hipMemcpyHostToDevice,
                           ctx->stream));
    ctx->mlut_valid = true;
  }
  const dim3 grid((stride + XYB_PX_WG - 1) / XYB_PX_WG, (padded_h + XYB_ROWS - 1) / XYB_ROWS);
  if (order == JXG_XYB_AVX2)
    k_xyb<true, JXG_XYB_AVX2><<<grid, 256, 0, ctx->stream>>>(d_rgb, row_bytes, channels, width, height, ctx->d_mlut, d_xyb, stride, padded_h);
  else if (fused)
    k_xyb<true, JXG_XYB_SCALAR><<<grid, 256, 0, ctx->stream>>>(d_rgb, row_bytes, channels, width, height, ctx->d_mlut, d_xyb, stride, padded_h);
  else
    k_xyb<false, JXG_XYB_SCALAR><<<grid, 256, 0, ctx->stream>>>(d_rgb, row_bytes, channels, width, height, ctx->d_mlut, d_xyb, stride, padded_h);
  LAUNCH_CHECK("k_xyb");
  return JXG_OK;
}

// The AQ chain with its pre-map (2 xb x 2 yb floats, every entry written by
// k_aq_pre before k_aq_block reads it) at `pre`.
static int fill_qf_into(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                        uint32_t W, uint32_t H, float distance, float inv_scale, int fused,
                        uint8_t *d_qf, float *pre) {
  const int xb = (W + 7) / 8, yb = (H + 7) / 8;
  const int pre_w = 2 * xb, pre_h = 2 * yb;
  const float *yp = d_xyb + (size_t)stride * padded_h;
  const dim3 g1((4 * pre_w + 255) / 256, pre_h);  // four lanes per pre-map value
  prof_begin(ctx, JXG_PROF_AQ_PRE, ctx->stream);
  if (fused)
    k_aq_pre<true><<<g1, 256, 0, ctx->stream>>>(yp, stride, W, H, pre_w, pre_h, pre);
  else
    k_aq_pre<false><<<g1, 256, 0, ctx->stream>>>(yp, stride, W, H, pre_w, pre_h, pre);
  LAUNCH_CHECK("k_aq_pre");
  prof_end(ctx, JXG_PROF_AQ_PRE, ctx->stream);
  const dim3 g2((xb + AQ_STRIP - 1) / AQ_STRIP, yb);
  prof_begin(ctx, JXG_PROF_AQ_BLOCK, ctx->stream);
  if (fused)
    k_aq_block<true><<<g2, 256, 0, ctx->stream>>>(d_xyb, stride, padded_h, W, H, pre, pre_w, distance, inv_scale, ctx->d_exp2, d_qf, xb, yb);
  else
    k_aq_block<false><<<g2, 256, 0, ctx->stream>>>(d_xyb, stride, padded_h, W, H, pre, pre_w, distance, inv_scale, ctx->d_exp2, d_qf, xb, yb);
  LAUNCH_CHECK("k_aq_block");
  prof_end(ctx, JXG_PROF_AQ_BLOCK, ctx->stream);
  return JXG_OK;
}

static int fill_qf(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h, uint32_t W,
                   uint32_t H, float distance, float inv_scale, int fused, uint8_t *d_qf) {
  const size_t need = (size_t)(2 * ((W + 7) / 8)) * (2 * ((H + 7) / 8));
  int rc = ensure(&ctx->d_pre, &ctx->pre_cap, need);
  if (rc) return rc;
  return fill_qf_into(ctx, d_xyb, stride, padded_h, W, H, distance, inv_scale, fused, d_qf,
                      ctx->d_pre);
}

int jxg_fill_quant_field_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                             uint32_t width, uint32_t height, float distance, float inv_scale,
                             int fused, uint8_t *d_qf) {
  int rc = check_image(width, height);  // EmptyImage before argument checks, as the reference
  if (rc) return rc;                    // validates (src/encode_image.rs:1785-1802)
  if (!ctx || !d_xyb || !d_qf) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  if (!(distance > 0.0f) || !(inv_scale > 0.0f))
    return set_err(JXG_ERR_INVALID_DISTANCE, "distance %g inv_scale %g", (double)distance, (double)inv_scale);
  return fill_qf(ctx, d_xyb, stride, padded_h, width, height, distance, inv_scale, fused, d_qf);
}

int jxg_fill_quant_field_noalloc_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride,
                                     uint32_t padded_h, uint32_t width, uint32_t height,
                                     float distance, float inv_scale, int fused, uint8_t *d_qf) {
  int rc = check_image(width, height);
  if (rc) return rc;
  if (!ctx || !d_xyb || !d_qf) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  if (!(distance > 0.0f) || !(inv_scale > 0.0f))
    return set_err(JXG_ERR_INVALID_DISTANCE, "distance %g inv_scale %g", (double)distance, (double)inv_scale);
  const size_t need = (size_t)(2 * ((width + 7) / 8)) * (2 * ((height + 7) / 8));
  if (need > JXG_NOALLOC_AQ_PRE_FLOATS)
    return set_err(JXG_ERR_TOO_LARGE,
                   "image %ux%u needs a %zu-float AQ pre-map; the no-allocation entry point "
                   "holds at most %u (about 67 megapixels)",
                   width, height, need, (unsigned)JXG_NOALLOC_AQ_PRE_FLOATS);
  if (!ctx->d_noalloc_pre) return set_err(JXG_ERR_DEVICE, "no-allocation scratch not mapped");
  return fill_qf_into(ctx, d_xyb, stride, padded_h, width, height, distance, inv_scale, fused,
                      d_qf, ctx->d_noalloc_pre);
}

static int select_impl(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h, uint32_t W,
                       uint32_t H, uint8_t *d_qf, const int8_t *ytox, const int8_t *ytob,
                       const jxg_frame_params *p, uint8_t *cells, float *c8, float *c16, float *c32) {
  int rc = upload_derived(ctx, p);
  if (rc) return rc;
  if ((rc = zero_cmaps(ctx, W, H, &ytox, &ytob))) return rc;
  const float *tables, *selq;
  if ((rc = tables_for(ctx, p, &tables, &selq))) return rc;
  const int xb = (W + 7) / 8, yb = (H + 7) / 8;
  const dim3 grid((xb + 3) / 4, (yb + 3) / 4);
  // The Fast / Fastest decision re-tiles the merge search's layout
  // (strategy decisions only: the cost dump keeps the search's own field).
  const int retile = cells ? ctx->h_derived->retile : 0;
  float *c32raw = nullptr, *c8raw = nullptr;
  if (retile) {
    const size_t nb = (size_t)xb * yb;
    if ((rc = ensure(&ctx->d_pc, &ctx->pc_cap, 4 * nb))) return rc;
    if ((rc = ensure(&ctx->d_c32, &ctx->c32_cap, nb))) return rc;
    c32raw = ctx->d_c32;
    c8raw = ctx->d_pc;  // the leaf plane of the piece costs
  }
  prof_begin(ctx, JXG_PROF_SELECT, ctx->stream);
  if (p->fused)
    k_select<true><<<grid, SEL_T, 0, ctx->stream>>>(d_xyb, stride, padded_h, W, H, d_qf, ytox, ytob, tables, ctx->d_scans, selq, ctx->d_derived, cells, c8, c16, c32, c32raw, c8raw, xb, yb);
  else
    k_select<false><<<grid, SEL_T, 0, ctx->stream>>>(d_xyb, stride, padded_h, W, H, d_qf, ytox, ytob, tables, ctx->d_scans, selq, ctx->d_derived, cells, c8, c16, c32, c32raw, c8raw, xb, yb);
  LAUNCH_CHECK("k_select");
  prof_end(ctx, JXG_PROF_SELECT, ctx->stream);
  if (retile) {
    if ((rc = retile_stages(ctx, ctx->stream, d_xyb, stride, padded_h, W, H, d_qf, ytox, ytob,
                            tables, p, retile, cells, ctx->d_c32, ctx->d_pc)))
      return rc;
    if (ctx->h_derived->refine_steps)
      return refine_stages(ctx, ctx->stream, d_xyb, stride, padded_h, W, H, d_qf, ytox, ytob,
                           tables, p, cells);
  }
  return JXG_OK;
}

int jxg_strategy_costs_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                           uint32_t width, uint32_t height, const uint8_t *d_qf,
                           const int8_t *d_ytox, const int8_t *d_ytob,
                           const jxg_frame_params *params, float *d_cost8, float *d_cost16,
                           float *d_cost32) {
  if (!ctx || !d_xyb || !d_qf || !d_cost8 || !d_cost16 || !d_cost32)
    return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  // The selection kernel rewrites the quant field it reads; give it a copy.
  const size_t nb = (size_t)((width + 7) / 8) * ((height + 7) / 8);
  uint8_t *tmp = nullptr;
  HIP_TRY(hipMallocAsync((void **)&tmp, nb, ctx->stream));
  HIP_TRY(hipMemcpyAsync(tmp, d_qf, nb, hipMemcpyDeviceToDevice, ctx->stream));
  rc = select_impl(ctx, d_xyb, stride, padded_h, width, height, tmp, d_ytox, d_ytob, params, nullptr,
                   d_cost8, d_cost16, d_cost32);
  (void)hipFreeAsync(tmp, ctx->stream);
  return rc;
}

int jxg_select_strategy_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                            uint32_t width, uint32_t height, uint8_t *d_qf, const int8_t *d_ytox,
                            const int8_t *d_ytob, const jxg_frame_params *params,
                            uint8_t *d_cells) {
  if (!ctx || !d_xyb || !d_qf || !d_cells) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  if ((rc = check_speed(params))) return rc;
  return select_impl(ctx, d_xyb, stride, padded_h, width, height, d_qf, d_ytox, d_ytob, params,
                     d_cells, nullptr, nullptr, nullptr);
}

int jxg_forward_quantize_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                             uint32_t width, uint32_t height, const uint8_t *d_cells,
                             const uint8_t *d_qf, const int8_t *d_ytox, const int8_t *d_ytob,
                             const jxg_frame_params *params, int32_t *d_levels[3],
                             float *d_coeffs[3], int16_t *d_dc[3]) {
  if (!ctx || !d_xyb || !d_cells || !d_qf || !d_levels || !d_coeffs || !d_dc)
    return set_err(JXG_ERR_INVALID_ARG, "null argument");
  for (int c = 0; c < 3; c++)
    if (!d_levels[c] || !d_coeffs[c] || !d_dc[c]) return set_err(JXG_ERR_INVALID_ARG, "null output");
  int rc = check_image(width, height);
  if (rc) return rc;
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  if ((rc = upload_derived(ctx, params))) return rc;
  if ((rc = zero_cmaps(ctx, width, height, &d_ytox, &d_ytob))) return rc;
  const float *tables, *selq;
  if ((rc = tables_for(ctx, params, &tables, &selq))) return rc;
  const int xb = (width + 7) / 8, yb = (height + 7) / 8;
  // At most one crossing transform per 4 blocks (the smallest multi-block
  // transform covers two blocks, and a crossing one at least two).
  const size_t cap = (size_t)xb * yb / 2 + 1;
  if ((rc = ensure(&ctx->d_xt, &ctx->xt_cap, cap))) return rc;
  QArgs A;
  A.xyb = d_xyb;
  A.stride = stride;
  A.padded_h = padded_h;
  A.W = width;
  A.H = height;
  A.xb = xb;
  A.yb = yb;
  A.cells = d_cells;
  A.qf = d_qf;
  A.ytox = d_ytox;
  A.ytob = d_ytob;
  A.tables = tables;
  A.D = ctx->d_derived;
  A.lv0 = d_levels[0];
  A.lv1 = d_levels[1];
  A.lv2 = d_levels[2];
  A.co0 = d_coeffs[0];
  A.co1 = d_coeffs[1];
  A.co2 = d_coeffs[2];
  A.dc0 = d_dc[0];
  A.dc1 = d_dc[1];
  A.dc2 = d_dc[2];
  A.xt_count = ctx->d_misc;  // zero between calls (k_quantize_xt clears it)
  A.bad = ctx->d_misc + 1;
  A.status = ctx->d_misc + 4;
  A.xt_list = ctx->d_xt;
  A.xt_cap = (uint32_t)cap;
  const dim3 grid((xb + 7) / 8, (yb + 7) / 8);
  prof_begin(ctx, JXG_PROF_QUANTIZE, ctx->stream);
  if (params->fused)
    k_quantize<true><<<grid, 256, 0, ctx->stream>>>(A);
  else
    k_quantize<false><<<grid, 256, 0, ctx->stream>>>(A);
  LAUNCH_CHECK("k_quantize");
  prof_end(ctx, JXG_PROF_QUANTIZE, ctx->stream);
  const int xt_grid = 256;
  prof_begin(ctx, JXG_PROF_QUANTIZE_XT, ctx->stream);
  if (params->fused)
    k_quantize_xt<true><<<xt_grid, 256, 0, ctx->stream>>>(A);
  else
    k_quantize_xt<false><<<xt_grid, 256, 0, ctx->stream>>>(A);
  LAUNCH_CHECK("k_quantize_xt");
  prof_end(ctx, JXG_PROF_QUANTIZE_XT, ctx->stream);
  return JXG_OK;
}

int jxg_strategy_cost_sites_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride,
                                uint32_t padded_h, uint32_t width, uint32_t height,
                                const int8_t *d_ytox, const int8_t *d_ytob,
                                const jxg_frame_params *params, int strategy, uint32_t n_sites,
                                const uint32_t *d_sites, const float *d_qac, float meta_r,
                                float *d_cost) {
  if (!ctx || !d_xyb || !params || (n_sites && (!d_sites || !d_qac || !d_cost)))
    return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  if (strategy < 0 || strategy >= JXG_NUM_STRATEGIES || strategy == JXG_DCT128X128 ||
      strategy == JXG_DCT128X64 || strategy == JXG_DCT64X128)
    return set_err(JXG_ERR_INVALID_ARG, "strategy %d has no cost kernel", strategy);
  if (n_sites == 0) return JXG_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = upload_derived(ctx, params))) return rc;
  if ((rc = zero_cmaps(ctx, width, height, &d_ytox, &d_ytob))) return rc;
  const float *tables, *selq;
  if ((rc = tables_for(ctx, params, &tables, &selq))) return rc;
  SiteArgs S;
  S.xyb = d_xyb;
  S.stride = (int)stride;
  S.padded_h = (int)padded_h;
  S.W = (int)width;
  S.H = (int)height;
  S.tiles_x = (int)((width + 63) / 64);
  S.tiles_y = (int)((height + 63) / 64);
  S.ytox = d_ytox;
  S.ytob = d_ytob;
  S.tables = tables;
  S.D = ctx->d_derived;
  S.scanpos = ctx->d_scanpos;
  S.s = strategy;
  S.n = n_sites;
  S.sites = d_sites;
  S.qac = d_qac;
  S.meta_r = meta_r;
  S.cost = d_cost;
  const bool big = is_dct64(strategy);
  if (params->fused) {
    if (big)
      k_site_cost<true, true><<<n_sites, 256, 0, ctx->stream>>>(S);
    else
      k_site_cost<true, false><<<n_sites, 256, 0, ctx->stream>>>(S);
  } else {
    if (big)
      k_site_cost<false, true><<<n_sites, 256, 0, ctx->stream>>>(S);
    else
      k_site_cost<false, false><<<n_sites, 256, 0, ctx->stream>>>(S);
  }
  LAUNCH_CHECK("k_site_cost");
  return JXG_OK;
}

int jxg_front_end_dev(jxg_ctx *ctx, const uint8_t *d_rgb, size_t row_bytes, int channels,
                      uint32_t width, uint32_t height, const jxg_frame_params *params,
                      int8_t *d_ytox, int8_t *d_ytob, uint8_t *d_qf, uint8_t *d_cells,
                      int32_t *d_levels[3], float *d_coeffs[3], int16_t *d_dc[3]) {
  if (!ctx || !d_rgb || !params || !d_qf || !d_cells) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  if ((rc = check_speed(params))) return rc;
  uint32_t stride, ph;
  jxg_plane_geometry(width, height, &stride, &ph);
  if ((rc = ensure(&ctx->d_xyb, &ctx->xyb_cap, (size_t)3 * stride * ph))) return rc;
  const size_t tiles = (size_t)((width + 63) / 64) * ((height + 63) / 64);
  if (!d_ytox || !d_ytob) {
    if ((rc = ensure(&ctx->d_cmap, &ctx->cmap_cap, 2 * tiles))) return rc;
    if (!d_ytox) d_ytox = ctx->d_cmap;
    if (!d_ytob) d_ytob = ctx->d_cmap + tiles;
  }
  float lut[256];
  jxg_srgb_lut(lut);
  prof_begin(ctx, JXG_PROF_XYB, ctx->stream);
  if ((rc = jxg_xyb_dev(ctx, d_rgb, row_bytes, channels, width, height, params->xyb_fwd, lut,
                        params->xyb_order, params->fused, ctx->d_xyb, stride, ph)))
    return rc;
  prof_end(ctx, JXG_PROF_XYB, ctx->stream);
  // Alpha of gray + alpha and RGBA input: fringe clearing, fill_invisible on
  // XYB and the alpha block factors (src/frame/encoder.rs:60-69,152-157).
  // k_xyb already snapped gray input to X = 0, B = Y; the fill is the same
  // weighted mean on every plane, so snapping first gives identical planes.
  const bool has_alpha = channels == 2 || channels == 4;
  ctx->alpha_used = 0;
  if (has_alpha) {
    prof_begin(ctx, JXG_PROF_ALPHA, ctx->stream);
    if ((rc = alpha_front(ctx, d_rgb, row_bytes, channels, width, height, params->flags,
                          ctx->d_xyb, stride, ph)))
      return rc;
    prof_end(ctx, JXG_PROF_ALPHA, ctx->stream);
    ctx->alpha_used = (size_t)width * height;
  }
  if ((rc = upload_derived(ctx, params))) return rc;
  // The CfL slopes (needed before the strategy search, as setup_dc_group
  // computes them, src/frame/dc_groups.rs:166-184) depend only on XYB: they
  // run on the side stream beside the quant field.
  HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream));
  HIP_TRY(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
  prof_begin(ctx, JXG_PROF_CMAP, ctx->side);
  if ((rc = fill_cmap(ctx, ctx->d_xyb, stride, ph, width, height, params, d_ytox, d_ytob,
                      ctx->side)))
    return rc;
  prof_end(ctx, JXG_PROF_CMAP, ctx->side);
  HIP_TRY(hipEventRecord(ctx->ev_join, ctx->side));
  if ((rc = fill_qf(ctx, ctx->d_xyb, stride, ph, width, height, params->distance,
                    ctx->h_derived->inv_scale, params->fused, d_qf)))
    return rc;
  if ((params->flags & JXG_FLAG_CHROMA_HF) &&
      (rc = chroma_hf(ctx, ctx->d_xyb, stride, ph, width, height, params->distance, params->fused,
                      d_qf)))
    return rc;
  HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
  if ((rc = jxg_select_strategy_dev(ctx, ctx->d_xyb, stride, ph, width, height, d_qf, d_ytox, d_ytob,
                                    params, d_cells)))
    return rc;
  // once every layout and quant-field decision is made (src/frame/dc_groups.rs:302-312)
  if (has_alpha && (params->flags & JXG_FLAG_FILL_INVISIBLE)) {
    prof_begin(ctx, JXG_PROF_ALPHA_QF, ctx->stream);
    if ((rc = alpha_finish(ctx, width, height, d_cells, d_qf))) return rc;
    prof_end(ctx, JXG_PROF_ALPHA_QF, ctx->stream);
  }
  return jxg_forward_quantize_dev(ctx, ctx->d_xyb, stride, ph, width, height, d_cells, d_qf, d_ytox,
                                  d_ytob, params, d_levels, d_coeffs, d_dc);
}

int jxg_front_end(jxg_ctx *ctx, const uint8_t *rgb, size_t row_bytes, int channels, uint32_t width,
                  uint32_t height, const jxg_frame_params *params, int8_t *ytox, int8_t *ytob,
                  uint8_t *qf, uint8_t *cells, int32_t *levels[3], float *coeffs[3],
                  int16_t *dc[3]) {
  if (!ctx || !rgb || !params || !qf || !cells || !levels || !coeffs || !dc)
    return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  if (channels < 1 || channels > 4) return set_err(JXG_ERR_INVALID_ARG, "channels %d", channels);
  const size_t xb = (width + 7) / 8, yb = (height + 7) / 8, nb = xb * yb;
  const size_t tiles = (size_t)((width + 63) / 64) * ((height + 63) / 64);
  const size_t in_bytes = (size_t)height * row_bytes;
  if ((rc = ensure(&ctx->d_rgb, &ctx->rgb_cap, in_bytes))) return rc;
  // one device block for every output: levels, coeffs, dc, qf, cells, maps
  const size_t bytes = nb * 64 * 3 * (sizeof(int32_t) + sizeof(float)) + nb * 3 * sizeof(int16_t) +
                       2 * nb + 2 * tiles + 256;
  uint8_t *dev = nullptr;
  HIP_TRY(hipMalloc((void **)&dev, bytes));
  uint8_t *p = dev;
  int32_t *d_lv[3];
  float *d_co[3];
  int16_t *d_dc[3];
  for (int c = 0; c < 3; c++) {
    d_lv[c] = (int32_t *)p;
    p += nb * 64 * sizeof(int32_t);
  }
  for (int c = 0; c < 3; c++) {
    d_co[c] = (float *)p;
    p += nb * 64 * sizeof(float);
  }
  for (int c = 0; c < 3; c++) {
    d_dc[c] = (int16_t *)p;
    p += nb * sizeof(int16_t);
  }
  uint8_t *d_qf = p;
  p += nb;
  uint8_t *d_cells = p;
  p += nb;
  int8_t *d_tx = (int8_t *)p;
  p += tiles;
  int8_t *d_tb = (int8_t *)p;
  auto run = [&]() -> int {
    HIP_TRY(hipMemcpyAsync(ctx->d_rgb, rgb, in_bytes, hipMemcpyHostToDevice, ctx->stream));
    int r = jxg_front_end_dev(ctx, ctx->d_rgb, row_bytes, channels, width, height, params, d_tx,
                              d_tb, d_qf, d_cells, d_lv, d_co, d_dc);
    if (r) return r;
    for (int c = 0; c < 3; c++) {
      HIP_TRY(hipMemcpyAsync(levels[c], d_lv[c], nb * 64 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(hipMemcpyAsync(coeffs[c], d_co[c], nb * 64 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(hipMemcpyAsync(dc[c], d_dc[c], nb * sizeof(int16_t), hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(hipMemcpyAsync(qf, d_qf, nb, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(cells, d_cells, nb, hipMemcpyDeviceToHost, ctx->stream));
    if (ytox) HIP_TRY(hipMemcpyAsync(ytox, d_tx, tiles, hipMemcpyDeviceToHost, ctx->stream));
    if (ytob) HIP_TRY(hipMemcpyAsync(ytob, d_tb, tiles, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return JXG_OK;
  };
  rc = run();
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(dev);
  return rc;
}

// Device-side status of the last jxg_forward_quantize_dev call:
// out[0] = transforms that crossed their tile, out[1] = 1 if an unsupported
// strategy cell was seen (its blocks are left unwritten).  Synchronizes.
int jxg_quantize_status(jxg_ctx *ctx, uint32_t out[2]) {
  if (!ctx || !out) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  uint32_t h[2];
  HIP_TRY(hipMemcpyAsync(h, ctx->d_misc + 4, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  out[0] = h[0];
  out[1] = h[1];
  return JXG_OK;
}

}  // extern "C"

