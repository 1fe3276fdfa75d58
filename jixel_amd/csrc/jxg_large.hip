// jxg_large.hip - the DCT64X64 transform of every 64x64 block of an image,
// in two forms:
//
//   k_dct64_plain  the reference's FP32 butterflies (dct64x64_scalar_input,
//                  src/dct.rs:2549-2573) in its operation order: the block
//                  in LDS, 64 column lines then 64 row lines (dct64_line).
//                  Bit-identical to the quantize stage's DCT64X64.
//   k_dct64_mfma   Y = C X C^T / 4096 as two 64x64x64 products on the bf16
//                  matrix cores with split precision: every f32 operand a is
//                  held as a_hi + a_lo (two bf16), each product takes the
//                  three terms hi*hi + hi*lo + lo*hi (the lo*lo term is below
//                  2^-16 relative) with f32 accumulation.  C is the scaled
//                  DCT-II matrix (c_0 = 1, c_k = sqrt 2), split on the host
//                  from its FP64 values.  One workgroup of four waves per
//                  (block, channel); each wave owns a 32x32 quadrant and runs
//                  v_mfma_f32_32x32x16_bf16 (4 K-steps x 3 terms per product).
//
// Both write the DCT64X64 coefficient storage (out[u * 64 + v], u the
// horizontal frequency) in the chunk layout of the quantize outputs, with
// xblocks = 8 * ceil(W / 64).
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <vector>

#include "jxg_ctx.h"
#include "jxg_device.h"

using namespace jxg;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int LP = 72;  // bf16 row pitch of the LDS operand tiles (16-byte aligned rows)

__device__ __forceinline__ size_t out_index(int xb, int bx, int by, int k) {
  // chunk j = k >> 9 of a transform covering 8 blocks in x (src/group.rs layout)
  return ((size_t)(by + (k >> 9)) * xb + bx) * 64 + (k & 511);
}

template <bool F>
__global__ __launch_bounds__(256) void k_dct64_plain(const float *__restrict__ xyb, int stride,
                                                     int padded_h, int xb, float *out0,
                                                     float *out1, float *out2) {
  __shared__ float t[64 * 65];
  const int tid = threadIdx.x, c = blockIdx.z;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 64;
  const float *src = xyb + (size_t)c * stride * padded_h;
  for (int e = tid; e < 64 * 16; e += 256) {
    const int r = e >> 4, q = e & 15;
    const float4 v = *reinterpret_cast<const float4 *>(src + (size_t)(y0 + r) * stride + x0 + 4 * q);
    float *d = &t[r * 65 + 4 * q];
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
    d[3] = v.w;
  }
  __syncthreads();
  if (tid < 64) dct64_line<F>(&t[tid], 65, 0.0f, false);  // column x = tid
  __syncthreads();
  if (tid < 64) dct64_line<F>(&t[tid * 65], 1, 1.0f / 4096.0f, true);  // row of frequency v
  __syncthreads();
  float *out = c == 0 ? out0 : (c == 1 ? out1 : out2);
  const int bx = blockIdx.x * 8, by = blockIdx.y * 8;
  for (int k = tid; k < 4096; k += 256) {
    const int u = k >> 6, v = k & 63;
    out[out_index(xb, bx, by, k)] = t[dct64_pos(v) * 65 + dct64_pos(u)];
  }
}

__device__ __forceinline__ bf16x8 lds8(const __bf16 *p) {
  return *reinterpret_cast<const bf16x8 *>(p);
}

// acc += A(32 x 64) * B(64 x 32) with split operands.  A row r of the
// quadrant is a_hi / a_lo + r * LP (K contiguous); B column n is
// b_hi / b_lo + n * LP (K contiguous, i.e. B stored transposed).
__device__ __forceinline__ f32x16 gemm_split(const __bf16 *a_hi, const __bf16 *a_lo,
                                             const __bf16 *b_hi, const __bf16 *b_lo) {
  const int l = threadIdx.x & 63;
  const int r = l & 31, kg = (l >> 5) * 8;
  f32x16 acc = {};
#pragma unroll
  for (int kk = 0; kk < 64; kk += 16) {
    const bf16x8 ah = lds8(a_hi + r * LP + kk + kg), al = lds8(a_lo + r * LP + kk + kg);
    const bf16x8 bh = lds8(b_hi + r * LP + kk + kg), bl = lds8(b_lo + r * LP + kk + kg);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
  }
  return acc;
}

__global__ __launch_bounds__(256) void k_dct64_mfma(const float *__restrict__ xyb, int stride,
                                                    int padded_h, int xb,
                                                    const __bf16 *__restrict__ cmat,
                                                    float *out0, float *out1, float *out2) {
  // X split (row y, K = x), C split (row k, K = i), T^T split (row u, K = y).
  __shared__ __attribute__((aligned(16))) __bf16 xh[64 * LP], xl[64 * LP];
  __shared__ __attribute__((aligned(16))) __bf16 ch[64 * LP], cl[64 * LP];
  __shared__ __attribute__((aligned(16))) __bf16 th[64 * LP], tl[64 * LP];
  const int tid = threadIdx.x, c = blockIdx.z;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 64;
  const float *src = xyb + (size_t)c * stride * padded_h;
  for (int e = tid; e < 64 * 16; e += 256) {
    const int r = e >> 4, q = e & 15;
    const float4 v = *reinterpret_cast<const float4 *>(src + (size_t)(y0 + r) * stride + x0 + 4 * q);
    const float f[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const __bf16 h = (__bf16)f[i];
      xh[r * LP + 4 * q + i] = h;
      xl[r * LP + 4 * q + i] = (__bf16)(f[i] - (float)h);
    }
  }
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e >> 6, q = e & 63;
    ch[r * LP + q] = cmat[e];
    cl[r * LP + q] = cmat[4096 + e];
  }
  __syncthreads();
  const int w = tid >> 6, l = tid & 63;
  const int qr = (w >> 1) * 32, qc = (w & 1) * 32;
  // T[y][u] = sum_x X[y][x] C[u][x]: A = X rows qr.., B columns = C rows qc..
  f32x16 acc = gemm_split(xh + qr * LP, xl + qr * LP, ch + qc * LP, cl + qc * LP);
  // D layout: element i is row (i / 4) * 8 + (l / 32) * 4 + i % 4, column l % 32
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int y = qr + (i >> 2) * 8 + (l >> 5) * 4 + (i & 3), u = qc + (l & 31);
    const float v = acc[i];
    const __bf16 h = (__bf16)v;
    th[u * LP + y] = h;  // stored transposed: row u, K = y
    tl[u * LP + y] = (__bf16)(v - (float)h);
  }
  __syncthreads();
  // Y[v][u] = sum_y C[v][y] T[y][u]: A = C rows qr.., B columns = T^T rows qc..
  acc = gemm_split(ch + qr * LP, cl + qr * LP, th + qc * LP, tl + qc * LP);
  float *out = c == 0 ? out0 : (c == 1 ? out1 : out2);
  const int bx = blockIdx.x * 8, by = blockIdx.y * 8;
  const int u = qc + (l & 31);
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const int v = qr + g * 8 + (l >> 5) * 4;  // four consecutive v, one u
    const int k = u * 64 + v;
    float4 o;
    o.x = acc[g * 4 + 0] * (1.0f / 4096.0f);
    o.y = acc[g * 4 + 1] * (1.0f / 4096.0f);
    o.z = acc[g * 4 + 2] * (1.0f / 4096.0f);
    o.w = acc[g * 4 + 3] * (1.0f / 4096.0f);
    *reinterpret_cast<float4 *>(out + out_index(xb, bx, by, k)) = o;
  }
}

// acc += A(32 x 64) * B(64 x 32) with each operand split in three bf16
// terms (hi + mid + lo holds an f32 exactly): the six products above 2^-24
// of the leading one (lo*hi, mid*mid, hi*lo, mid*hi, hi*mid, hi*hi, the
// smallest first), f32 accumulation.
__device__ __forceinline__ f32x16 gemm_split3(const __bf16 *a0, const __bf16 *a1,
                                              const __bf16 *a2, const __bf16 *b0,
                                              const __bf16 *b1, const __bf16 *b2) {
  const int l = threadIdx.x & 63;
  const int r = l & 31, kg = (l >> 5) * 8;
  f32x16 acc = {};
#pragma unroll
  for (int kk = 0; kk < 64; kk += 16) {
    const int o = r * LP + kk + kg;
    const bf16x8 ah = lds8(a0 + o), am = lds8(a1 + o), al = lds8(a2 + o);
    const bf16x8 bh = lds8(b0 + o), bm = lds8(b1 + o), bl = lds8(b2 + o);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bm, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bm, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
  }
  return acc;
}

// f = h + m + l exactly for a finite f32 (8 significand bits each).
__device__ __forceinline__ void split3(float f, __bf16 &h, __bf16 &m, __bf16 &l) {
  h = (__bf16)f;
  const float r1 = f - (float)h;
  m = (__bf16)r1;
  l = (__bf16)(r1 - (float)m);
}

// Y = C X C^T / 4096 with three-term operands (JXG_DCT64_MFMA_BF16X6).  The
// X tiles are reused for T^T once the first product has read them.
__global__ __launch_bounds__(256) void k_dct64_mfma6(const float *__restrict__ xyb, int stride,
                                                     int padded_h, int xb,
                                                     const __bf16 *__restrict__ cmat,
                                                     float *out0, float *out1, float *out2) {
  __shared__ __attribute__((aligned(16))) __bf16 x0s[64 * LP], x1s[64 * LP], x2s[64 * LP];
  __shared__ __attribute__((aligned(16))) __bf16 c0s[64 * LP], c1s[64 * LP], c2s[64 * LP];
  const int tid = threadIdx.x, c = blockIdx.z;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 64;
  const float *src = xyb + (size_t)c * stride * padded_h;
  for (int e = tid; e < 64 * 16; e += 256) {
    const int r = e >> 4, q = e & 15;
    const float4 v = *reinterpret_cast<const float4 *>(src + (size_t)(y0 + r) * stride + x0 + 4 * q);
    const float f[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int o = r * LP + 4 * q + i;
      split3(f[i], x0s[o], x1s[o], x2s[o]);
    }
  }
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e >> 6, q = e & 63;
    c0s[r * LP + q] = cmat[e];
    c1s[r * LP + q] = cmat[4096 + e];
    c2s[r * LP + q] = cmat[8192 + e];
  }
  __syncthreads();
  const int w = tid >> 6, l = tid & 63;
  const int qr = (w >> 1) * 32, qc = (w & 1) * 32;
  // T[y][u] = sum_x X[y][x] C[u][x]
  f32x16 acc = gemm_split3(x0s + qr * LP, x1s + qr * LP, x2s + qr * LP, c0s + qc * LP,
                           c1s + qc * LP, c2s + qc * LP);
  __syncthreads();  // every wave has read X
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int y = qr + (i >> 2) * 8 + (l >> 5) * 4 + (i & 3), u = qc + (l & 31);
    const int o = u * LP + y;  // T^T: row u, K = y
    split3(acc[i], x0s[o], x1s[o], x2s[o]);
  }
  __syncthreads();
  // Y[v][u] = sum_y C[v][y] T^T[u][y]
  acc = gemm_split3(c0s + qr * LP, c1s + qr * LP, c2s + qr * LP, x0s + qc * LP, x1s + qc * LP,
                    x2s + qc * LP);
  float *out = c == 0 ? out0 : (c == 1 ? out1 : out2);
  const int bx = blockIdx.x * 8, by = blockIdx.y * 8;
  const int u = qc + (l & 31);
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const int v = qr + g * 8 + (l >> 5) * 4;
    const int k = u * 64 + v;
    float4 o;
    o.x = acc[g * 4 + 0] * (1.0f / 4096.0f);
    o.y = acc[g * 4 + 1] * (1.0f / 4096.0f);
    o.z = acc[g * 4 + 2] * (1.0f / 4096.0f);
    o.w = acc[g * 4 + 3] * (1.0f / 4096.0f);
    *reinterpret_cast<float4 *>(out + out_index(xb, bx, by, k)) = o;
  }
}

// ---------------------------------------------------------------------------
// DCT128X128 (the 128 family's square member) in the same two forms.

// dct1d_128 (src/dct128.rs:330-352) on a line in LDS (stride s), in place:
// the sum / difference halves, two dct1d_64 and the odd half's running
// sums, with frequency f left at dct128_pos(f) (the reference's final
// even / odd interleave folded into the output index).
__host__ __device__ __forceinline__ int dct128_pos(int f) {
  return (f & 1) ? 64 + dct64_pos(f >> 1) : dct64_pos(f >> 1);
}

template <bool F>
__device__ __forceinline__ void dct128_line(float *p, int s, const float *wc, float scale,
                                            bool do_scale) {
  // positions {i, 127 - i, 63 - i, 64 + i} are read and written by one group
  for (int i = 0; i < 32; i++) {
    const int j = 63 - i;
    const float a = p[i * s], b = p[(127 - i) * s], c = p[j * s], d = p[(127 - j) * s];
    p[i * s] = a + b;
    p[(64 + i) * s] = (a - b) * wc[i];
    p[j * s] = c + d;
    p[(64 + j) * s] = (c - d) * wc[j];
  }
  dct64_line<F>(p, s, 0.0f, false);
  float *q = p + 64 * s;
  dct64_line<F>(q, s, 0.0f, false);
  q[dct64_pos(0) * s] = fmla<F>(q[dct64_pos(0) * s], JXG_SQRT2, q[dct64_pos(1) * s]);
  for (int k = 1; k < 63; k++) q[dct64_pos(k) * s] += q[dct64_pos(k + 1) * s];
  if (do_scale)
    for (int i = 0; i < 128; i++) p[i * s] = p[i * s] * scale;
}

__device__ __forceinline__ size_t out_index128(int xb, int bx, int by, int k) {
  // chunk j = k >> 10 of a transform covering 16 blocks in x
  return ((size_t)(by + (k >> 10)) * xb + bx) * 64 + (k & 1023);
}

constexpr int P128 = 129;

template <bool F>
__global__ __launch_bounds__(128) void k_dct128_plain(const float *__restrict__ xyb, int stride,
                                                      int padded_h, int xb,
                                                      const float *__restrict__ wc_g,
                                                      float *out0, float *out1, float *out2) {
  __shared__ float t[128 * P128];
  __shared__ float wc[64];
  const int tid = threadIdx.x, c = blockIdx.z;
  const int x0 = blockIdx.x * 128, y0 = blockIdx.y * 128;
  const float *src = xyb + (size_t)c * stride * padded_h;
  if (tid < 64) wc[tid] = wc_g[tid];
  for (int e = tid; e < 128 * 32; e += 128) {
    const int r = e >> 5, q = e & 31;
    const float4 v = *reinterpret_cast<const float4 *>(src + (size_t)(y0 + r) * stride + x0 + 4 * q);
    float *d = &t[r * P128 + 4 * q];
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
    d[3] = v.w;
  }
  __syncthreads();
  dct128_line<F>(&t[tid], P128, wc, 0.0f, false);  // column x = tid
  __syncthreads();
  dct128_line<F>(&t[tid * P128], 1, wc, 1.0f / 16384.0f, true);  // LDS row tid
  __syncthreads();
  float *out = c == 0 ? out0 : (c == 1 ? out1 : out2);
  const int bx = blockIdx.x * 16, by = blockIdx.y * 16;
  for (int k = tid; k < 16384; k += 128) {
    const int u = k >> 7, v = k & 127;
    out[out_index128(xb, bx, by, k)] = t[dct128_pos(v) * P128 + dct128_pos(u)];
  }
}

constexpr int LP128 = 136;  // bf16 row pitch of the 128-wide operand tiles

// acc[ti][tj] += A(rows ar + 32 ti, K) * B(rows br + 32 tj, K)^T over K = 128,
// split operands, three products per K-step.
__device__ __forceinline__ void gemm128_split(const __bf16 *a_hi, const __bf16 *a_lo,
                                              const __bf16 *b_hi, const __bf16 *b_lo,
                                              f32x16 acc[2][2]) {
  const int l = threadIdx.x & 63;
  const int r = l & 31, kg = (l >> 5) * 8;
#pragma unroll 2
  for (int kk = 0; kk < 128; kk += 16) {
    bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      ah[t] = lds8(a_hi + (32 * t + r) * LP128 + kk + kg);
      al[t] = lds8(a_lo + (32 * t + r) * LP128 + kk + kg);
      bh[t] = lds8(b_hi + (32 * t + r) * LP128 + kk + kg);
      bl[t] = lds8(b_lo + (32 * t + r) * LP128 + kk + kg);
    }
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
      for (int tj = 0; tj < 2; tj++) {
        acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[ti], bh[tj], acc[ti][tj], 0, 0, 0);
        acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[ti], bl[tj], acc[ti][tj], 0, 0, 0);
        acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[ti], bh[tj], acc[ti][tj], 0, 0, 0);
      }
  }
}

// Y = C X C^T / 16384 on the bf16 matrix cores, split operands.  Four
// waves, each a 64x64 quadrant as 2x2 tiles of 32x32.  The X tiles are
// reused for T^T once the first product has read them.
__global__ __launch_bounds__(256) void k_dct128_mfma(const float *__restrict__ xyb, int stride,
                                                     int padded_h, int xb,
                                                     const __bf16 *__restrict__ cmat,
                                                     float *out0, float *out1, float *out2) {
  __shared__ __attribute__((aligned(16))) __bf16 xh[128 * LP128], xl[128 * LP128];
  __shared__ __attribute__((aligned(16))) __bf16 ch[128 * LP128], cl[128 * LP128];
  const int tid = threadIdx.x, c = blockIdx.z;
  const int x0 = blockIdx.x * 128, y0 = blockIdx.y * 128;
  const float *src = xyb + (size_t)c * stride * padded_h;
  for (int e = tid; e < 128 * 32; e += 256) {
    const int r = e >> 5, q = e & 31;
    const float4 v = *reinterpret_cast<const float4 *>(src + (size_t)(y0 + r) * stride + x0 + 4 * q);
    const float f[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const __bf16 h = (__bf16)f[i];
      xh[r * LP128 + 4 * q + i] = h;
      xl[r * LP128 + 4 * q + i] = (__bf16)(f[i] - (float)h);
    }
  }
  for (int e = tid; e < 128 * 128; e += 256) {
    const int r = e >> 7, q = e & 127;
    ch[r * LP128 + q] = cmat[e];
    cl[r * LP128 + q] = cmat[16384 + e];
  }
  __syncthreads();
  const int w = tid >> 6, l = tid & 63;
  const int qr = (w >> 1) * 64, qc = (w & 1) * 64;
  f32x16 acc[2][2] = {};
  // T[y][u] = sum_x X[y][x] C[u][x]
  gemm128_split(xh + qr * LP128, xl + qr * LP128, ch + qc * LP128, cl + qc * LP128, acc);
  __syncthreads();  // every wave has read X
#pragma unroll
  for (int ti = 0; ti < 2; ti++)
#pragma unroll
    for (int tj = 0; tj < 2; tj++)
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int y = qr + 32 * ti + (i >> 2) * 8 + (l >> 5) * 4 + (i & 3);
        const int u = qc + 32 * tj + (l & 31);
        const float v = acc[ti][tj][i];
        const __bf16 h = (__bf16)v;
        xh[u * LP128 + y] = h;  // T^T: row u, K = y
        xl[u * LP128 + y] = (__bf16)(v - (float)h);
      }
  __syncthreads();
  f32x16 acc2[2][2] = {};
  // Y[v][u] = sum_y C[v][y] T^T[u][y]
  gemm128_split(ch + qr * LP128, cl + qr * LP128, xh + qc * LP128, xl + qc * LP128, acc2);
  float *out = c == 0 ? out0 : (c == 1 ? out1 : out2);
  const int bx = blockIdx.x * 16, by = blockIdx.y * 16;
#pragma unroll
  for (int ti = 0; ti < 2; ti++)
#pragma unroll
    for (int tj = 0; tj < 2; tj++) {
      const int u = qc + 32 * tj + (l & 31);
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const int v = qr + 32 * ti + g * 8 + (l >> 5) * 4;  // four consecutive v, one u
        const int k = u * 128 + v;
        float4 o;
        o.x = acc2[ti][tj][g * 4 + 0] * (1.0f / 16384.0f);
        o.y = acc2[ti][tj][g * 4 + 1] * (1.0f / 16384.0f);
        o.z = acc2[ti][tj][g * 4 + 2] * (1.0f / 16384.0f);
        o.w = acc2[ti][tj][g * 4 + 3] * (1.0f / 16384.0f);
        *reinterpret_cast<float4 *>(out + out_index128(xb, bx, by, k)) = o;
      }
    }
}

// Round-to-nearest-even f32 -> bf16 bits (finite inputs) and back.
uint16_t bf16_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float bf16_value(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

}  // namespace

extern "C" {

int jxg_dct64_image_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                        uint32_t width, uint32_t height, int method, int fused,
                        float *d_out[3]) {
  if (!ctx || !d_xyb || !d_out || !d_out[0] || !d_out[1] || !d_out[2])
    return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  if (method != JXG_DCT64_PLAIN && method != JXG_DCT64_MFMA_BF16X3 &&
      method != JXG_DCT64_MFMA_BF16X6)
    return set_err(JXG_ERR_INVALID_ARG, "unknown DCT64 method %d", method);
  HIP_TRY(hipSetDevice(ctx->device));
  const int tx = (int)((width + 63) / 64), ty = (int)((height + 63) / 64);
  const int xb = 8 * tx;
  const dim3 grid(tx, ty, 3);
  if (method == JXG_DCT64_PLAIN) {
    if (fused)
      k_dct64_plain<true><<<grid, 256, 0, ctx->stream>>>(d_xyb, (int)stride, (int)padded_h, xb,
                                                         d_out[0], d_out[1], d_out[2]);
    else
      k_dct64_plain<false><<<grid, 256, 0, ctx->stream>>>(d_xyb, (int)stride, (int)padded_h, xb,
                                                          d_out[0], d_out[1], d_out[2]);
    LAUNCH_CHECK("k_dct64_plain");
    return JXG_OK;
  }
  if (method == JXG_DCT64_MFMA_BF16X6) {
    if (!ctx->d_c64x6) {
      // The scaled DCT-II matrix in FP64, split into bf16 hi + mid + lo.
      std::vector<uint16_t> h(3 * 4096);
      for (int k = 0; k < 64; k++)
        for (int i = 0; i < 64; i++) {
          const double ck = k == 0 ? 1.0 : sqrt(2.0);
          double v = ck * cos(M_PI * k * (2 * i + 1) / 128.0);
          for (int t = 0; t < 3; t++) {
            const uint16_t b = bf16_bits((float)v);
            h[t * 4096 + k * 64 + i] = b;
            v -= (double)bf16_value(b);
          }
        }
      HIP_TRY(hipMalloc((void **)&ctx->d_c64x6, h.size() * sizeof(uint16_t)));
      HIP_TRY(hipMemcpy(ctx->d_c64x6, h.data(), h.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    }
    k_dct64_mfma6<<<grid, 256, 0, ctx->stream>>>(d_xyb, (int)stride, (int)padded_h, xb,
                                                 (const __bf16 *)ctx->d_c64x6, d_out[0], d_out[1],
                                                 d_out[2]);
    LAUNCH_CHECK("k_dct64_mfma6");
    return JXG_OK;
  }
  if (!ctx->d_c64) {
    // The scaled DCT-II matrix in FP64, split into bf16 hi + lo.
    uint16_t h[2 * 4096];
    for (int k = 0; k < 64; k++)
      for (int i = 0; i < 64; i++) {
        const double ck = k == 0 ? 1.0 : sqrt(2.0);
        const double v = ck * cos(M_PI * k * (2 * i + 1) / 128.0);
        const uint16_t hi = bf16_bits((float)v);
        h[k * 64 + i] = hi;
        h[4096 + k * 64 + i] = bf16_bits((float)(v - (double)bf16_value(hi)));
      }
    HIP_TRY(hipMalloc((void **)&ctx->d_c64, sizeof(h)));
    HIP_TRY(hipMemcpy(ctx->d_c64, h, sizeof(h), hipMemcpyHostToDevice));
  }
  k_dct64_mfma<<<grid, 256, 0, ctx->stream>>>(d_xyb, (int)stride, (int)padded_h, xb,
                                              (const __bf16 *)ctx->d_c64, d_out[0], d_out[1],
                                              d_out[2]);
  LAUNCH_CHECK("k_dct64_mfma");
  return JXG_OK;
}

int jxg_dct128_image_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                         uint32_t width, uint32_t height, int method, int fused,
                         float *d_out[3]) {
  if (!ctx || !d_xyb || !d_out || !d_out[0] || !d_out[1] || !d_out[2])
    return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  const int tx = (int)((width + 127) / 128), ty = (int)((height + 127) / 128);
  if (stride < (uint32_t)tx * 128 || padded_h < (uint32_t)ty * 128 || (stride & 3))
    return set_err(JXG_ERR_INVALID_ARG,
                   "DCT128 planes need stride >= %d and padded_h >= %d (128-px blocks)", tx * 128,
                   ty * 128);
  if (method != JXG_DCT64_PLAIN && method != JXG_DCT64_MFMA_BF16X3)
    return set_err(JXG_ERR_INVALID_ARG, "unknown DCT128 method %d", method);
  HIP_TRY(hipSetDevice(ctx->device));
  const int xb = 16 * tx;
  const dim3 grid(tx, ty, 3);
  if (method == JXG_DCT64_PLAIN) {
    if (!ctx->d_wc128) {
      float wc[64];  // 1 / (2 cos((2i+1) pi / 256)), the reference's WC128 values
      for (int i = 0; i < 64; i++) wc[i] = (float)(0.5 / cos((2 * i + 1) * M_PI / 256.0));
      HIP_TRY(hipMalloc((void **)&ctx->d_wc128, sizeof(wc)));
      HIP_TRY(hipMemcpy(ctx->d_wc128, wc, sizeof(wc), hipMemcpyHostToDevice));
    }
    if (fused)
      k_dct128_plain<true><<<grid, 128, 0, ctx->stream>>>(d_xyb, (int)stride, (int)padded_h, xb,
                                                          ctx->d_wc128, d_out[0], d_out[1],
                                                          d_out[2]);
    else
      k_dct128_plain<false><<<grid, 128, 0, ctx->stream>>>(d_xyb, (int)stride, (int)padded_h, xb,
                                                           ctx->d_wc128, d_out[0], d_out[1],
                                                           d_out[2]);
    LAUNCH_CHECK("k_dct128_plain");
    return JXG_OK;
  }
  if (!ctx->d_c128) {
    std::vector<uint16_t> h(2 * 16384);
    for (int k = 0; k < 128; k++)
      for (int i = 0; i < 128; i++) {
        const double ck = k == 0 ? 1.0 : sqrt(2.0);
        const double v = ck * cos(M_PI * k * (2 * i + 1) / 256.0);
        const uint16_t hi = bf16_bits((float)v);
        h[k * 128 + i] = hi;
        h[16384 + k * 128 + i] = bf16_bits((float)(v - (double)bf16_value(hi)));
      }
    HIP_TRY(hipMalloc((void **)&ctx->d_c128, h.size() * sizeof(uint16_t)));
    HIP_TRY(hipMemcpy(ctx->d_c128, h.data(), h.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  }
  k_dct128_mfma<<<grid, 256, 0, ctx->stream>>>(d_xyb, (int)stride, (int)padded_h, xb,
                                               (const __bf16 *)ctx->d_c128, d_out[0], d_out[1],
                                               d_out[2]);
  LAUNCH_CHECK("k_dct128_mfma");
  return JXG_OK;
}

}  // extern "C"
