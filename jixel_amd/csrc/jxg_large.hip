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
  if (method != JXG_DCT64_PLAIN && method != JXG_DCT64_MFMA_BF16X3)
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

}  // extern "C"
