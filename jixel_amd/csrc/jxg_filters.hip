// jxg_filters.hip - whole-plane filters of the front end.
//
//   k_gaborish  gaborish pre-inversion (src/gaborish.rs:44-279): normalised
//               symmetric 5x5 kernel with mirror-reflected edges.  One
//               256-thread workgroup per 64x16 output tile of one plane; the
//               tile and its 2-px halo are staged in LDS, each thread writes
//               four rows of one column.  Output is out of place, so every
//               output sees source pixels, as the reference's delayed row
//               ring guarantees (src/gaborish.rs:94-135).
#include <float.h>
#include <hip/hip_runtime.h>

#include "jxg_ctx.h"
#include "jxg_device.h"

using namespace jxg;

namespace {

constexpr int GAB_TW = 64, GAB_TH = 16;
constexpr int GAB_LW = GAB_TW + 4, GAB_LH = GAB_TH + 4;

// reflect (src/gaborish.rs:80-91).
__host__ __device__ __forceinline__ int gab_reflect(int i, int len) {
  int j = i;
  while (j < 0 || j >= len) j = j < 0 ? -j - 1 : 2 * len - j - 1;
  return j;
}

struct GabWeights {
  float c, r, d, r2, l, d2;
};

// The weights of gaborish_inverse (src/gaborish.rs:44-70), in f32 and in the
// reference's order.
GabWeights gab_weights(float mul) {
  const float K0 = -0.09495815671340026f, K1 = -0.041031725066768575f,
              K2 = 0.013710004822696948f, K3 = 0.006510206083837737f,
              K4 = -0.0014789063378272242f;
  const float m = mul;
  const float w_r = m * K0, w_d = m * K1, w_r2 = m * K2, w_l = m * K3, w_d2 = m * K4;
  float sum = 1.0f + m * 4.0f * (K0 + K1 + K2 + K4 + 2.0f * K3);
  if (sum < 1e-5f) sum = 1e-5f;
  const float norm = 1.0f / sum;
  return {norm, norm * w_r, norm * w_d, norm * w_r2, norm * w_l, norm * w_d2};
}

// Output pixel (x, y) of the padded plane is the filtered image at
// (min(x, W-1), min(y, H-1)): the padding stays an edge replication.
template <bool F>
__global__ __launch_bounds__(256) void k_gaborish(const float *__restrict__ src,
                                                  float *__restrict__ dst, int stride,
                                                  int padded_h, int W, int H, GabWeights k) {
  __shared__ float t[GAB_LH][GAB_LW + 1];
  const size_t plane = (size_t)stride * padded_h;
  const float *s = src + blockIdx.z * plane;
  float *o = dst + blockIdx.z * plane;
  const int x0 = blockIdx.x * GAB_TW, y0 = blockIdx.y * GAB_TH;
  // Source columns / rows used by this tile start at the clamped origin.
  const int bx = min(x0, W - 1), by = min(y0, H - 1);
  for (int e = threadIdx.x; e < GAB_LH * GAB_LW; e += 256) {
    const int ly = e / GAB_LW, lx = e - ly * GAB_LW;
    const int sy = gab_reflect(by - 2 + ly, H), sx = gab_reflect(bx - 2 + lx, W);
    t[ly][lx] = s[(size_t)sy * stride + sx];
  }
  __syncthreads();
  const int x = x0 + (threadIdx.x & 63);
  if (x >= stride) return;
  const int cx = min(x, W - 1) - bx + 2;
#pragma unroll
  for (int i = 0; i < GAB_TH / 4; i++) {
    const int y = y0 + (threadIdx.x >> 6) + 4 * i;
    if (y >= padded_h) break;
    const int cy = min(y, H - 1) - by + 2;
    const float *r0 = t[cy], *rm1 = t[cy - 1], *rp1 = t[cy + 1], *rm2 = t[cy - 2],
                *rp2 = t[cy + 2];
    const float s_r = rm1[cx] + rp1[cx] + r0[cx - 1] + r0[cx + 1];
    const float s_d = rm1[cx - 1] + rm1[cx + 1] + rp1[cx - 1] + rp1[cx + 1];
    const float s_r2 = rm2[cx] + rp2[cx] + r0[cx - 2] + r0[cx + 2];
    const float s_d2 = rm2[cx - 2] + rm2[cx + 2] + rp2[cx - 2] + rp2[cx + 2];
    const float s_l = rm2[cx - 1] + rm2[cx + 1] + rm1[cx - 2] + rm1[cx + 2] + rp1[cx - 2] +
                      rp1[cx + 2] + rp2[cx - 1] + rp2[cx + 1];
    float acc = k.c * r0[cx];
    acc = fmla<F>(k.r, s_r, acc);
    acc = fmla<F>(k.d, s_d, acc);
    acc = fmla<F>(k.r2, s_r2, acc);
    acc = fmla<F>(k.d2, s_d2, acc);
    acc = fmla<F>(k.l, s_l, acc);
    o[(size_t)y * stride + x] = acc;
  }
}

}  // namespace

extern "C" {

int jxg_gaborish_inverse_dev(jxg_ctx *ctx, const float *d_src, float *d_dst, uint32_t stride,
                             uint32_t padded_h, uint32_t width, uint32_t height, float mul,
                             int fused) {
  if (!ctx || !d_src || !d_dst) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  if (!(mul >= 0.0f) || !(mul <= FLT_MAX))
    return set_err(JXG_ERR_INVALID_ARG, "gaborish multiplier %g must be finite and >= 0",
                   (double)mul);
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t n = (size_t)3 * stride * padded_h;
  const bool in_place = d_src == d_dst;
  if (in_place && (rc = ensure(&ctx->d_tmp, &ctx->tmp_cap, n))) return rc;
  float *out = in_place ? ctx->d_tmp : d_dst;
  const GabWeights k = gab_weights(mul);
  const dim3 grid((stride + GAB_TW - 1) / GAB_TW, (padded_h + GAB_TH - 1) / GAB_TH, 3);
  if (fused)
    k_gaborish<true><<<grid, 256, 0, ctx->stream>>>(d_src, out, (int)stride, (int)padded_h,
                                                    (int)width, (int)height, k);
  else
    k_gaborish<false><<<grid, 256, 0, ctx->stream>>>(d_src, out, (int)stride, (int)padded_h,
                                                     (int)width, (int)height, k);
  LAUNCH_CHECK("k_gaborish");
  if (in_place)
    HIP_TRY(hipMemcpyAsync(d_dst, ctx->d_tmp, n * sizeof(float), hipMemcpyDeviceToDevice,
                           ctx->stream));
  return JXG_OK;
}

}  // extern "C"
