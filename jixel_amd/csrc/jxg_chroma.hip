// jxg_chroma.hip - chroma-HF protection of the adaptive quant field
// (apply_chroma_hf_protection_with_stats, src/adaptive_quant.rs:344-408).
//
//   k_chroma_hf  one thread per 8x8 block: the three statistics of
//                chroma_hf_stats_scalar (src/adaptive_quant.rs:222-279) as
//                sequential f32 sums in its order (magnitudes row by row,
//                then horizontal pairs, then vertical pairs), the risk weights
//                and the raised quant value.  Whole-image quant field; the
//                reference's per-DC-group calls cover the same blocks.
#include <hip/hip_runtime.h>

#include "jxg_ctx.h"
#include "jxg_device.h"

using namespace jxg;

namespace {

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

template <bool F>
__global__ __launch_bounds__(256) void k_chroma_hf(const float *__restrict__ xyb, int stride,
                                                   int padded_h, int W, int H, int xb, int yb,
                                                   float ds, uint8_t *__restrict__ qf) {
  const int bx = blockIdx.x * 256 + threadIdx.x, by = blockIdx.y;
  if (bx >= xb) return;
  const int px = bx * 8, py = by * 8;
  const int w = min(W - px, 8), h = min(H - py, 8);
  if (w * h == 1) return;  // no edges
  const size_t plane = (size_t)stride * padded_h;
  const float *X = xyb + (size_t)py * stride + px, *Y = X + plane, *B = X + 2 * plane;
  float chroma_sum = 0.0f, chroma_grad = 0.0f, luma_grad = 0.0f;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const int i = y * stride + x;
      chroma_sum += fabsf(X[i]) + fabsf(B[i] - Y[i]);
    }
  int edges = 0;
  if (w > 1) {
    for (int y = 0; y < h; y++)
      for (int x = 0; x + 1 < w; x++) {
        const int i = y * stride + x, j = i + 1;
        chroma_grad += fabsf(X[j] - X[i]) + fabsf((B[j] - Y[j]) - (B[i] - Y[i]));
        luma_grad += fabsf(Y[j] - Y[i]);
      }
    edges += (w - 1) * h;
  }
  if (h > 1) {
    for (int y = 0; y + 1 < h; y++)
      for (int x = 0; x < w; x++) {
        const int i = y * stride + x, j = i + stride;
        chroma_grad += fabsf(X[j] - X[i]) + fabsf((B[j] - Y[j]) - (B[i] - Y[i]));
        luma_grad += fabsf(Y[j] - Y[i]);
      }
    edges += w * (h - 1);
  }
  const float mean_chroma = chroma_sum / (float)(w * h);
  const float mean_gradient = chroma_grad / (float)edges;
  const float chroma_share = chroma_grad / (chroma_grad + luma_grad + 1e-4f);
  const float chroma_w = clamp01((mean_chroma - 0.025f) / 0.075f);
  const float gradient_w = clamp01((mean_gradient - 0.006f) / 0.030f);
  const float share_w = clamp01((chroma_share - 0.25f) / 0.55f);
  const float risk = chroma_w * gradient_w * share_w;
  if (!(risk > 0.0f)) return;
  const float ceiling = fmla<F>(1.4f, share_w * share_w, 0.35f);  // ceil_share, ceil_base
  const float gain = fmla<F>(ceiling * ds, risk, 1.0f);
  uint8_t *q = qf + (size_t)by * xb + bx;
  *q = (uint8_t)fminf(fmaxf(roundf((float)*q * gain), 1.0f), 255.0f);
}

}  // namespace

// The distance strength: fades in over 0.35..0.75 and out over 1..2
// (chroma_hf_fit, src/adaptive_quant.rs:331-342,354-366); 0 skips the stage.
JXG_HIDDEN float chroma_hf_strength(float d) {
  if (!(d >= 0.35f)) return 0.0f;
  const float fade_in = fminf(fmaxf((d - 0.35f) / 0.40f, 0.0f), 1.0f);
  const float fade_out = 1.0f - fminf(fmaxf((d - 1.0f) / 1.0f, 0.0f), 1.0f);
  return 0.40f * fade_in * fade_out;
}

JXG_HIDDEN int chroma_hf(jxg_ctx *ctx, const float *xyb, uint32_t stride, uint32_t ph, uint32_t W,
                         uint32_t H, float distance, int fused, uint8_t *qf) {
  const float ds = chroma_hf_strength(distance);
  if (!(ds > 0.0f)) return JXG_OK;
  const int xb = (int)((W + 7) / 8), yb = (int)((H + 7) / 8);
  const dim3 grid((xb + 255) / 256, yb);
  if (fused)
    k_chroma_hf<true><<<grid, 256, 0, ctx->stream>>>(xyb, (int)stride, (int)ph, (int)W, (int)H, xb,
                                                     yb, ds, qf);
  else
    k_chroma_hf<false><<<grid, 256, 0, ctx->stream>>>(xyb, (int)stride, (int)ph, (int)W, (int)H,
                                                      xb, yb, ds, qf);
  LAUNCH_CHECK("k_chroma_hf");
  return JXG_OK;
}

extern "C" int jxg_chroma_hf_protection_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride,
                                            uint32_t padded_h, uint32_t width, uint32_t height,
                                            float distance, int fused, uint8_t *d_qf) {
  if (!ctx || !d_xyb || !d_qf) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  if (!(distance > 0.0f)) return set_err(JXG_ERR_INVALID_DISTANCE, "distance %g", (double)distance);
  HIP_TRY(hipSetDevice(ctx->device));
  return chroma_hf(ctx, d_xyb, stride, padded_h, width, height, distance, fused, d_qf);
}
