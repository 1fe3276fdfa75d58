// jxg_cfl.hip - chroma-from-luma slope maps (fill_cmap,
// src/color_correlation.rs:901-960, regression path compute_cmap_tile
// :441-508, the path XYB frames take at Fast and faster).
//
//   k_cmap  one 256-thread workgroup per 64x64-px tile: the three planes of
//           the tile are staged in LDS, every 8x8 block gets its three DCT8s
//           as 1-D lines (rows, then columns with the 1/64 scale), one thread
//           per block runs the regression sums in coefficient order, and one
//           thread accumulates the blocks in raster order and solves the two
//           slopes, so every sum keeps the reference's sequential f32 order.
#include <hip/hip_runtime.h>

#include "jxg_ctx.h"
#include "jxg_device.h"

using namespace jxg;

namespace {

constexpr int CM_P = 65;  // LDS row pitch

struct CmapArgs {
  float distance;
  float base_x, base_b;
  float color_factor;
};

// cfl_deadzone (src/color_correlation.rs:383-394)
__device__ __forceinline__ float cfl_deadzone(float d) {
  const float AMOUNT = 1.5f, LO = 1.0f, HI = 2.0f;
  if (d <= LO) return AMOUNT;
  if (d >= HI) return 0.0f;
  const float SCALE = AMOUNT / (HI - LO);
  return SCALE * (HI - d);
}

// solve_multiplier (src/color_correlation.rs:397-409)
template <bool F>
__device__ __forceinline__ int solve_multiplier(float ca, float cb, int num, float dz,
                                                float units) {
  if (num == 0) return 0;
  float x = -cb / fmla<F>((float)num * 1e-3f, 0.5f, ca);
  if (fabsf(x) < dz) x = 0.0f;
  const float r = fminf(fmaxf(round_away(x * units), -128.0f), 127.0f);
  return (int)r;
}

template <bool F>
__global__ __launch_bounds__(256) void k_cmap(const float *__restrict__ xyb, int stride,
                                              int padded_h, int W, int H, int tiles_x,
                                              const float *__restrict__ qm_x,
                                              const float *__restrict__ qm_b, CmapArgs a,
                                              int8_t *__restrict__ ytox,
                                              int8_t *__restrict__ ytob) {
  __shared__ float t[3][64 * CM_P];  // X, Y, B
  __shared__ float sums[64][4];
  __shared__ int valid[64];
  const int tid = threadIdx.x;
  const int tx = blockIdx.x, ty = blockIdx.y;
  const int x0 = tx * 64, y0 = ty * 64;
  const size_t plane = (size_t)stride * padded_h;
  // stride and padded_h are multiples of 64: the whole tile is in bounds.
  for (int e = tid; e < 3 * 64 * 16; e += 256) {
    const int c = e / (64 * 16), r = (e / 16) % 64, q = e % 16;
    const float4 v =
        *reinterpret_cast<const float4 *>(xyb + c * plane + (size_t)(y0 + r) * stride + x0 + 4 * q);
    float *d = &t[c][r * CM_P + 4 * q];
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
    d[3] = v.w;
  }
  __syncthreads();
  // DCT8 rows (src/dct.rs:631-660: rows first, columns scaled by 1/64)
  for (int e = tid; e < 3 * 64 * 8; e += 256) {
    const int c = e / 512, rem = e % 512, b = rem >> 3, r = rem & 7;
    dct_line<F, 8>(&t[c][((b >> 3) * 8 + r) * CM_P + (b & 7) * 8], 1, 0.0f, false);
  }
  __syncthreads();
  for (int e = tid; e < 3 * 64 * 8; e += 256) {
    const int c = e / 512, rem = e % 512, b = rem >> 3, col = rem & 7;
    dct_line<F, 8>(&t[c][((b >> 3) * 8) * CM_P + (b & 7) * 8 + col], CM_P, 1.0f / 64.0f, true);
  }
  __syncthreads();
  if (tid < 64) {
    const int bx = tid & 7, by = tid >> 3;
    const int px = x0 + bx * 8, py = y0 + by * 8;
    const bool ok = px + 8 <= W && py + 8 <= H;
    valid[tid] = ok;
    if (ok) {
      const float K_INV = 1.0f / 84.0f;
      const float sx = -a.base_x, sb = 1.0f - a.base_b;
      float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
      // coefficient k = u * 8 + v of the square DCT8 sits at LDS (v, u)
      for (int k = 0; k < 64; k++) {
        const int at = (by * 8 + (k & 7)) * CM_P + bx * 8 + (k >> 3);
        float y = t[1][at], x = t[0][at], b = t[2][at];
        if (k == 0) y = x = b = 0.0f;  // the regression ignores the DC
        if (sx != 0.0f) x = fmla<F>(sx, y, x);
        if (sb != 0.0f) b = fmla<F>(sb, y, b);
        // cfl_regression_scalar (src/color_correlation.rs:333-357)
        const float m_x = y * qm_x[k], s_x = x * qm_x[k];
        const float a_x = K_INV * m_x, b_x = -s_x;
        s0 = fmla<F>(a_x, a_x, s0);
        s1 = fmla<F>(a_x, b_x, s1);
        const float m_b = y * qm_b[k], s_b = b * qm_b[k];
        const float a_b = K_INV * m_b, b_b = fmla<F>(1.0f, m_b, -s_b);
        s2 = fmla<F>(a_b, a_b, s2);
        s3 = fmla<F>(a_b, b_b, s3);
      }
      sums[tid][0] = s0;
      sums[tid][1] = s1;
      sums[tid][2] = s2;
      sums[tid][3] = s3;
    }
  }
  __syncthreads();
  if (tid == 0) {
    const int xb = (W + 7) / 8, yb = (H + 7) / 8;
    const int bxc = min(8, xb - tx * 8), byc = min(8, yb - ty * 8);
    float ca_x = 0.0f, cb_x = 0.0f, ca_b = 0.0f, cb_b = 0.0f;
    int num = 0;
    for (int by = 0; by < byc; by++)
      for (int bx = 0; bx < bxc; bx++) {
        const int b = by * 8 + bx;
        if (!valid[b]) continue;
        ca_x += sums[b][0];
        cb_x += sums[b][1];
        ca_b += sums[b][2];
        cb_b += sums[b][3];
        num += 64;
      }
    const float dz = cfl_deadzone(a.distance);
    const float units = a.color_factor / 84.0f;
    ytox[ty * tiles_x + tx] = (int8_t)solve_multiplier<F>(ca_x, cb_x, num, dz, units);
    ytob[ty * tiles_x + tx] = (int8_t)solve_multiplier<F>(ca_b, cb_b, num, dz, units);
  }
}

}  // namespace

JXG_HIDDEN int fill_cmap(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                         uint32_t width, uint32_t height, const jxg_frame_params *p,
                         int8_t *d_ytox, int8_t *d_ytob) {
  const int tiles_x = (int)((width + 63) / 64), tiles_y = (int)((height + 63) / 64);
  const uint32_t off = jxg_qt_offset(JXG_QT_DCT8);
  const float *qm_x = ctx->d_tables + off + 3 * 64;  // inv_matrix(0) of DCT8
  const float *qm_b = ctx->d_tables + off + 5 * 64;  // inv_matrix(2)
  const CmapArgs a{p->distance, p->cfl_base_x, p->cfl_base_b, p->cfl_color_factor};
  const dim3 grid(tiles_x, tiles_y);
  if (p->fused)
    k_cmap<true><<<grid, 256, 0, ctx->stream>>>(d_xyb, (int)stride, (int)padded_h, (int)width,
                                                (int)height, tiles_x, qm_x, qm_b, a, d_ytox,
                                                d_ytob);
  else
    k_cmap<false><<<grid, 256, 0, ctx->stream>>>(d_xyb, (int)stride, (int)padded_h, (int)width,
                                                 (int)height, tiles_x, qm_x, qm_b, a, d_ytox,
                                                 d_ytob);
  LAUNCH_CHECK("k_cmap");
  return JXG_OK;
}

extern "C" {

int jxg_fill_cmap_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                      uint32_t width, uint32_t height, const jxg_frame_params *params,
                      int8_t *d_ytox, int8_t *d_ytob) {
  if (!ctx || !d_xyb || !params || !d_ytox || !d_ytob)
    return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  if ((rc = upload_derived(ctx, params))) return rc;  // validates the parameters
  HIP_TRY(hipSetDevice(ctx->device));
  return fill_cmap(ctx, d_xyb, stride, padded_h, width, height, params, d_ytox, d_ytob);
}

}  // extern "C"
