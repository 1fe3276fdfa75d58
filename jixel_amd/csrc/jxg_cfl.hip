// jxg_cfl.hip - chroma-from-luma slope maps (fill_cmap,
// src/color_correlation.rs:901-960, regression path compute_cmap_tile
// :441-508, the path XYB frames take at Fast and faster).
//
//   k_cmap_blocks  one 128-thread workgroup per quarter tile (64 x 16 px): the
//                  three planes staged in LDS, every 8x8 block's three DCT8s
//                  as 1-D lines (rows, then columns with the 1/64 scale), and
//                  the block's four regression sums, one lane each, in
//                  coefficient order.
//   k_cmap_solve   one wavefront per tile: the block sums in raster order and
//                  the two slopes.  Every sum keeps the reference's
//                  sequential f32 order.
#include <hip/hip_runtime.h>

#include "jxg_batch.h"
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

// A quarter tile (64 px wide, 16 px tall = 2 block rows of 8 blocks) per
// 128-thread workgroup: DCT8s of the three planes in LDS, then four lanes per
// block run the four regression sums (cfl_regression_scalar,
// src/color_correlation.rs:333-357) in coefficient order.  Block sums go to
// `part`; k_cmap_solve adds a tile's blocks in raster order and solves.
#define CM_ROWS 16
#define CM_T 128

template <bool F, int ROWS>
__device__ __forceinline__ void cmap_blocks_body(const float *__restrict__ xyb, int stride,
                                                      int padded_h, int W, int H,
                                                      const float *__restrict__ qm_x,
                                                      const float *__restrict__ qm_b, CmapArgs a,
                                                      float *__restrict__ part, int tiles_x_, int bix, int biy) {
  __shared__ float t[3][ROWS * CM_P];  // X, Y, B
  __shared__ float qms[2][64];         // the X and B quantization weights (LDS: read per step)
  const int tid = threadIdx.x;
  static_assert(CM_T >= 128, "one thread per weight");
  if (tid < 128) qms[tid >> 6][tid & 63] = tid < 64 ? qm_x[tid] : qm_b[tid - 64];
  constexpr int PARTS = 64 / ROWS;  // sub-tiles per 64x64 tile
  const int tx = bix, ty = biy / PARTS, quarter = biy % PARTS;
  const int x0 = tx * 64, y0 = ty * 64 + quarter * ROWS;
  const size_t plane = (size_t)stride * padded_h;
  // stride and padded_h are multiples of 64: the whole quarter tile is in bounds.
  for (int e = tid; e < 3 * ROWS * 16; e += CM_T) {
    const int c = e / (ROWS * 16), r = (e / 16) % ROWS, q = e % 16;
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
  constexpr int NB = ROWS / 8 * 8;  // blocks per workgroup
  for (int e = tid; e < 3 * NB * 8; e += CM_T) {
    const int c = e / (NB * 8), rem = e % (NB * 8), b = rem >> 3, r = rem & 7;
    dct_line<F, 8>(&t[c][((b >> 3) * 8 + r) * CM_P + (b & 7) * 8], 1, 0.0f, false);
  }
  __syncthreads();
  for (int e = tid; e < 3 * NB * 8; e += CM_T) {
    const int c = e / (NB * 8), rem = e % (NB * 8), b = rem >> 3, col = rem & 7;
    dct_line<F, 8>(&t[c][((b >> 3) * 8) * CM_P + (b & 7) * 8 + col], CM_P, 1.0f / 64.0f, true);
  }
  __syncthreads();
  if (tid < 4 * NB) {
    const int b = tid >> 2, which = tid & 3;
    const int bx = b & 7, by = b >> 3;
    const float K_INV = 1.0f / 84.0f;
    const float shift = which < 2 ? -a.base_x : 1.0f - a.base_b;
    const float *oc = which < 2 ? t[0] : t[2];
    float acc = 0.0f;
    // coefficient k = u * 8 + v of the square DCT8 sits at LDS (v, u); the
    // lanes of a block differ only in selects, so the wavefront stays uniform
#pragma unroll 16
    for (int k = 0; k < 64; k++) {
      const int at = (by * 8 + (k & 7)) * CM_P + bx * 8 + (k >> 3);
      float y = t[1][at], o = oc[at];
      if (k == 0) y = o = 0.0f;  // the regression ignores the DC
      const float fo = fmla<F>(shift, y, o);
      o = shift != 0.0f ? fo : o;
      const float qm = qms[which < 2 ? 0 : 1][k];
      const float m = y * qm, s_o = o * qm;
      const float am = K_INV * m;
      // which 0 / 2: a * a; 1: a * (-s_x); 3: a * (m_b - s_b) (fmla(1, m, -s) is exact)
      const float second = (which & 1) ? (which == 1 ? -s_o : fmla<F>(1.0f, m, -s_o)) : am;
      acc = fmla<F>(am, second, acc);
    }
    const int tiles_x = tiles_x_;
    part[((size_t)(ty * tiles_x + tx) * 64 + quarter * NB + b) * 4 + which] = acc;
  }
}

template <bool F>
__global__ __launch_bounds__(CM_T) void k_cmap_blocks(const float *__restrict__ xyb, int stride,
                                                      int padded_h, int W, int H,
                                                      const float *__restrict__ qm_x,
                                                      const float *__restrict__ qm_b, CmapArgs a,
                                                      float *__restrict__ part) {
  cmap_blocks_body<F, CM_ROWS>(xyb, stride, padded_h, W, H, qm_x, qm_b, a, part, (int)gridDim.x,
                               blockIdx.x, blockIdx.y);
}

// The tile's block sums in raster order, then the two slopes
// (compute_cmap_tile, src/color_correlation.rs:470-508).  One wavefront per
// tile: lane b loads block b's four sums, lane 0 adds them in order.
template <bool F>
__device__ __forceinline__ void cmap_solve_body(const float *__restrict__ part, int W, int H,
                                                   int tiles_x, int tiles_y, CmapArgs a,
                                                   int8_t *__restrict__ ytox,
                                                   int8_t *__restrict__ ytob, int bix, int biy) {
  __shared__ float4 s4[64];
  const int ti = bix, lane = threadIdx.x;
  const int tx = ti % tiles_x, ty = ti / tiles_x;
  s4[lane] = reinterpret_cast<const float4 *>(part)[(size_t)ti * 64 + lane];
  __syncthreads();
  if (lane != 0) return;
  const int xb = (W + 7) / 8, yb = (H + 7) / 8;
  const int bxc = min(8, xb - tx * 8), byc = min(8, yb - ty * 8);
  float ca_x = 0.0f, cb_x = 0.0f, ca_b = 0.0f, cb_b = 0.0f;
  int num = 0;
  for (int by = 0; by < byc; by++)
    for (int bx = 0; bx < bxc; bx++) {
      const int px = tx * 64 + bx * 8, py = ty * 64 + by * 8;
      if (px + 8 > W || py + 8 > H) continue;
      const float4 s = s4[by * 8 + bx];
      ca_x += s.x;
      cb_x += s.y;
      ca_b += s.z;
      cb_b += s.w;
      num += 64;
    }
  const float dz = cfl_deadzone(a.distance);
  const float units = a.color_factor / 84.0f;
  ytox[ti] = (int8_t)solve_multiplier<F>(ca_x, cb_x, num, dz, units);
  ytob[ti] = (int8_t)solve_multiplier<F>(ca_b, cb_b, num, dz, units);
}

template <bool F>
__global__ __launch_bounds__(64) void k_cmap_solve(const float *__restrict__ part, int W, int H,
                                                   int tiles_x, int tiles_y, CmapArgs a,
                                                   int8_t *__restrict__ ytox,
                                                   int8_t *__restrict__ ytob) {
  cmap_solve_body<F>(part, W, H, tiles_x, tiles_y, a, ytox, ytob, blockIdx.x, blockIdx.y);
}

// Batched forms: one launch per stage over every image (jxg_batch.h).
// ROWS (16 or 32) pixel rows per workgroup: a tile shape the batch
// autotuner chooses.
template <bool F, int ROWS>
__global__ __launch_bounds__(CM_T) void k_cmap_blocks_b(const BatchImg *__restrict__ img,
                                                        const uint32_t *__restrict__ pfx, int n) {
  const int i = batch_image(pfx, n, blockIdx.x);
  const BatchImg &I = img[i];
  const float *qm_x = I.tables + 3 * 64;  // DCT8 inv_matrix(0); the DCT8 kind is at offset 0
  const float *qm_b = I.tables + 5 * 64;  // inv_matrix(2)
  const uint32_t l = blockIdx.x - pfx[i];
  const int gw = I.gw[BS_CMAPB];
  const CmapArgs a{I.distance, I.cfl_base_x, I.cfl_base_b, I.cfl_color_factor};
  cmap_blocks_body<F, ROWS>(I.xyb, I.stride, I.ph, I.W, I.H, qm_x, qm_b, a, I.part, gw,
                            (int)(l % gw), (int)(l / gw));
}

template <bool F>
__global__ __launch_bounds__(64) void k_cmap_solve_b(const BatchImg *__restrict__ img,
                                                     const uint32_t *__restrict__ pfx, int n) {
  const int i = batch_image(pfx, n, blockIdx.x);
  const BatchImg &I = img[i];
  const CmapArgs a{I.distance, I.cfl_base_x, I.cfl_base_b, I.cfl_color_factor};
  cmap_solve_body<F>(I.part, I.W, I.H, I.tiles_x, I.tiles_y, a, I.ytox, I.ytob,
                     (int)(blockIdx.x - pfx[i]), 0);
}

}  // namespace

template <bool F>
static void launch_cmap(hipStream_t st, const BatchImg *d_img, const uint32_t *d_pfx, int n,
                        const uint32_t total[BS_N], int cm_rows) {
  const int m = n + 1;
  if (cm_rows == 32)
    k_cmap_blocks_b<F, 32><<<total[BS_CMAPB], CM_T, 0, st>>>(d_img, d_pfx + BS_CMAPB * m, n);
  else
    k_cmap_blocks_b<F, 16><<<total[BS_CMAPB], CM_T, 0, st>>>(d_img, d_pfx + BS_CMAPB * m, n);
  k_cmap_solve_b<F><<<total[BS_CMAPS], 64, 0, st>>>(d_img, d_pfx + BS_CMAPS * m, n);
}

JXG_HIDDEN int batch_cmap(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                          const uint32_t *d_pfx, int n, const uint32_t total[BS_N], int fused,
                          int cm_rows) {
  (void)ctx;
  static_assert(JXG_QT_DCT8 == 0, "the batch CfL stage reads the DCT8 kind at blob offset 0");
  if (fused)
    launch_cmap<true>(st, d_img, d_pfx, n, total, cm_rows);
  else
    launch_cmap<false>(st, d_img, d_pfx, n, total, cm_rows);
  LAUNCH_CHECK("batch cmap");
  return JXG_OK;
}

JXG_HIDDEN int fill_cmap(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                         uint32_t width, uint32_t height, const jxg_frame_params *p,
                         int8_t *d_ytox, int8_t *d_ytob, hipStream_t st) {
  const int tiles_x = (int)((width + 63) / 64), tiles_y = (int)((height + 63) / 64);
  const size_t need = (size_t)tiles_x * tiles_y * 64 * 4;
  int rc = ensure(&ctx->d_tmp, &ctx->tmp_cap, need);
  if (rc) return rc;
  const float *tables, *selq;
  if ((rc = tables_for(ctx, p, &tables, &selq))) return rc;
  const uint32_t off = jxg_qt_offset(JXG_QT_DCT8);
  const float *qm_x = tables + off + 3 * 64;  // inv_matrix(0) of DCT8
  const float *qm_b = tables + off + 5 * 64;  // inv_matrix(2)
  const CmapArgs a{p->distance, p->cfl_base_x, p->cfl_base_b, p->cfl_color_factor};
  const dim3 grid(tiles_x, 4 * tiles_y);
  const unsigned sgrid = (unsigned)(tiles_x * tiles_y);
  if (p->fused) {
    k_cmap_blocks<true><<<grid, CM_T, 0, st>>>(d_xyb, (int)stride, (int)padded_h,
                                                       (int)width, (int)height, qm_x, qm_b, a,
                                                       ctx->d_tmp);
    k_cmap_solve<true><<<sgrid, 64, 0, st>>>(ctx->d_tmp, (int)width, (int)height,
                                                      tiles_x, tiles_y, a, d_ytox, d_ytob);
  } else {
    k_cmap_blocks<false><<<grid, CM_T, 0, st>>>(d_xyb, (int)stride, (int)padded_h,
                                                        (int)width, (int)height, qm_x, qm_b, a,
                                                        ctx->d_tmp);
    k_cmap_solve<false><<<sgrid, 64, 0, st>>>(ctx->d_tmp, (int)width, (int)height,
                                                       tiles_x, tiles_y, a, d_ytox, d_ytob);
  }
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
  return fill_cmap(ctx, d_xyb, stride, padded_h, width, height, params, d_ytox, d_ytob,
                   ctx->stream);
}

}  // extern "C"
