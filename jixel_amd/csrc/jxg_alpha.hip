// jxg_alpha.hip - the alpha-dependent stages of 2- and 4-channel input.
//
//   k_alpha_extract   the alpha samples of the interleaved input as a W x H plane
//   k_fringe<INV>     one sweep of alpha_fringe::clear (src/alpha_fringe.rs:105-225):
//                     every fringe sample (level, or 255 - level, in 1..4) whose
//                     clipped 9x9 window of fringe levels has mean <= 1/2 is
//                     replaced.  The reference replaces a row only after it has
//                     left the running window, so every window reads original
//                     samples and the sweep is a pure stencil: 64x16 output tiles
//                     with a 4-px halo in LDS, column sums then row sums, all in
//                     integers.  The first sweep also counts near-opaque samples
//                     of its output, the second its replacements; both counts
//                     gate the second sweep and its commit on the device
//                     (clear, src/alpha_fringe.rs:77-103).
//   k_alpha_vis       visibility bits (alpha > 0), 32 per word, and the plane's
//                     invisible count, minimum and maximum.
//   k_fill_invisible  fill_invisible (src/invisible.rs:34-146).  Each invisible
//                     sample reads its already filled left and upper-right
//                     neighbours, so sample (x, y) can run once (x - 1, y) and
//                     (x + 1, y - 1) are done: the anti-diagonals x + 2y are the
//                     parallel steps.  One workgroup per channel, 1024 rows per
//                     pass, lane = row: lane r handles column s - 2r at step s,
//                     its upper-right value is lane r - 1's previous output (one
//                     DPP wave shift), and the first lane of each wave reads the
//                     wave above through a 64-entry LDS ring, the wave running
//                     144 steps behind so a barrier every 16 steps orders every
//                     ring write before its read and every read before the slot
//                     is reused.  Each sample is computed once with the
//                     reference's operations in its order, so the result is
//                     bit-identical.  Critical path: W + 2 * 1024 + 15 * 16 steps
//                     per pass.
//   k_repad           refreshes the edge padding of the planes from the filled
//                     edge samples.
//   k_alpha_blocks    alpha_block_factors (src/adaptive_quant.rs:953-1030).
//   k_alpha_apply     apply_alpha_factors (src/adaptive_quant.rs:1036-1070), for
//                     non-constant alpha only (src/frame/encoder.rs:60-69).
#include <hip/hip_runtime.h>

#include "jxg_ctx.h"
#include "jxg_device.h"

using namespace jxg;

namespace {

// stat words: [0] near-opaque count after the first sweep, [1] replacements of
// the second sweep, [2] minimum alpha, [3] maximum alpha, [4] invisible count
enum { ST_NEAR = 0, ST_CHANGED = 1, ST_MIN = 2, ST_MAX = 3, ST_ZERO = 4, ST_WORDS = 8 };

constexpr uint32_t FR_REACH = 4;

__global__ void k_astat_init(uint32_t *st) {
  if (threadIdx.x < ST_WORDS) st[threadIdx.x] = threadIdx.x == ST_MIN ? 255u : 0u;
}

__global__ __launch_bounds__(256) void k_alpha_extract(const uint8_t *__restrict__ in,
                                                       size_t row_bytes, int ch, int W, int H,
                                                       uint8_t *__restrict__ a) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= W) return;
  a[(size_t)y * W + x] = in[(size_t)y * row_bytes + (size_t)x * ch + ch - 1];
}

__device__ __forceinline__ uint32_t fringe_level(uint32_t a, bool inv) {
  const uint32_t level = inv ? 255u - a : a;
  return level <= FR_REACH ? level : 0u;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

#define FT_W 64
#define FT_H 16
#define FT_PW (FT_W + 8)
#define FT_PH (FT_H + 8)

// INV = false: the first sweep (fringe = levels 1..4, replaced by 0);
// INV = true: the second (fringe = 255 - level in 1..4, replaced by 255),
// run only when the first sweep left near_opaque * 12 >= W * H samples.
template <bool INV>
__global__ __launch_bounds__(256) void k_fringe(const uint8_t *__restrict__ src,
                                                uint8_t *__restrict__ dst, int W, int H,
                                                uint32_t *__restrict__ st) {
  const size_t n = (size_t)W * H;
  if (INV && (size_t)st[ST_NEAR] * 12 < n) return;
  __shared__ uint8_t lv[FT_PH][FT_PW];
  __shared__ uint16_t cs[FT_H][FT_PW];
  const int tid = threadIdx.x;
  const int x0 = blockIdx.x * FT_W, y0 = blockIdx.y * FT_H;
  for (int e = tid; e < FT_PH * FT_PW; e += 256) {
    const int r = e / FT_PW, c = e % FT_PW;
    const int y = y0 - 4 + r, x = x0 - 4 + c;
    lv[r][c] = (x >= 0 && x < W && y >= 0 && y < H) ? (uint8_t)fringe_level(src[(size_t)y * W + x], INV) : 0;
  }
  __syncthreads();
  for (int e = tid; e < FT_H * FT_PW; e += 256) {
    const int r = e / FT_PW, c = e % FT_PW;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) s += lv[r + k][c];
    cs[r][c] = (uint16_t)s;
  }
  __syncthreads();
  uint32_t counted = 0;
  for (int e = tid; e < FT_H * FT_W; e += 256) {
    const int r = e / FT_W, c = e % FT_W;
    const int y = y0 + r, x = x0 + c;
    if (x >= W || y >= H) continue;
    const uint32_t a = src[(size_t)y * W + x];
    uint32_t out = a;
    if (lv[r + 4][c + 4] != 0) {
      uint32_t s = 0;
#pragma unroll
      for (int k = 0; k < 9; k++) s += cs[r][c + k];
      const uint32_t rows = (uint32_t)(min(H, y + 5) - max(0, y - 4));
      const uint32_t cols = (uint32_t)(min(W, x + 5) - max(0, x - 4));
      // sum * 2 * 255 <= rows * cols * 255 (MEAN_DIVISOR 2, max 255)
      if ((uint64_t)s * 2u <= (uint64_t)rows * cols) out = INV ? 255u : 0u;
    }
    dst[(size_t)y * W + x] = (uint8_t)out;
    if (INV) {
      counted += out != a;
    } else {
      const uint32_t d = 255u - out;
      counted += d >= 1 && d <= FR_REACH;
    }
  }
  counted = wave_sum(counted);
  if ((tid & 63) == 0 && counted) atomicAdd(&st[INV ? ST_CHANGED : ST_NEAR], counted);
}

// The second sweep's result replaces the first's when it changed at least
// 1/12 of the samples.
__global__ __launch_bounds__(256) void k_fringe_commit(const uint8_t *__restrict__ filled,
                                                       uint8_t *__restrict__ alpha, size_t n,
                                                       const uint32_t *__restrict__ st) {
  if ((size_t)st[ST_NEAR] * 12 < n || (size_t)st[ST_CHANGED] * 12 < n) return;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    alpha[i] = filled[i];
}

// Visibility words: bit x & 31 of vis[y * vw + (x >> 5)] = alpha(x, y) > 0,
// zero past the row (vw = ceil(W / 32) + 1).  Also the invisible count and
// the alpha range.
__global__ __launch_bounds__(256) void k_alpha_vis(const uint8_t *__restrict__ a, int W, int H,
                                                   int vw, uint32_t *__restrict__ vis,
                                                   uint32_t *__restrict__ st) {
  const int wx = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  uint32_t bits = 0, zeros = 0, lo = 255, hi = 0;
  if (wx < vw) {
    const uint8_t *row = a + (size_t)y * W;
    for (int i = 0; i < 32; i++) {
      const int x = wx * 32 + i;
      if (x >= W) break;
      const uint32_t v = row[x];
      bits |= (v > 0 ? 1u : 0u) << i;
      zeros += v == 0;
      lo = min(lo, v);
      hi = max(hi, v);
    }
    vis[(size_t)y * vw + wx] = bits;
  }
  zeros = wave_sum(zeros);
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, o));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, o));
  }
  if ((threadIdx.x & 63) == 0) {
    if (zeros) atomicAdd(&st[ST_ZERO], zeros);
    atomicMin(&st[ST_MIN], lo);
    atomicMax(&st[ST_MAX], hi);
  }
}

constexpr int FI_WAVES = 16;
constexpr int FI_ROWS = FI_WAVES * 64;
constexpr int FI_G = 8;    // steps per prefetch group
constexpr int FI_K = 16;   // steps between barriers
constexpr int FI_D = 144;  // lag of each wave behind the one above (>= 127 + FI_K, multiple of FI_G)
constexpr int FI_R = 64;   // ring entries per wave boundary (>= 128 + FI_K - FI_D + FI_K + 1)

// lane r receives lane r - 1's value; lane 0 keeps `old`
__device__ __forceinline__ float wave_shr1(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}

// Nine visibility bits (columns x0 .. x0 + 8) of one row's words.
__device__ __forceinline__ uint32_t vis_window(const uint32_t *__restrict__ row, int vw, int x0,
                                               bool valid) {
  if (!valid || x0 < -FI_G || x0 >= 32 * (vw - 1)) return 0;
  if (x0 < 0) return (row[0] << (-x0)) & 0x1FFu;
  const int wi = x0 >> 5;
  const uint64_t w = (uint64_t)row[wi] | ((uint64_t)row[wi + 1] << 32);
  return (uint32_t)(w >> (x0 & 31)) & 0x1FFu;
}

struct FillCols {
  uint32_t vb[3];  // visibility windows of rows y - 1, y, y + 1
  float v0[FI_G + 1], v1[FI_G + 1];  // samples of rows y and y + 1
};

__device__ __forceinline__ void fill_load(FillCols &c, const float *__restrict__ plane, int stride,
                                          const uint32_t *__restrict__ vis, int vw, int W, int H,
                                          int y, int x0) {
  const int yc = min(y, H - 1), yd = min(y + 1, H - 1);
  c.vb[0] = vis_window(vis + (size_t)max(y - 1, 0) * vw, vw, x0, y >= 1 && y - 1 < H);
  c.vb[1] = vis_window(vis + (size_t)yc * vw, vw, x0, y < H);
  c.vb[2] = vis_window(vis + (size_t)yd * vw, vw, x0, y + 1 < H);
  const float *r0 = plane + (size_t)yc * stride, *r1 = plane + (size_t)yd * stride;
#pragma unroll
  for (int j = 0; j <= FI_G; j++) {
    const int x = min(max(x0 + j, 0), W - 1);
    c.v0[j] = r0[x];
    c.v1[j] = r1[x];
  }
}

__global__ __launch_bounds__(FI_ROWS) void k_fill_invisible(float *__restrict__ xyb, int stride,
                                                            int padded_h, int W, int H,
                                                            const uint32_t *__restrict__ vis,
                                                            int vw,
                                                            const uint32_t *__restrict__ st) {
  if (st[ST_ZERO] == 0) return;  // every sample visible: nothing to fill
  __shared__ float ring[FI_WAVES][FI_R];
  float *plane = xyb + (size_t)blockIdx.x * stride * padded_h;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int steps = (FI_WAVES - 1) * FI_D + W + 126;
  const int groups = (steps + FI_K - 1) / FI_K * (FI_K / FI_G);
  for (int y0 = 0; y0 < H; y0 += FI_ROWS) {
    const int y = y0 + w * 64 + lane;
    const bool up = y > 0, down = y + 1 < H, row_ok = y < H;
    float *row = plane + (size_t)min(y, H - 1) * stride;
    float prev_out = 0.0f, left_val = 0.0f, up_val = 0.0f;
    bool left_vis = false;
    // Every lane meets column -1 before column 0 and takes its upper sample
    // from that step's exchange, except the pass's first row, which starts at
    // column 0: it reads (0, y0 - 1) from the previous pass.
    if (w == 0 && lane == 0 && y0 > 0)
      up_val = __hip_atomic_load(plane + (size_t)(y0 - 1) * stride, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
    FillCols cur, nxt;
    // group g of this wave covers local steps 8g .. 8g + 7, global steps
    // 8g + w * FI_D ..; global group gg runs local group gg - w * FI_D / 8
    const int lag_groups = w * (FI_D / FI_G);
    fill_load(nxt, plane, stride, vis, vw, W, H, y, -lag_groups * FI_G - 2 * lane);
    for (int gg = 0; gg < groups; gg++) {
      cur = nxt;
      const int x0 = (gg - lag_groups) * FI_G - 2 * lane;
      fill_load(nxt, plane, stride, vis, vw, W, H, y, x0 + FI_G);
#pragma unroll
      for (int j = 0; j < FI_G; j++) {
        const int x = x0 + j;
        float ur = wave_shr1(prev_out);  // (x + 1, y - 1), filled
        if (lane == 0) {
          if (w > 0)
            ur = ring[w - 1][(x + 1) & (FI_R - 1)];
          else if (y0 > 0)
            ur = __hip_atomic_load(plane + (size_t)(y0 - 1) * stride + min(max(x + 1, 0), W - 1),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const bool own_vis = (cur.vb[1] >> j) & 1;
        float val = cur.v0[j];
        if (row_ok && x >= 0 && x < W && !own_vis) {
          float sum = 0.0f, weight = 0.0f;
          if (x > 0) {
            sum += left_val;
            weight += 1.0f;
            if (left_vis) {
              sum += left_val;
              weight += 1.0f;
            }
          }
          if (x + 1 < W) {
            if (up) {
              sum += ur;
              weight += 1.0f;
            }
            if ((cur.vb[1] >> (j + 1)) & 1) {
              sum += 2.0f * cur.v0[j + 1];
              weight += 2.0f;
            }
            if (up && ((cur.vb[0] >> (j + 1)) & 1)) {
              sum += 2.0f * ur;
              weight += 2.0f;
            }
            if (down && ((cur.vb[2] >> (j + 1)) & 1)) {
              sum += 2.0f * cur.v1[j + 1];
              weight += 2.0f;
            }
          }
          if (up && ((cur.vb[0] >> j) & 1)) {
            sum += 2.0f * up_val;
            weight += 2.0f;
          }
          if (down && ((cur.vb[2] >> j) & 1)) {
            sum += 2.0f * cur.v1[j];
            weight += 2.0f;
          }
          val = weight > 1.0f ? sum / weight : sum;
          row[x] = val;
        }
        left_val = val;
        left_vis = own_vis;
        prev_out = val;
        up_val = ur;
        if (lane == 63 && w + 1 < FI_WAVES && x >= 0 && x < W) ring[w][x & (FI_R - 1)] = val;
      }
      if (gg & 1) __syncthreads();  // every FI_K = 2 groups
    }
    __threadfence();
    __syncthreads();
  }
}

// Edge padding of the three planes from the (filled) edge samples.
__global__ __launch_bounds__(256) void k_repad(float *__restrict__ xyb, int stride, int padded_h,
                                               int W, int H, const uint32_t *__restrict__ st) {
  if (st[ST_ZERO] == 0) return;
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= stride || (x < W && y < H)) return;
  float *plane = xyb + (size_t)blockIdx.z * stride * padded_h;
  plane[(size_t)y * stride + x] = plane[(size_t)min(y, H - 1) * stride + min(x, W - 1)];
}

// alpha_block_factors: the block's largest alpha / 255, 1 for fully
// transparent or opaque blocks, at least 0.1 otherwise.
__global__ __launch_bounds__(256) void k_alpha_blocks(const uint8_t *__restrict__ a, int W, int H,
                                                      int xb, int yb, float *__restrict__ f) {
  const int bx = blockIdx.x * 256 + threadIdx.x, by = blockIdx.y;
  if (bx >= xb) return;
  const int x = bx * 8, end = min(x + 8, W);
  uint32_t peak = 0;
  for (int yy = by * 8; yy < min(by * 8 + 8, H); yy++)
    for (int xx = x; xx < end; xx++) peak = max(peak, (uint32_t)a[(size_t)yy * W + xx]);
  const float al = (float)peak / 255.0f;
  f[(size_t)by * xb + bx] = (al == 0.0f || al >= 1.0f) ? 1.0f : fmaxf(al, 0.1f);
}

// apply_alpha_factors on first blocks; a multi-block transform takes the
// largest factor of its footprint and writes that one value to every cell.
// With `st`, only for non-constant alpha.
__global__ __launch_bounds__(256) void k_alpha_apply(const uint8_t *__restrict__ cells,
                                                     const float *__restrict__ f, int xb, int yb,
                                                     uint8_t *__restrict__ qf,
                                                     const uint32_t *__restrict__ st) {
  if (st && st[ST_MIN] == st[ST_MAX]) return;
  const int bx = blockIdx.x * 256 + threadIdx.x, by = blockIdx.y;
  if (bx >= xb) return;
  const uint8_t cell = cells[(size_t)by * xb + bx];
  const int s = cell >> 1;
  if (!(cell & 1) || s >= JXG_NUM_STRATEGIES) return;
  const int cx = cov_x(s), cy = cov_y(s);
  float k;
  if (bx + cx > xb || by + cy > yb) {
    k = 1.0f;
  } else {
    k = 0.0f;
    for (int iy = 0; iy < cy; iy++)
      for (int ix = 0; ix < cx; ix++) k = fmaxf(k, f[(size_t)(by + iy) * xb + bx + ix]);
  }
  const int q = min(max((int)roundf((float)qf[(size_t)by * xb + bx] * k), 1), 255);
  // a legal layout keeps every footprint inside the field; clip one that is not
  for (int iy = 0; iy < min(cy, yb - by); iy++)
    for (int ix = 0; ix < min(cx, xb - bx); ix++) qf[(size_t)(by + iy) * xb + bx + ix] = (uint8_t)q;
}

int ensure_stat(jxg_ctx *ctx) {
  if (ctx->d_astat) return JXG_OK;
  HIP_TRY(hipMalloc((void **)&ctx->d_astat, ST_WORDS * sizeof(uint32_t)));
  return JXG_OK;
}

// Fringe clearing of `a` (W x H) in place, with ctx->d_alpha2 / d_alpha3 as
// the sweeps' outputs.  Counters must have been reset.
int fringe_clear(jxg_ctx *ctx, uint8_t *a, uint32_t W, uint32_t H) {
  const size_t n = (size_t)W * H;
  int rc;
  if ((rc = ensure(&ctx->d_alpha2, &ctx->alpha2_cap, 2 * n))) return rc;
  uint8_t *s1 = ctx->d_alpha2, *s2 = ctx->d_alpha2 + n;
  const dim3 grid((W + FT_W - 1) / FT_W, (H + FT_H - 1) / FT_H);
  k_fringe<false><<<grid, 256, 0, ctx->stream>>>(a, s1, (int)W, (int)H, ctx->d_astat);
  k_fringe<true><<<grid, 256, 0, ctx->stream>>>(s1, s2, (int)W, (int)H, ctx->d_astat);
  // first sweep's output back into `a`, then the second's when it is kept
  HIP_TRY(hipMemcpyAsync(a, s1, n, hipMemcpyDeviceToDevice, ctx->stream));
  const unsigned cg = (unsigned)min((n + 255) / 256, (size_t)4096);
  k_fringe_commit<<<cg, 256, 0, ctx->stream>>>(s2, a, n, ctx->d_astat);
  LAUNCH_CHECK("k_fringe");
  return JXG_OK;
}

int visibility(jxg_ctx *ctx, const uint8_t *a, uint32_t W, uint32_t H, int *vw_out) {
  const int vw = (int)((W + 31) / 32) + 1;
  int rc;
  if ((rc = ensure(&ctx->d_vis, &ctx->vis_cap, (size_t)vw * H))) return rc;
  k_alpha_vis<<<dim3((vw + 255) / 256, H), 256, 0, ctx->stream>>>(a, (int)W, (int)H, vw, ctx->d_vis,
                                                                  ctx->d_astat);
  LAUNCH_CHECK("k_alpha_vis");
  *vw_out = vw;
  return JXG_OK;
}

int fill_planes(jxg_ctx *ctx, float *xyb, uint32_t stride, uint32_t ph, uint32_t W, uint32_t H,
                int vw) {
  k_fill_invisible<<<3, FI_ROWS, 0, ctx->stream>>>(xyb, (int)stride, (int)ph, (int)W, (int)H,
                                                   ctx->d_vis, vw, ctx->d_astat);
  LAUNCH_CHECK("k_fill_invisible");
  k_repad<<<dim3((stride + 255) / 256, ph, 3), 256, 0, ctx->stream>>>(xyb, (int)stride, (int)ph,
                                                                       (int)W, (int)H, ctx->d_astat);
  LAUNCH_CHECK("k_repad");
  return JXG_OK;
}

int block_factors(jxg_ctx *ctx, const uint8_t *a, uint32_t W, uint32_t H, float *f) {
  const int xb = (int)((W + 7) / 8), yb = (int)((H + 7) / 8);
  k_alpha_blocks<<<dim3((xb + 255) / 256, yb), 256, 0, ctx->stream>>>(a, (int)W, (int)H, xb, yb, f);
  LAUNCH_CHECK("k_alpha_blocks");
  return JXG_OK;
}

int apply_factors(jxg_ctx *ctx, const uint8_t *cells, const float *f, uint32_t W, uint32_t H,
                  uint8_t *qf, const uint32_t *st) {
  const int xb = (int)((W + 7) / 8), yb = (int)((H + 7) / 8);
  k_alpha_apply<<<dim3((xb + 255) / 256, yb), 256, 0, ctx->stream>>>(cells, f, xb, yb, qf, st);
  LAUNCH_CHECK("k_alpha_apply");
  return JXG_OK;
}

}  // namespace

// Front-end alpha stages (called by jxg_front_end_dev after XYB): the alpha
// plane into ctx->d_alpha, fringe clearing, the fill and the block factors.
JXG_HIDDEN int alpha_front(jxg_ctx *ctx, const uint8_t *d_in, size_t row_bytes, int ch,
                           uint32_t W, uint32_t H, int flags, float *xyb, uint32_t stride,
                           uint32_t ph) {
  const size_t n = (size_t)W * H;
  int rc;
  if ((rc = ensure_stat(ctx))) return rc;
  if ((rc = ensure(&ctx->d_alpha, &ctx->alpha_cap, n))) return rc;
  k_astat_init<<<1, 64, 0, ctx->stream>>>(ctx->d_astat);
  k_alpha_extract<<<dim3((W + 255) / 256, H), 256, 0, ctx->stream>>>(d_in, row_bytes, ch, (int)W,
                                                                     (int)H, ctx->d_alpha);
  LAUNCH_CHECK("k_alpha_extract");
  if ((flags & JXG_FLAG_LOSSY_ALPHA) && (rc = fringe_clear(ctx, ctx->d_alpha, W, H))) return rc;
  if (!(flags & JXG_FLAG_FILL_INVISIBLE)) return JXG_OK;
  int vw;
  if ((rc = visibility(ctx, ctx->d_alpha, W, H, &vw))) return rc;
  if ((rc = fill_planes(ctx, xyb, stride, ph, W, H, vw))) return rc;
  const size_t nb = (size_t)((W + 7) / 8) * ((H + 7) / 8);
  if ((rc = ensure(&ctx->d_afac, &ctx->afac_cap, nb))) return rc;
  return block_factors(ctx, ctx->d_alpha, W, H, ctx->d_afac);
}

// After the layout and quant-field decisions: the alpha factors on the
// final quant field, when the frame's alpha is not constant.
JXG_HIDDEN int alpha_finish(jxg_ctx *ctx, uint32_t W, uint32_t H, const uint8_t *cells,
                            uint8_t *qf) {
  return apply_factors(ctx, cells, ctx->d_afac, W, H, qf, ctx->d_astat);
}

extern "C" {

int jxg_alpha_fringe_clear_dev(jxg_ctx *ctx, uint8_t *d_alpha, uint32_t width, uint32_t height) {
  if (!ctx || !d_alpha) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = ensure_stat(ctx))) return rc;
  k_astat_init<<<1, 64, 0, ctx->stream>>>(ctx->d_astat);
  return fringe_clear(ctx, d_alpha, width, height);
}

int jxg_fill_invisible_dev(jxg_ctx *ctx, float *d_xyb, uint32_t stride, uint32_t padded_h,
                           uint32_t width, uint32_t height, const uint8_t *d_alpha) {
  if (!ctx || !d_xyb || !d_alpha) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = ensure_stat(ctx))) return rc;
  k_astat_init<<<1, 64, 0, ctx->stream>>>(ctx->d_astat);
  int vw;
  if ((rc = visibility(ctx, d_alpha, width, height, &vw))) return rc;
  return fill_planes(ctx, d_xyb, stride, padded_h, width, height, vw);
}

int jxg_alpha_block_factors_dev(jxg_ctx *ctx, const uint8_t *d_alpha, uint32_t width,
                                uint32_t height, float *d_factors) {
  if (!ctx || !d_alpha || !d_factors) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  return block_factors(ctx, d_alpha, width, height, d_factors);
}

int jxg_apply_alpha_factors_dev(jxg_ctx *ctx, const uint8_t *d_cells, const float *d_factors,
                                uint32_t width, uint32_t height, uint8_t *d_qf) {
  if (!ctx || !d_cells || !d_factors || !d_qf) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  int rc = check_image(width, height);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  return apply_factors(ctx, d_cells, d_factors, width, height, d_qf, nullptr);
}

int jxg_alpha_plane(jxg_ctx *ctx, uint8_t *dst, size_t bytes, int to_host) {
  if (!ctx || !dst) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  if (!ctx->d_alpha || ctx->alpha_used == 0)
    return set_err(JXG_ERR_INVALID_ARG, "the last front end call had no alpha");
  if (bytes < ctx->alpha_used)
    return set_err(JXG_ERR_INVALID_ARG, "buffer of %zu bytes, alpha plane has %zu", bytes,
                   ctx->alpha_used);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMemcpyAsync(dst, ctx->d_alpha, ctx->alpha_used,
                         to_host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, ctx->stream));
  if (to_host) HIP_TRY(hipStreamSynchronize(ctx->stream));
  return JXG_OK;
}

}  // extern "C"
