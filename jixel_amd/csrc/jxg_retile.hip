// jxg_retile.hip - the coefficient-model re-tiling of the Fast and Fastest
// strategy decision (layout_dp::retile_coefficient,
// src/ac_strategy/selection/layout_dp.rs:1737-2430, called from
// src/ac_strategy/selection.rs:2194-2208), then adjust_quant_field
// (selection.rs:503).
//
//   k_piece_costs  one 256-thread workgroup per 32x32-px region: the
//                  coefficient-model cost (strategy_cost with META_R and the
//                  region quantizer of the field before its adjustment) of a
//                  DCT8, a DCT8X16 (two blocks wide), a DCT16X8 (two tall)
//                  and a DCT16X16 anchored at each of its 16 blocks.  The
//                  region and its one-block right / bottom margin are staged
//                  in LDS once; each shape's sites are transformed there in
//                  the reference's pass order and scored like k_select's
//                  candidates (sums in storage order, f32).
//   k_retile       one wavefront per 8x8-block tile of a DC group (the
//                  aligned grid, then at Fast the grid shifted by (4, 4)):
//                  the exact tiling of plan_small (a broken-profile DP over
//                  the cells in row-major order, 2^(w+1) frontier states)
//                  around and over the incumbent 32x32 blocks; the new pieces
//                  replace the old ones when they price lower.
//   k_adjust_qf    per transform: the aggregate quant value over its blocks.
//
// The DP runs one lane per frontier state, pulling each state's at most five
// predecessors.  The reference relaxes states in the order they entered the
// layer and keeps the first strictly cheaper arrival, so each state keeps
// the minimum of (cost, rank of the predecessor * 8 + choice), and its rank
// in the next layer is the rank of its first arrival among all states, read
// off a bitmap of arrival keys with population counts.  That reproduces the
// reference's choice among equal-cost tilings.
#include <hip/hip_runtime.h>

#include "jxg_batch.h"
#include "jxg_ctx.h"
#include "jxg_device.h"
#include "jxg_internal.h"

using namespace jxg;

#define RT_RD_LAMBDA 0.08086717f
#define RT_META_R 4.0f

namespace {

// ---------------------------------------------------------------------------
// Piece costs

// The four shapes in the planner's choice order (leaf, wide, tall, square):
// pixel width / height, raw strategy, first pass, and whether the storage
// index is column-major in place (k = c * H + r) or row-major (r * W + c).
// src/dct.rs: DCT8 rows first, stored transposed; DCT8X16 (16 wide) rows
// first, row-major; DCT16X8 (16 tall) columns first, stored transposed;
// DCT16X16 columns first, stored transposed.
template <int SH>
struct PieceShape;
template <>
struct PieceShape<0> {
  static constexpr int W = 8, H = 8, S = JXG_DCT, NS = 16;
  static constexpr bool ROWS_FIRST = true, TRANS = true;
};
template <>
struct PieceShape<1> {
  static constexpr int W = 16, H = 8, S = JXG_DCT8X16, NS = 16;
  static constexpr bool ROWS_FIRST = true, TRANS = false;
};
template <>
struct PieceShape<2> {
  static constexpr int W = 8, H = 16, S = JXG_DCT16X8, NS = 16;
  static constexpr bool ROWS_FIRST = false, TRANS = true;
};
template <>
struct PieceShape<3> {
  static constexpr int W = 16, H = 16, S = JXG_DCT16X16, NS = 8;  // two batches of 8 sites
  static constexpr bool ROWS_FIRST = false, TRANS = true;
};

#define PC_WIN 40
#define PC_WP 41
#define PC_WORK 2304  // the largest batch: 16 sites of 16 x 9 (tall pieces)

struct PcLds {
  float win[3][PC_WIN * PC_WP];  // the region and its right / bottom block margin
  float wy[PC_WORK];             // luma coefficients of the batch's sites
  float wc[PC_WORK];             // X or B coefficients
  float sq[2048], mg[2048];      // per coefficient: squared error, rate term (storage order)
  float quot[3][64];             // 0, 1 - one, numerator / a
  float rate[256];               // log2(1 + i)
  float qac[16];                 // per site of the shape: region quantizer
  uint32_t nz[16], mx[16];       // nonzero count, 1 + last nonzero scan position
  float sse[3][16], bits[3][16];
};

struct PieceArgs {
  const float *xyb;
  int stride, ph, W, H, xb, yb;
  const uint8_t *qf;  // the field before adjust_quant_field
  const int8_t *ytox, *ytob;
  const float *tables;
  const jxg_derived *D;
  const uint16_t *scanpos;  // storage index -> scan position (DCT8 at 0, 2x1 at 64, 2x2 at 192)
  float *pc;                // 4 planes of xb * yb: leaf, wide, tall, square
};

__device__ __forceinline__ uint8_t rt_aggregate_quant(uint8_t max_q, uint32_t sum, int covered,
                                                      float d) {
  if (covered < 4) return max_q;
  const float mixer = fmaxf(1.0f - fmaxf(d - 1.54138f, 0.0f) * 0.56391f, 0.0f);
  const float mean = (float)sum / (float)covered;
  const float mixed = (float)max_q * mixer + mean * (1.0f - mixer);
  return (uint8_t)fminf(fmaxf(mixed + 0.5f, 1.0f), 255.0f);
}

// One batch of one shape: sites b * NS .. b * NS + NS - 1 of the region's 16
// anchors (anchor a at block (a & 3, a >> 2)).  Writes the sites' costs to
// cost[] (NaN where the piece leaves the image or a 256-px AC group).
template <bool F, int SH>
__device__ void piece_batch(PcLds &L, const PieceArgs &A, int b, int rbx, int rby, float fx,
                            float fb, float *cost) {
  using P = PieceShape<SH>;
  constexpr int W = P::W, H = P::H, NS = P::NS, PITCH = W + 1, SS = H * PITCH, SIZE = W * H;
  constexpr int SW = W > H ? W : H, SHH = W > H ? H : W;  // storage width / height
  constexpr int CX = SW / 8, CY = SHH / 8;
  constexpr int CW = W / 8, CH = H / 8;  // covered blocks
  constexpr int KIND = SH == 0 ? JXG_QT_DCT8 : (SH == 3 ? JXG_QT_DCT16X16 : JXG_QT_DCT16X8);
  constexpr int SPOFF = SH == 0 ? 0 : (SH == 3 ? 192 : 64);
  const int tid = threadIdx.x, lane = tid & 63;
  const jxg_derived *D = A.D;
  const int cls = jxg_area_class(CX, CY);
  const uint32_t qoff = D->qt_off[KIND], qn = D->qt_n[KIND];
  // the sites' quantizers: region_qac of the footprint (max starts at 1)
  if (tid < NS) {
    const int a = b * NS + tid, ax = rbx + (a & 3), ay = rby + (a >> 2);
    uint8_t max_q = 1;
    uint32_t sum = 0;
    for (int y = ay; y < ay + CH; y++)
      for (int x = ax; x < ax + CW; x++) {
        const uint8_t q = (x < A.xb && y < A.yb) ? A.qf[(size_t)y * A.xb + x] : 1;
        max_q = q > max_q ? q : max_q;
        sum += q;
      }
    L.qac[tid] = D->scale * (float)rt_aggregate_quant(max_q, sum, CW * CH, D->distance);
    L.nz[tid] = 0;
    L.mx[tid] = 0;
  }
  auto stage = [&](float *dst, int c) {
    for (int e = tid; e < NS * SIZE; e += 256) {
      const int site = e / SIZE, rem = e - site * SIZE, r = rem / W, col = rem - r * W;
      const int a = b * NS + site;
      dst[site * SS + r * PITCH + col] = L.win[c][((a >> 2) * 8 + r) * PC_WP + (a & 3) * 8 + col];
    }
  };
  auto transform = [&](float *buf) {
    const float scale = 1.0f / (float)(W * H);
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
      const bool rows = P::ROWS_FIRST ? (pass == 0) : (pass == 1);
      if (rows) {
        for (int e = tid; e < NS * H; e += 256) {
          const int site = e / H, r = e - site * H;
          dct_line<F, W>(buf + site * SS + r * PITCH, 1, scale, pass == 1);
        }
      } else {
        for (int e = tid; e < NS * W; e += 256) {
          const int site = e / W, col = e - site * W;
          dct_line<F, H>(buf + site * SS + col, PITCH, scale, pass == 1);
        }
      }
      __syncthreads();
    }
  };
  // Coefficient terms of channel c (sse_and_rate_scalar_impl,
  // src/inflated_cost.rs:435-500) and the per-site sums in storage order.
  auto channel = [&](int c) {
    const float *inv = A.tables + qoff + (3 + c) * qn;
    const float qm = D->qm_mult[c], one_num = D->dequant_numerator;
    const float f = c == 0 ? fx : fb;
    const float t0 = D->thr[c][cls][0], t1 = D->thr[c][cls][1];
    const float t2 = D->thr[c][cls][2], t3 = D->thr[c][cls][3];
    for (int e0 = 0; e0 < NS * SIZE; e0 += 256) {
      const int e = e0 + tid;
      const int site = __builtin_amdgcn_readfirstlane((e0 + (tid & ~63)) / SIZE);
      const int k = e - site * SIZE;
      const int r = P::TRANS ? k % H : k / W, col = P::TRANS ? k / H : k % W;
      const int srow = k / SW, scol = k - srow * SW;
      const int off = site * SS + r * PITCH + col;
      const float yv = L.wy[off];
      const float coef = c == 1 ? yv : fmla<F>(-f, yv, L.wc[off]);
      const bool ac = !(scol < CX && srow < CY);
      const int yfix = srow >= SHH / 2 ? 2 : 0;
      const float thr = scol >= SW / 2 ? (yfix ? t3 : t1) : (yfix ? t2 : t0);
      const float qs = L.qac[site] * qm;
      const float a = inv[k] * qs * coef;
      const float q = (ac && fabsf(a) >= thr) ? round_away(a) : 0.0f;
      const float aq = fabsf(q);
      float lev;
      if (aq < 64.0f)
        lev = q - copysignf(L.quot[c][(int)aq], q);
      else
        lev = q - one_num / q;
      const float diff = a - lev;
      L.sq[e] = ac ? diff * diff : 0.0f;
      float rt = 0.0f;
      if (q != 0.0f) {
        if (aq < 256.0f)
          rt = L.rate[(int)aq];
        else {
          const uint32_t ki = aq >= 4294967295.0f ? 0xffffffffu : (uint32_t)aq;
          rt = ki < JXG_RATE_LUT_N ? D->rate_lut[ki] : log2f(1.0f + aq);
        }
      }
      L.mg[e] = rt;
      const uint64_t m = __ballot(q != 0.0f);
      uint32_t pos1 = q != 0.0f ? (uint32_t)A.scanpos[SPOFF + k] + 1u : 0u;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) pos1 = max(pos1, (uint32_t)__shfl_xor((int)pos1, o));
      if (lane == 0) {
        atomicAdd(&L.nz[site], (uint32_t)__popcll(m));
        atomicMax(&L.mx[site], pos1);
      }
    }
    __syncthreads();
    if (tid < 2 * NS) {
      const int site = tid >> 1, which = tid & 1;
      const float4 *src = reinterpret_cast<const float4 *>((which ? L.mg : L.sq) + site * SIZE);
      float acc = 0.0f;
      for (int k4 = 0; k4 < SIZE / 4; k4 += 4) {
        float4 v[4];
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] = src[k4 + i];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          acc += v[i].x;
          acc += v[i].y;
          acc += v[i].z;
          acc += v[i].w;
        }
      }
      if (which == 0) {
        L.sse[c][site] = acc;
      } else {
        // model_bits (src/inflated_cost.rs:540-546)
        const uint32_t nzc = L.nz[site];
        const float nzf = (float)nzc;
        const float header =
            0.4f * (nzc < 256u ? L.rate[nzc] : (nzc < JXG_RATE_LUT_N ? D->rate_lut[nzc] : log2f(1.0f + nzf)));
        float visited = 0.0f;
        if (nzc != 0) {
          const int64_t span = (int64_t)L.mx[site] - (int64_t)(CX * CY + nzc);
          visited = span > 0 ? (float)span : 0.0f;
        }
        L.bits[c][site] = nzf * 1.6f + 1.0f * acc + header + 0.5f * visited;
      }
    }
    __syncthreads();
    if (tid < NS) {
      L.nz[tid] = 0;
      L.mx[tid] = 0;
    }
  };
  // Y: transform and keep; then X and B against it.
  stage(L.wy, 1);
  __syncthreads();
  transform(L.wy);
  channel(1);
  for (int c = 0; c <= 2; c += 2) {
    stage(L.wc, c);
    __syncthreads();
    transform(L.wc);
    channel(c);
  }
  if (tid < NS) {
    const int a = b * NS + tid, ax = rbx + (a & 3), ay = rby + (a >> 2);
    float d_total = 0.0f, r_total = 0.0f;
    for (int c = 0; c < 3; c++) {
      d_total += D->channel_weight[c] * L.sse[c][tid];
      r_total += L.bits[c][tid];
    }
    // Footprint::visible_fraction (src/inflated_cost.rs:650-660)
    const int fx0 = ax * 8, fy0 = ay * 8;
    const int vw = min(W, max(A.W - fx0, 0)), vh = min(H, max(A.H - fy0, 0));
    const float visible = (vw == W && vh == H) ? 1.0f : (float)(vw * vh) / (float)(W * H);
    const bool valid = ax + CW <= A.xb && ay + CH <= A.yb && (ax >> 5) == ((ax + CW - 1) >> 5) &&
                       (ay >> 5) == ((ay + CH - 1) >> 5);
    cost[a] = valid ? fmla<F>(RT_RD_LAMBDA, r_total + RT_META_R, d_total * visible) : __int_as_float(0x7fc00000);
  }
  __syncthreads();
}

template <bool F>
__device__ void piece_costs_body(const PieceArgs &A, int bix, int biy) {
  __shared__ PcLds L;
  __shared__ float cost[4][16];
  const int tid = threadIdx.x;
  const int rbx = bix * 4, rby = biy * 4;
  const jxg_derived *D = A.D;
  if (rbx >= A.xb || rby >= A.yb) return;
  {
    const float num = D->dequant_numerator;
    if (tid < 192) {
      const int c = tid >> 6, a = tid & 63;
      L.quot[c][a] = a >= 2 ? num / (float)a : (a == 1 ? 1.0f - D->dequant_one[c] : 0.0f);
    }
    L.rate[tid] = D->rate_lut[tid];
  }
  // the region's 40x40 window; float4 chunks past the padded planes are 0
  // (no piece that reads them lies inside the image)
  const size_t plane = (size_t)A.stride * A.ph;
  const int px0 = rbx * 8, py0 = rby * 8;
  for (int e = tid; e < 3 * PC_WIN * (PC_WIN / 4); e += 256) {
    const int c = e / (PC_WIN * 10), rem = e - c * PC_WIN * 10, r = rem / 10, q = rem - r * 10;
    const int y = py0 + r, x = px0 + 4 * q;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (y < A.ph && x + 4 <= A.stride)
      v = *reinterpret_cast<const float4 *>(A.xyb + c * plane + (size_t)y * A.stride + x);
    float *dst = &L.win[c][r * PC_WP + 4 * q];
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
    dst[3] = v.w;
  }
  // cmap_factors (src/ac_strategy.rs:1838-1855): the region's 64x64 tile,
  // clamped to its DC group's tile grid
  const int gx = rbx >> 8, gy = rby >> 8;
  const int xsize = min(256, A.xb - gx * 256), ysize = min(256, A.yb - gy * 256);
  const int lbx = rbx - gx * 256, lby = rby - gy * 256;
  const int tiles_x = (A.W + 63) / 64;
  const int ntx = (xsize + 7) / 8, nty = (ysize + 7) / 8;
  const int tx = min(lbx / 8, ntx - 1), ty = min(lby / 8, nty - 1);
  const size_t ti = (size_t)(gy * 32 + ty) * tiles_x + gx * 32 + tx;
  const float fx = fmla<F>((float)A.ytox[ti], D->cfl_inv_color_factor, D->cfl_base_x);
  const float fb = fmla<F>((float)A.ytob[ti], D->cfl_inv_color_factor, D->cfl_base_b);
  __syncthreads();
  piece_batch<F, 0>(L, A, 0, rbx, rby, fx, fb, cost[0]);
  piece_batch<F, 1>(L, A, 0, rbx, rby, fx, fb, cost[1]);
  piece_batch<F, 2>(L, A, 0, rbx, rby, fx, fb, cost[2]);
  piece_batch<F, 3>(L, A, 0, rbx, rby, fx, fb, cost[3]);
  piece_batch<F, 3>(L, A, 1, rbx, rby, fx, fb, cost[3]);
  if (tid < 64) {
    const int s = tid >> 4, a = tid & 15, x = rbx + (a & 3), y = rby + (a >> 2);
    if (x < A.xb && y < A.yb) A.pc[(size_t)s * A.xb * A.yb + (size_t)y * A.xb + x] = cost[s][a];
  }
}

template <bool F>
__global__ __launch_bounds__(256) void k_piece_costs(PieceArgs A) {
  piece_costs_body<F>(A, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// The tile DP

#define RT_STATES 512

struct RtLds {
  uint8_t choices[64 * RT_STATES];  // per cell and state: the winning choice
  float cost[2][RT_STATES];         // per state: path cost (INF: not reached), double-buffered
  uint16_t rank[2][RT_STATES];      // per state: its rank in the layer
  uint64_t bits[2][64];             // arrival keys of the layer being built, double-buffered
  uint32_t pfx[64];
  float pcost[4][64];  // per cell: leaf, wide, tall, square (INF: not a candidate)
  float score[64];     // per cell: the score of the piece anchored there (current / new)
  uint8_t live[64];
  uint8_t plan[64];  // per cell: the saved plan's choice (0: none)
};

struct RetileArgs {
  uint8_t *cells;
  const float *pc;   // 4 planes of xb * yb
  const float *c32;  // raw DCT32X32 cost at each 32x32 region's first block
  int xb, yb;
  int crossing, unaligned;
};

__device__ __forceinline__ int rt_strategy(int choice) {
  return choice == 1 ? JXG_DCT : (choice == 2 ? JXG_DCT8X16 : (choice == 3 ? JXG_DCT16X8 : JXG_DCT16X16));
}

__device__ __forceinline__ int rt_slot(int s) {
  return s == JXG_DCT ? 0 : (s == JXG_DCT8X16 ? 1 : (s == JXG_DCT16X8 ? 2 : (s == JXG_DCT16X16 ? 3 : -1)));
}

__device__ __forceinline__ uint64_t rt_mask(int x, int y, int s, int w) {
  const uint64_t row = (1ull << cov_x(s)) - 1ull;
  uint64_t m = 0;
  for (int r = 0; r < cov_y(s); r++) m |= row << ((y + r) * w + x);
  return m;
}

__device__ __forceinline__ uint64_t wave_or64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= (uint64_t)__shfl_xor((long long)v, o);
  return v;
}

// plan_small on a w x h tile around the `fixed` cells.  Returns the optimum;
// when it is finite and below replace_below, L.plan receives the plan.
__device__ float rt_plan(RtLds &L, uint64_t fixed, int w, int h, float replace_below) {
  const int lane = threadIdx.x;
  const int cells = w * h, states = 1 << (w + 1), mask = states - 1;
  const int below_bit = 1 << w;
  const int start = (int)(fixed & (uint64_t)mask);
  for (int t = lane; t < RT_STATES; t += 64) {
    L.cost[0][t] = t == start ? 0.0f : INFINITY;
    L.rank[0][t] = 0;
  }
  L.bits[0][lane] = 0;
  L.bits[1][lane] = 0;
  __syncthreads();
  const int jn = (states + 63) >> 6;
  for (int i = 0; i < cells; i++) {
    const int cur = i & 1, nxt = cur ^ 1;
    // the fixed flag of cell i + w + 1, entering the window (0 past the tile)
    const int corner = (w + 1 + i < 64) ? (int)((fixed >> (w + 1 + i)) & 1ull) : 0;
    const int shapes = L.live[i];
    const float c_leaf = L.pcost[0][i], c_wide = L.pcost[1][i];
    const float c_tall = L.pcost[2][i], c_square = L.pcost[3][i];
    float best[8];
    uint32_t first[8];
    uint8_t choice[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      best[j] = INFINITY;
      first[j] = 0xffffffffu;
      choice[j] = 0;
      const int t = lane + 64 * j;
      if (j >= jn || t >= states) continue;
      const int tw = (t >> w) & 1;
      const int low = (t & (below_bit - 1)) << 1;
      // Candidate (predecessor, choice): the least cost wins, then the
      // earliest arrival (predecessor rank, then choice order); the state's
      // own arrival key is the earliest finite candidate.
      uint32_t wkey = 0xffffffffu;
      float wcost = INFINITY;
      uint8_t wch = 0;
      uint32_t fkey = 0xffffffffu;
      auto take = [&](int from, float add, bool plain, int ch) {
        const float c0 = L.cost[cur][from];
        if (!(c0 < INFINITY)) return;
        const float c = plain ? c0 : c0 + add;
        if (!(c < INFINITY)) return;
        const uint32_t key = (uint32_t)L.rank[cur][from] * 8u + (uint32_t)ch;
        if (c < wcost || (c == wcost && key < wkey)) {
          wcost = c;
          wkey = key;
          wch = (uint8_t)ch;
        }
        fkey = min(fkey, key);
      };
      if (tw == corner) {
        take(low | 1, 0.0f, true, 0);  // the cell was filled before
        take(low, c_leaf, false, 1);
        if ((t & 1) && (shapes & 1)) take(low & ~2, c_wide, false, 2);
        if (((t >> (w - 1)) & 1) && (shapes & 2)) take(low & ~below_bit, c_tall, false, 3);
      }
      if (corner == 0 && tw == 1 && (t & 1) && ((t >> (w - 1)) & 1) && (shapes & 4))
        take(low & ~2 & ~below_bit, c_square, false, 4);
      best[j] = wcost;
      first[j] = fkey;
      choice[j] = wch;
    }
    // arrival bitmap of this layer
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (first[j] != 0xffffffffu)
        atomicOr((unsigned long long *)&L.bits[cur][first[j] >> 6], 1ull << (first[j] & 63));
    __syncthreads();
    {
      const uint32_t cnt = (uint32_t)__popcll(L.bits[cur][lane]);
      uint32_t inc = cnt;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)inc, o);
        if (lane >= o) inc += v;
      }
      L.pfx[lane] = inc - cnt;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int t = lane + 64 * j;
      if (j >= jn || t >= states) continue;
      float c = INFINITY;
      uint16_t r = 0;
      if (first[j] != 0xffffffffu) {
        const uint32_t f = first[j], wd = f >> 6;
        r = (uint16_t)(L.pfx[wd] + (uint32_t)__popcll(L.bits[cur][wd] & ((1ull << (f & 63)) - 1ull)));
        c = best[j];
        L.choices[i * RT_STATES + t] = choice[j];
      }
      L.cost[nxt][t] = c;
      L.rank[nxt][t] = r;
    }
    __syncthreads();
    L.bits[cur][lane] = 0;
  }
  const float total = L.cost[cells & 1][0];
  __syncthreads();
  if (!(total < INFINITY && total < replace_below)) return total;
  if (lane == 0) {
    const int added[5] = {0, 0, 2, below_bit, 2 | below_bit | (2 << w)};
    for (int i = 0; i < 64; i++) L.plan[i] = 0;
    int state = 0;
    for (int i = cells - 1; i >= 0; i--) {
      const int ch = L.choices[i * RT_STATES + state];
      if (ch != 0) L.plan[i] = (uint8_t)ch;
      state = (((state << 1) ^ added[ch]) & mask) | (ch == 0 ? 1 : 0);
    }
  }
  __syncthreads();
  return total;
}

// search_tile_coefficient (layout_dp.rs:2033-2199) for tile slot (sx, sy)
// of the pass: 32 slots per DC group and axis.
__device__ void retile_body(RtLds &L, const RetileArgs A, int sx, int sy) {
  const int lane = threadIdx.x;
  const int gx = sx >> 5, gy = sy >> 5;
  const int o = A.crossing ? 4 : 0;
  const int lx = o + 8 * (sx & 31), ly = o + 8 * (sy & 31);
  const int xsize = min(256, A.xb - gx * 256), ysize = min(256, A.yb - gy * 256);
  if (xsize <= 0 || ysize <= 0 || lx >= xsize || ly >= ysize) return;
  const int w = min(8, xsize - lx), h = min(8, ysize - ly), cells = w * h;
  if (cells < 2) return;
  const int X0 = gx * 256 + lx, Y0 = gy * 256 + ly;
  const size_t plane = (size_t)A.xb * A.yb;
  // the tile's current pieces (first blocks; on a shifted grid, those that
  // fit inside the tile)
  const bool in = lane < cells;
  const int x = in ? lane % w : 0, y = in ? lane / w : 0;
  uint8_t cell = 0;
  if (in) cell = A.cells[(size_t)(Y0 + y) * A.xb + X0 + x];
  const int s = cell >> 1;
  const bool first = in && (cell & 1);
  const bool fits = first && x + cov_x(s) <= w && y + cov_y(s) <= h;
  const bool curp = first && (!A.crossing || fits);
  if (!A.crossing) {
    // footprint_is_self_contained (src/ac_strategy/selection.rs:3235-3258)
    if (__ballot(first && !fits) != 0) return;
    uint32_t area = first ? (uint32_t)(cov_x(s) * cov_y(s)) : 0u;
#pragma unroll
    for (int k = 32; k > 0; k >>= 1) area += (uint32_t)__shfl_xor((int)area, k);
    if ((int)area != cells) return;
  }
  if (__ballot(curp && (s == JXG_IDENTITY || s == JXG_DCT2X2)) != 0) return;
  const uint64_t covered = wave_or64(curp ? rt_mask(x, y, s, w) : 0ull);
  const uint64_t full = cells == 64 ? ~0ull : ((1ull << cells) - 1ull);
  const uint64_t blocked = A.crossing ? (full & ~covered) : 0ull;
  // candidate costs of the four shapes at each cell
  {
    const int gxl = lx + x, gyl = ly + y;  // group-relative
#pragma unroll
    for (int ch = 1; ch <= 4; ch++) {
      const int ps = rt_strategy(ch), cw = cov_x(ps), chh = cov_y(ps);
      float v = INFINITY;
      if (in && x + cw <= w && y + chh <= h && !(blocked && (rt_mask(x, y, ps, w) & blocked))) {
        const bool off_grid = ch == 2 ? (gxl & 1) : (ch == 3 ? (gyl & 1) : (ch == 4 ? ((gxl | gyl) & 1) : false));
        bool legal = !(off_grid && !A.unaligned);
        if (legal) {
          legal = gxl + cw <= xsize && gyl + chh <= ysize && (gxl >> 5) == ((gxl + cw - 1) >> 5) &&
                  (gyl >> 5) == ((gyl + chh - 1) >> 5);
          if (!A.crossing)
            legal = legal && (gxl >> 3) == ((gxl + cw - 1) >> 3) && (gyl >> 3) == ((gyl + chh - 1) >> 3);
        }
        if (legal) v = A.pc[(size_t)(ch - 1) * plane + (size_t)(Y0 + y) * A.xb + X0 + x];
      }
      L.pcost[ch - 1][lane] = v;
    }
  }
  // the incumbent's scores
  float my_score = 0.0f;
  if (curp) {
    const int sl = rt_slot(s);
    const size_t at = (size_t)(Y0 + y) * A.xb + X0 + x;
    my_score = sl >= 0 ? A.pc[(size_t)sl * plane + at] : A.c32[at];
  }
  const bool fixedp = curp && rt_slot(s) < 0;
  const uint64_t fixed = wave_or64(fixedp ? rt_mask(x, y, s, w) : 0ull);
  L.score[lane] = my_score;
  __syncthreads();
  // SmallCosts::live_masks (layout_dp.rs:1860-1916)
  {
    uint8_t lv = 0;
    if (in) {
      const int i = lane;
      const float *lf = L.pcost[0], *wd = L.pcost[1], *tl = L.pcost[2], *sq = L.pcost[3];
      if (y + 1 == h) {
        if (x + 1 < w) lv = wd[i] < lf[i] + lf[i + 1];
      } else if (x + 1 < w) {
        const int bl = i + w;
        const float top = lf[i] + lf[i + 1];
        const float leaves = (top + lf[bl]) + lf[bl + 1];
        const float wides = wd[i] + wd[bl];
        const float talls = tl[i] + tl[i + 1];
        const float bst = fminf(fminf(leaves, wides), talls);
        lv = (uint8_t)((wd[i] < top) | ((tl[i] < lf[i] + lf[bl]) << 1) | ((sq[i] < bst) << 2));
      } else {
        lv = (uint8_t)((tl[i] < lf[i] + lf[i + w]) << 1);
      }
    }
    L.live[lane] = lv;
  }
  // fixed_cost: the incumbents outside the candidate shapes, in raster order
  float fixed_cost = 0.0f;
  {
    uint64_t fm = __ballot(fixedp);
    while (fm) {
      const int i = __builtin_ctzll(fm);
      fm &= fm - 1;
      fixed_cost += __shfl(my_score, i);
    }
  }
  __syncthreads();
  const float around = rt_plan(L, fixed | blocked, w, h, INFINITY) + fixed_cost;
  bool keep_fixed = true;
  float winner = around;
  if (fixed) {
    const float over = rt_plan(L, blocked, w, h, around);
    if (over < around) {
      winner = over;
      keep_fixed = false;
    }
  }
  if (!(winner < INFINITY)) return;  // winner is never NaN: sums of costs
  // the new layout at each cell, and the old / new sums over the pieces
  // that differ, both in raster order (the sorted layout's order)
  const int pch = in ? L.plan[lane] : 0;
  int ns = -1;
  if (pch) ns = rt_strategy(pch);
  else if (keep_fixed && fixedp) ns = s;
  const bool in_now = curp && ns == s;  // the current piece survives
  const bool in_was = ns >= 0 && curp && s == ns;
  float new_score = 0.0f;
  if (ns >= 0 && !in_was) new_score = L.pcost[rt_slot(ns)][lane];
  float old_cost = 0.0f, new_cost = 0.0f;
  {
    uint64_t om = __ballot(curp && !in_now), nm = __ballot(ns >= 0 && !in_was);
    while (om) {
      const int i = __builtin_ctzll(om);
      om &= om - 1;
      old_cost += __shfl(my_score, i);
    }
    while (nm) {
      const int i = __builtin_ctzll(nm);
      nm &= nm - 1;
      new_cost += __shfl(new_score, i);
    }
  }
  if (!(new_cost < old_cost)) return;
  if (ns >= 0 && !in_was) {
    const int cw = cov_x(ns), chh = cov_y(ns);
    for (int yy = 0; yy < chh; yy++)
      for (int xx = 0; xx < cw; xx++)
        A.cells[(size_t)(Y0 + y + yy) * A.xb + X0 + x + xx] =
            (uint8_t)((ns << 1) | ((xx == 0 && yy == 0) ? 1 : 0));
  }
}

__global__ __launch_bounds__(64) void k_retile(RetileArgs A) {
  __shared__ RtLds L;
  retile_body(L, A, blockIdx.x, blockIdx.y);
}

// adjust_quant_field (src/ac_strategy/selection.rs:503-527): one thread per
// block; a multi-block transform's first block writes the aggregate over its
// blocks (transforms are disjoint, so the update is in place).
__device__ __forceinline__ void adjust_body(const uint8_t *cells, uint8_t *qf, int xb, int yb,
                                            float d, int i) {
  if (i >= xb * yb) return;
  const uint8_t cell = cells[i];
  if (!(cell & 1)) return;
  const int s = cell >> 1, cw = cov_x(s), chh = cov_y(s);
  if (cw == 1 && chh == 1) return;
  const int x = i % xb, y = i / xb;
  uint8_t max_q = 0;
  uint32_t sum = 0;
  for (int yy = 0; yy < chh; yy++)
    for (int xx = 0; xx < cw; xx++) {
      const uint8_t q = qf[(size_t)(y + yy) * xb + x + xx];
      max_q = q > max_q ? q : max_q;
      sum += q;
    }
  const uint8_t v = rt_aggregate_quant(max_q, sum, cw * chh, d);
  for (int yy = 0; yy < chh; yy++)
    for (int xx = 0; xx < cw; xx++) qf[(size_t)(y + yy) * xb + x + xx] = v;
}

__global__ __launch_bounds__(256) void k_adjust_qf(const uint8_t *cells, uint8_t *qf, int xb, int yb,
                                                   float d) {
  adjust_body(cells, qf, xb, yb, d, blockIdx.x * 256 + threadIdx.x);
}

// ---- batch forms: the same bodies behind the flat index map ----

template <bool F>
__global__ __launch_bounds__(256) void k_piece_costs_b(const BatchImg *__restrict__ img,
                                                       const uint32_t *__restrict__ pfx, int n,
                                                       const uint16_t *__restrict__ scanpos) {
  const int bi = batch_image(pfx, n, blockIdx.x);
  const BatchImg &I = img[bi];
  if (!I.D->retile) return;
  const uint32_t bl = blockIdx.x - pfx[bi];
  PieceArgs A;
  A.xyb = I.xyb;
  A.stride = I.stride;
  A.ph = I.ph;
  A.W = I.W;
  A.H = I.H;
  A.xb = I.xb;
  A.yb = I.yb;
  A.qf = I.qf;
  A.ytox = I.ytox;
  A.ytob = I.ytob;
  A.tables = I.tables;
  A.D = I.D;
  A.scanpos = scanpos;
  A.pc = I.pc;
  const int gw = (I.xb + 3) / 4;
  piece_costs_body<F>(A, (int)(bl % (uint32_t)gw), (int)(bl / (uint32_t)gw));
}

__global__ __launch_bounds__(64) void k_retile_b(const BatchImg *__restrict__ img,
                                                 const uint32_t *__restrict__ pfx, int n,
                                                 int crossing) {
  __shared__ RtLds L;
  const int bi = batch_image(pfx, n, blockIdx.x);
  const BatchImg &I = img[bi];
  if (!I.D->retile || (crossing && I.D->retile < 2)) return;
  const uint32_t bl = blockIdx.x - pfx[bi];
  RetileArgs A;
  A.cells = I.cells;
  A.pc = I.pc;
  A.c32 = I.c32;
  A.xb = I.xb;
  A.yb = I.yb;
  A.crossing = crossing;
  A.unaligned = I.D->retile >= 2;
  const int gw = 32 * ((I.xb + 255) / 256);
  retile_body(L, A, (int)(bl % (uint32_t)gw), (int)(bl / (uint32_t)gw));
}

__global__ __launch_bounds__(256) void k_adjust_qf_b(const BatchImg *__restrict__ img,
                                                     const uint32_t *__restrict__ pfx, int n) {
  const int bi = batch_image(pfx, n, blockIdx.x);
  const BatchImg &I = img[bi];
  if (!I.D->retile) return;
  const uint32_t bl = blockIdx.x - pfx[bi];
  adjust_body(I.cells, I.qf, I.xb, I.yb, I.distance, (int)(bl * 256 + threadIdx.x));
}

}  // namespace

// Host side: the re-tiling stages after k_select (single image), on `st`.
JXG_HIDDEN int retile_stages(jxg_ctx *ctx, hipStream_t st, const float *xyb, uint32_t stride,
                             uint32_t ph, uint32_t W, uint32_t H, uint8_t *qf, const int8_t *ytox,
                             const int8_t *ytob, const float *tables, const jxg_frame_params *p,
                             int mode, uint8_t *cells, const float *c32, float *pc) {
  const int xb = (int)((W + 7) / 8), yb = (int)((H + 7) / 8);
  PieceArgs A;
  A.xyb = xyb;
  A.stride = (int)stride;
  A.ph = (int)ph;
  A.W = (int)W;
  A.H = (int)H;
  A.xb = xb;
  A.yb = yb;
  A.qf = qf;
  A.ytox = ytox;
  A.ytob = ytob;
  A.tables = tables;
  A.D = ctx->d_derived;
  A.scanpos = ctx->d_scanpos;
  A.pc = pc;
  const dim3 g1((xb + 3) / 4, (yb + 3) / 4);
  prof_begin(ctx, JXG_PROF_RETILE, st);
  if (p->fused)
    k_piece_costs<true><<<g1, 256, 0, st>>>(A);
  else
    k_piece_costs<false><<<g1, 256, 0, st>>>(A);
  LAUNCH_CHECK("k_piece_costs");
  RetileArgs R;
  R.cells = cells;
  R.pc = pc;
  R.c32 = c32;
  R.xb = xb;
  R.yb = yb;
  R.unaligned = mode >= 2;
  const dim3 g2(32 * ((xb + 255) / 256), 32 * ((yb + 255) / 256));
  R.crossing = 0;
  k_retile<<<g2, 64, 0, st>>>(R);
  LAUNCH_CHECK("k_retile");
  if (mode >= 2) {
    R.crossing = 1;
    k_retile<<<g2, 64, 0, st>>>(R);
    LAUNCH_CHECK("k_retile (shifted)");
  }
  k_adjust_qf<<<(xb * yb + 255) / 256, 256, 0, st>>>(cells, qf, xb, yb, p->distance);
  LAUNCH_CHECK("k_adjust_qf");
  prof_end(ctx, JXG_PROF_RETILE, st);
  return JXG_OK;
}

// Batch: the same stages over every image (images whose tier does not
// re-tile return at once).
JXG_HIDDEN int batch_retile(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                            const uint32_t *d_pfx, int n, const uint32_t total[], int fused) {
  const uint32_t *p = d_pfx;
  const int m = n + 1;
  if (fused)
    k_piece_costs_b<true><<<total[BS_PIECE], 256, 0, st>>>(d_img, p + BS_PIECE * m, n, ctx->d_scanpos);
  else
    k_piece_costs_b<false><<<total[BS_PIECE], 256, 0, st>>>(d_img, p + BS_PIECE * m, n, ctx->d_scanpos);
  k_retile_b<<<total[BS_RETILE], 64, 0, st>>>(d_img, p + BS_RETILE * m, n, 0);
  k_retile_b<<<total[BS_RETILE], 64, 0, st>>>(d_img, p + BS_RETILE * m, n, 1);
  k_adjust_qf_b<<<total[BS_ADJUST], 256, 0, st>>>(d_img, p + BS_ADJUST * m, n);
  LAUNCH_CHECK("batch retile");
  return JXG_OK;
}
