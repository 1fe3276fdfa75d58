// jxg_retile.hip - the coefficient-model re-tiling of the Fast and Fastest
// strategy decision (layout_dp::retile_coefficient,
// src/ac_strategy/selection/layout_dp.rs:1737-2430, called from
// src/ac_strategy/selection.rs:2194-2208), then adjust_quant_field
// (selection.rs:503).
//
//   k_piece_costs  one 256-thread workgroup per 32x32-px region: the
//                  coefficient-model cost (strategy_cost with META_R and the
//                  region quantizer of the field before its adjustment) of a
//                  DCT8X16 (two blocks wide), a DCT16X8 (two tall) and a
//                  DCT16X16 anchored at each of its 16 blocks (and a DCT8
//                  where the strategy search has not priced one).  Each
//                  shape's sites are staged from the planes into LDS,
//                  transformed in the reference's pass order and scored like
//                  k_select's candidates; the squared errors are summed in
//                  storage order, the rate terms over the nonzero levels
//                  only (a zero term adds nothing to a non-negative f32 sum).
//   k_retile       one 256-thread workgroup per 8x8-block tile of a DC
//                  group (the aligned grid, then at Fast the grid shifted by
//                  (4, 4)): the exact tiling of plan_small (a broken-profile
//                  DP over the cells in row-major order, 2^(w+1) frontier
//                  states) around and over the incumbent 32x32 blocks; the
//                  new pieces replace the old ones when they price lower.
//   k_adjust_qf    per transform: the aggregate quant value over its blocks.
//
// The DP gives each thread two frontier states and has each state pull its
// at most five predecessors.  The reference relaxes states in the order they
// entered the layer and keeps the first strictly cheaper arrival, so each
// state keeps the minimum of (cost, rank of the predecessor * 8 + choice),
// and its rank in the next layer is the rank of its first arrival among all
// states, read off a bitmap of arrival keys with population counts.  That
// reproduces the reference's choice among equal-cost tilings.  A cell that
// is filled before only shifts the states (cost and order kept).
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

#define PC_WORK 2304  // the largest batch: 16 sites of 16 x 9 (tall pieces)

struct PcLds {
  float wy[PC_WORK];         // luma coefficients of the batch's sites
  float wc[PC_WORK];         // X or B coefficients
  float mg[2048];            // per coefficient: rate term (storage order); the squared
                             // errors go to wc once its coefficients are in registers
  uint8_t sp[2048];          // per coefficient: its scan position (< 256 for these shapes)
  uint64_t nzm[16][4];       // per site and 64-coefficient chunk: the nonzero levels
  float quot[3][64];         // 0, 1 - one, numerator / a
  float rate[256];           // log2(1 + i)
  float qac[16];             // per site of the shape: region quantizer
  float sse[3][16], bits[3][16];
  float cost[4][16];
  int need_leaf;             // some block of the region lacks its DCT8 cost
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
  return jxg_aggregate_quant(max_q, sum, covered, d);
}

// One batch of one shape: sites b * NS .. b * NS + NS - 1 of the region's 16
// anchors (anchor a at block (a & 3, a >> 2)).  Writes the sites' costs to
// L.cost[SH] (NaN where the piece leaves the image or a 256-px AC group).
template <bool F, int SH>
__device__ __forceinline__ void piece_batch(PcLds &L, const PieceArgs &A, int b, int rbx, int rby,
                                            float fx, float fb) {
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
  const size_t plane = (size_t)A.stride * A.ph;
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
  }
  // the sites' pixels from the (L2-resident) planes; past the padded
  // planes 0 (no piece that reads them lies inside the image)
  auto stage = [&](float *dst, int c) {
    for (int e = tid; e < NS * SIZE; e += 256) {
      const int site = e / SIZE, rem = e - site * SIZE, r = rem / W, col = rem - r * W;
      const int a = b * NS + site;
      const int y = (rby + (a >> 2)) * 8 + r, x = (rbx + (a & 3)) * 8 + col;
      dst[site * SS + r * PITCH + col] =
          (y < A.ph && x < A.stride) ? A.xyb[c * plane + (size_t)y * A.stride + x] : 0.0f;
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
  // Per thread, its coefficients' geometry (the same for every channel):
  // offset in the transform buffer, AC flag, threshold quadrant; and
  // their scan positions into LDS.
  constexpr int ITERS = NS * SIZE / 256;
  int geo[ITERS];
#pragma unroll
  for (int it = 0; it < ITERS; it++) {
    const int e = it * 256 + tid, site = e / SIZE, k = e - site * SIZE;
    const int r = P::TRANS ? k % H : k / W, col = P::TRANS ? k / H : k % W;
    const int srow = k / SW, scol = k - srow * SW;
    const int ac = !(scol < CX && srow < CY);
    const int tsel = (srow >= SHH / 2 ? 2 : 0) + (scol >= SW / 2 ? 1 : 0);
    geo[it] = (site * SS + r * PITCH + col) | (tsel << 12) | (ac << 14);
    L.sp[e] = (uint8_t)A.scanpos[SPOFF + k];
  }
  // Coefficient terms of channel c (sse_and_rate_scalar_impl,
  // src/inflated_cost.rs:435-500) and the per-site sums in storage order.
  auto channel = [&](int c) {
    const float *inv = A.tables + qoff + (3 + c) * qn;
    const float qm = D->qm_mult[c], one_num = D->dequant_numerator;
    const float f = c == 0 ? fx : fb;
    const float t0 = D->thr[c][cls][0], t1 = D->thr[c][cls][1];
    const float t2 = D->thr[c][cls][2], t3 = D->thr[c][cls][3];
    float invk[ITERS];
#pragma unroll
    for (int it = 0; it < ITERS; it++) {
      const int e = it * 256 + tid;
      invk[it] = inv[e - (e / SIZE) * SIZE];
    }
    float coefs[ITERS];
#pragma unroll
    for (int it = 0; it < ITERS; it++) {
      const int off = geo[it] & 4095;
      const float yv = L.wy[off];
      coefs[it] = c == 1 ? yv : fmla<F>(-f, yv, L.wc[off]);
    }
    __syncthreads();  // every coefficient read before wc takes the squared errors
#pragma unroll
    for (int it = 0; it < ITERS; it++) {
      const int e0 = it * 256;
      const int e = e0 + tid;
      const int site = __builtin_amdgcn_readfirstlane((e0 + (tid & ~63)) / SIZE);
      const int tsel = (geo[it] >> 12) & 3;
      const bool ac = (geo[it] >> 14) & 1;
      const float coef = coefs[it];
      const float thr = tsel >= 2 ? (tsel == 3 ? t3 : t2) : (tsel == 1 ? t1 : t0);
      const float qs = L.qac[site] * qm;
      const float a = invk[it] * qs * coef;
      const float q = (ac && fabsf(a) >= thr) ? round_away(a) : 0.0f;
      const float aq = fabsf(q);
      float lev;
      if (aq < 64.0f)
        lev = q - copysignf(L.quot[c][(int)aq], q);
      else
        lev = q - one_num / q;
      const float diff = a - lev;
      L.wc[e] = ac ? diff * diff : 0.0f;
      float rt = 0.0f;
      if (q != 0.0f) {
        if (aq < 256.0f)
          rt = L.rate[(int)aq];
        else {
          const uint32_t ki = aq >= 4294967295.0f ? 0xffffffffu : (uint32_t)aq;
          rt = jxg_rate_log2(D, aq);
        }
      }
      L.mg[e] = rt;
      const uint64_t m = __ballot(q != 0.0f);
      if (lane == 0) L.nzm[site][(e - site * SIZE) >> 6] = m;
    }
    __syncthreads();
    if (tid < NS) {
      // squared errors: every term, in storage order
      const float4 *src = reinterpret_cast<const float4 *>(L.wc + tid * SIZE);
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
      L.sse[c][tid] = acc;
    } else if (tid >= 64 && tid < 64 + NS) {
      // rate terms: only the nonzero levels contribute (adding a zero term
      // to a non-negative sum changes no bit), in storage order; their
      // count and the last one's scan position on the way
      const int site = tid - 64;
      float acc = 0.0f;
      uint32_t nzc = 0, mx = 0;
#pragma unroll
      for (int ch = 0; ch < SIZE / 64; ch++) {
        uint64_t m = L.nzm[site][ch];
        while (m) {
          const int i = __builtin_ctzll(m);
          m &= m - 1;
          const int e = site * SIZE + ch * 64 + i;
          acc += L.mg[e];
          mx = max(mx, (uint32_t)L.sp[e] + 1u);
          nzc++;
        }
      }
      // model_bits (src/inflated_cost.rs:540-546)
      const float nzf = (float)nzc;
      const float header =
          0.4f * (nzc < 256u ? L.rate[nzc] : jxg_rate_log2(D, nzf));
      float visited = 0.0f;
      if (nzc != 0) {
        const int64_t span = (int64_t)mx - (int64_t)(CX * CY + nzc);
        visited = span > 0 ? (float)span : 0.0f;
      }
      L.bits[c][site] = nzf * 1.6f + 1.0f * acc + header + 0.5f * visited;
    }
    __syncthreads();
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
    L.cost[SH][a] = valid ? fmla<F>(RT_RD_LAMBDA, r_total + RT_META_R, d_total * visible) : __int_as_float(0x7fc00000);
  }
  __syncthreads();
}

template <bool F>
__device__ __forceinline__ void piece_costs_body(const PieceArgs &A, int bix, int biy) {
  __shared__ PcLds L;
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
  // The strategy search has priced the DCT8 of (nearly) every block in the
  // leaf plane already (the same strategy_cost); the leaves are priced here
  // only when one of the region's blocks lacks one.
  if (tid < 64) {
    bool nl = false;
    if (tid < 16) {
      const int x = rbx + (tid & 3), y = rby + (tid >> 2);
      if (x < A.xb && y < A.yb) nl = isnan(A.pc[(size_t)y * A.xb + x]);
    }
    const uint64_t m = __ballot(nl);
    if (tid == 0) L.need_leaf = m != 0 ? 1 : 0;
  }
  __syncthreads();
  const bool need_leaf = L.need_leaf != 0;
  if (need_leaf) piece_batch<F, 0>(L, A, 0, rbx, rby, fx, fb);
  piece_batch<F, 1>(L, A, 0, rbx, rby, fx, fb);
  piece_batch<F, 2>(L, A, 0, rbx, rby, fx, fb);
  piece_batch<F, 3>(L, A, 0, rbx, rby, fx, fb);
  piece_batch<F, 3>(L, A, 1, rbx, rby, fx, fb);
  if (tid < 64) {
    const int s = tid >> 4, a = tid & 15, x = rbx + (a & 3), y = rby + (a >> 2);
    if ((s > 0 || need_leaf) && x < A.xb && y < A.yb)
      A.pc[(size_t)s * A.xb * A.yb + (size_t)y * A.xb + x] = L.cost[s][a];
  }
}

template <bool F>
__global__ __launch_bounds__(256) void k_piece_costs(PieceArgs A) {
  piece_costs_body<F>(A, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// The tile DP

#define RT_STATES 512

#define RT_T 256  // threads per tile: two frontier states each

struct RtLds {
  // per cell and thread: the winning choices of the thread's two states
  // (state tid in bits 0..3, state tid + 256 in bits 4..7)
  uint8_t choices[64][RT_T];
  uint2 st[2][RT_STATES];  // per state: path cost bits (INF: not reached), rank in the layer
  uint64_t bits[2][64];    // arrival keys of the layer being built, double-buffered
  uint32_t pfx[64];
  float pcost[4][64];  // per cell: leaf, wide, tall, square (INF: not a candidate)
  float score[64];     // per cell: the score of the piece anchored there
  uint8_t live[64];
  uint8_t plan[64];  // per cell: the saved plan's choice (0: none)
  uint64_t fixed, blocked;
  float fixed_cost;
  int go;
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
// Thread t holds states t and t + 256; each state pulls its (at most five)
// predecessors, all of a layer's reads issued before any is used.  A cell
// that is filled before (fixed) only shifts the states, keeping their cost
// and order.
__device__ __forceinline__ float rt_plan(RtLds &L, uint64_t fixed, int w, int h, float replace_below) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cells = w * h, states = 1 << (w + 1), mask = states - 1;
  const int below_bit = 1 << w;
  const int start = (int)(fixed & (uint64_t)mask);
  const uint32_t inf_bits = 0x7f800000u;
  for (int t = tid; t < RT_STATES; t += RT_T) L.st[0][t] = make_uint2(t == start ? 0u : inf_bits, 0u);
  if (tid < 128) (&L.bits[0][0])[tid] = 0;
  __syncthreads();
  const int jn = (states + RT_T - 1) / RT_T;  // 1 or 2, uniform
  for (int i = 0; i < cells; i++) {
    const int cur = i & 1, nxt = cur ^ 1;
    // the fixed flag of cell i + w + 1, entering the window (0 past the tile)
    const int corner = (w + 1 + i < 64) ? (int)((fixed >> (w + 1 + i)) & 1ull) : 0;
    if ((fixed >> i) & 1ull) {
#pragma clang loop unroll(full)
      for (int j = 0; j < 2; j++) {
        const int t = tid + RT_T * j;
        if (j < jn && t < states) {
          const int low = (t & (below_bit - 1)) << 1;
          L.st[nxt][t] = ((t >> w) & 1) == corner ? L.st[cur][low | 1] : make_uint2(inf_bits, 0u);
        }
      }
      L.choices[i][tid] = 0;
      __syncthreads();
      continue;
    }
    const int shapes = L.live[i];
    const float add[5] = {0.0f, L.pcost[0][i], L.pcost[1][i], L.pcost[2][i], L.pcost[3][i]};
    uint2 sv[2][5];
    int valid[2];
#pragma clang loop unroll(full)
    for (int j = 0; j < 2; j++) {
      const int t = tid + RT_T * j;
      const int low = (t & (below_bit - 1)) << 1;
      const int tw = (t >> w) & 1, b0 = t & 1, bw = (t >> (w - 1)) & 1;
      const bool in = j < jn && t < states;
      const bool same = in && tw == corner;
      valid[j] = (same ? 3 : 0) | ((same && b0 && (shapes & 1)) ? 4 : 0) |
                 ((same && bw && (shapes & 2)) ? 8 : 0) |
                 ((in && corner == 0 && tw == 1 && b0 && bw && (shapes & 4)) ? 16 : 0);
      if (j < jn) {
        sv[j][0] = L.st[cur][low | 1];                // the cell was filled before
        sv[j][1] = L.st[cur][low];                    // leaf
        sv[j][2] = L.st[cur][low & ~2];               // wide
        sv[j][3] = L.st[cur][low & ~below_bit];       // tall
        sv[j][4] = L.st[cur][low & ~2 & ~below_bit];  // square
      }
    }
    float best[2];
    uint32_t first[2], choice[2];
#pragma clang loop unroll(full)
    for (int j = 0; j < 2; j++) {
      // the least cost wins, then the earliest arrival (predecessor rank,
      // then choice order); the state's own arrival key is its earliest
      // finite candidate
      float wcost = INFINITY;
      uint32_t wkey = 0xffffffffu, wch = 0, fkey = 0xffffffffu;
#pragma clang loop unroll(full)
      for (int k = 0; k < 5; k++) {
        const float c0 = __uint_as_float(sv[j][k].x);
        const float c = k == 0 ? c0 : c0 + add[k];
        const bool ok = ((valid[j] >> k) & 1) && c0 < INFINITY && c < INFINITY;
        const uint32_t key = sv[j][k].y * 8u + (uint32_t)k;
        const bool better = ok && (c < wcost || (c == wcost && key < wkey));
        wcost = better ? c : wcost;
        wkey = better ? key : wkey;
        wch = better ? (uint32_t)k : wch;
        fkey = ok ? min(fkey, key) : fkey;
      }
      best[j] = wcost;
      first[j] = fkey;
      choice[j] = wch;
      if (fkey != 0xffffffffu)
        atomicOr((unsigned long long *)&L.bits[cur][fkey >> 6], 1ull << (fkey & 63));
    }
    __syncthreads();
    if (wave == 0) {
      // ranks: the arrival keys in increasing order
      const uint32_t cnt = (uint32_t)__popcll(L.bits[cur][lane]);
      uint32_t inc = cnt;
#pragma clang loop unroll(full)
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)inc, o);
        if (lane >= o) inc += v;
      }
      L.pfx[lane] = inc - cnt;
    }
    __syncthreads();
    uint32_t packed = 0;
#pragma clang loop unroll(full)
    for (int j = 0; j < 2; j++) {
      const int t = tid + RT_T * j;
      if (j >= jn || t >= states) continue;
      uint2 out = make_uint2(inf_bits, 0u);
      if (first[j] != 0xffffffffu) {
        const uint32_t f = first[j], wd = f >> 6;
        out.y = L.pfx[wd] + (uint32_t)__popcll(L.bits[cur][wd] & ((1ull << (f & 63)) - 1ull));
        out.x = __float_as_uint(best[j]);
        packed |= choice[j] << (4 * j);
      }
      L.st[nxt][t] = out;
    }
    L.choices[i][tid] = (uint8_t)packed;
    __syncthreads();
    if (tid < 64) L.bits[cur][tid] = 0;
  }
  const float total = __uint_as_float(L.st[cells & 1][0].x);
  __syncthreads();
  if (!(total < INFINITY && total < replace_below)) return total;
  if (tid == 0) {
    for (int i = 0; i < 64; i++) L.plan[i] = 0;
    int state = 0;
    for (int i = cells - 1; i >= 0; i--) {
      const int ch = (int)((L.choices[i][state & (RT_T - 1)] >> (4 * (state >> 8))) & 15u);
      if (ch != 0) L.plan[i] = (uint8_t)ch;
      // the predecessor: undo the shift, remove the cells the choice filled
      const int added = ch == 2 ? 2 : (ch == 3 ? below_bit : (ch == 4 ? (2 | below_bit) : 0));
      state = (((state << 1) ^ added) & mask) | (ch == 0 ? 1 : 0);
    }
  }
  __syncthreads();
  return total;
}

// search_tile_coefficient (layout_dp.rs:2033-2199) for tile slot (sx, sy)
// of the pass: 32 slots per DC group and axis.
__device__ __forceinline__ void retile_body(RtLds &L, const RetileArgs A, int sx, int sy) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int gx = sx >> 5, gy = sy >> 5;
  const int o = A.crossing ? 4 : 0;
  const int lx = o + 8 * (sx & 31), ly = o + 8 * (sy & 31);
  const int xsize = min(256, A.xb - gx * 256), ysize = min(256, A.yb - gy * 256);
  if (xsize <= 0 || ysize <= 0 || lx >= xsize || ly >= ysize) return;
  const int w = min(8, xsize - lx), h = min(8, ysize - ly), cells = w * h;
  if (cells < 2) return;
  const int X0 = gx * 256 + lx, Y0 = gy * 256 + ly;
  const size_t plane = (size_t)A.xb * A.yb;
  // Wave 0, one lane per cell: the tile's current pieces (first blocks; on a
  // shifted grid, those that fit inside the tile), the candidate costs and
  // the live masks.
  const bool in = wave == 0 && lane < cells;
  const int x = in ? lane % w : 0, y = in ? lane / w : 0;
  uint8_t cell = 0;
  if (in) cell = A.cells[(size_t)(Y0 + y) * A.xb + X0 + x];
  const int s = cell >> 1;
  const bool first = in && (cell & 1);
  const bool fits = first && x + cov_x(s) <= w && y + cov_y(s) <= h;
  const bool curp = first && (!A.crossing || fits);
  const bool fixedp = curp && rt_slot(s) < 0;
  float my_score = 0.0f;
  if (wave == 0) {
    bool go = true;
    if (!A.crossing) {
      // footprint_is_self_contained (src/ac_strategy/selection.rs:3235-3258)
      uint32_t area = first ? (uint32_t)(cov_x(s) * cov_y(s)) : 0u;
#pragma unroll
      for (int k = 32; k > 0; k >>= 1) area += (uint32_t)__shfl_xor((int)area, k);
      go = __ballot(first && !fits) == 0 && (int)area == cells;
    }
    go = go && __ballot(curp && (s == JXG_IDENTITY || s == JXG_DCT2X2)) == 0;
    const uint64_t covered = wave_or64(curp ? rt_mask(x, y, s, w) : 0ull);
    const uint64_t full = cells == 64 ? ~0ull : ((1ull << cells) - 1ull);
    const uint64_t blocked = A.crossing ? (full & ~covered) : 0ull;
    // candidate costs of the four shapes at each cell
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
    // the incumbent's scores
    if (curp) {
      const int sl = rt_slot(s);
      const size_t at = (size_t)(Y0 + y) * A.xb + X0 + x;
      my_score = sl >= 0 ? A.pc[(size_t)sl * plane + at] : A.c32[at];
    }
    const uint64_t fixed = wave_or64(fixedp ? rt_mask(x, y, s, w) : 0ull);
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
    if (lane == 0) {
      L.go = go;
      L.fixed = fixed;
      L.blocked = blocked;
      L.fixed_cost = fixed_cost;
    }
  }
  __syncthreads();
  if (!L.go) return;
  if (wave == 0) {
    // SmallCosts::live_masks (layout_dp.rs:1860-1916)
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
  const uint64_t fixed = L.fixed, blocked = L.blocked;
  const float fixed_cost = L.fixed_cost;
  // (rt_plan's first barrier publishes the live masks)
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
  if (!(winner < INFINITY) || wave != 0) return;  // winner is never NaN: sums of costs
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

__global__ __launch_bounds__(RT_T) void k_retile(RetileArgs A) {
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

__global__ __launch_bounds__(RT_T) void k_retile_b(const BatchImg *__restrict__ img,
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
  k_retile<<<g2, RT_T, 0, st>>>(R);
  LAUNCH_CHECK("k_retile");
  if (mode >= 2) {
    R.crossing = 1;
    k_retile<<<g2, RT_T, 0, st>>>(R);
    LAUNCH_CHECK("k_retile (shifted)");
  }
  if (ctx->h_derived->rerank) {
    const int rc = rerank_stages(ctx, st, xyb, stride, ph, W, H, qf, ytox, ytob, tables, p, cells);
    if (rc) return rc;
  }
  k_adjust_qf<<<(xb * yb + 255) / 256, 256, 0, st>>>(cells, qf, xb, yb, p->distance);
  LAUNCH_CHECK("k_adjust_qf");
  prof_end(ctx, JXG_PROF_RETILE, st);
  return JXG_OK;
}

// Batch: the same stages over every image (images whose tier does not
// re-tile have no workgroups).
JXG_HIDDEN int batch_retile(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                            const uint32_t *d_pfx, int n, const uint32_t total[], int fused) {
  const uint32_t *p = d_pfx;
  const int m = n + 1;
  if (!total[BS_RETILE]) return JXG_OK;
  if (fused)
    k_piece_costs_b<true><<<total[BS_PIECE], 256, 0, st>>>(d_img, p + BS_PIECE * m, n, ctx->d_scanpos);
  else
    k_piece_costs_b<false><<<total[BS_PIECE], 256, 0, st>>>(d_img, p + BS_PIECE * m, n, ctx->d_scanpos);
  k_retile_b<<<total[BS_RETILE], RT_T, 0, st>>>(d_img, p + BS_RETILE * m, n, 0);
  k_retile_b<<<total[BS_RETILE], RT_T, 0, st>>>(d_img, p + BS_RETILE * m, n, 1);
  const int rc = batch_rerank(ctx, st, d_img, d_pfx, n, total, fused);
  if (rc) return rc;
  k_adjust_qf_b<<<total[BS_ADJUST], 256, 0, st>>>(d_img, p + BS_ADJUST * m, n);
  LAUNCH_CHECK("batch retile");
  return JXG_OK;
}
