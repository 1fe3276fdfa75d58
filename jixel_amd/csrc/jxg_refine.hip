// jxg_refine.hip - refine_quant_field with qf_trellis, the last stage of the
// Fast tier's strategy decision (src/ac_strategy/selection.rs:534-896),
// run on the field after adjust_quant_field.
//
//   k_refine_list   per block: clears the block's refinement record and puts
//                   every eligible transform (a first block, not the DCT64
//                   family, field value above 1) on the list of its size
//                   class (up to 16x16, or 32x32).
//   k_refine_costs  one workgroup per listed transform (64 threads up to
//                   16x16, 256 for 32x32): the transform's three channels
//                   with the tile's CfL, then for the current field value
//                   and each coarser candidate the reconstruction cost of
//                   reconstruction_dist_and_rate (src/ac_strategy.rs:1538;
//                   src/inflated_cost.rs:1136-1500) in the reference's scalar
//                   order: the biased quantizer's coefficient error of Y, of
//                   X and B against the luma error (combine_error), the
//                   coefficient rate model, the scalar inverse transforms
//                   and the 8x8-window SSIM deficit per channel, then
//                   rd_cost with the reconstruction lambda.
//   k_qf_trellis    one workgroup per DC group: the raster chain of its first
//                   blocks, two rounds of QF residual prices (a histogram of
//                   the chain) and the Viterbi search over the candidates,
//                   then the chosen values written over each transform.
//
// The Viterbi recursion is a sequential f32 computation whose every
// rounding decides ties downstream, so one lane runs it, reading states
// that the other waves stage into LDS a chunk ahead.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "jxg_batch.h"
#include "jxg_ctx.h"
#include "jxg_device.h"
#include "jxg_internal.h"

using namespace jxg;

namespace {

struct RefineArgs {
  const float *xyb;
  int stride, ph, W, H, xb, yb;
  const uint8_t *cells;
  const uint8_t *qf;  // the adjusted field
  const int8_t *ytox, *ytob;
  const float *tables;
  const jxg_derived *D;
  const uint16_t *scanpos;  // storage index -> scan position per shape
  uint32_t *list[2];        // transforms by size class (block indices)
  uint32_t *count;          // [0], [1]: list lengths
  float *rcost;             // 4 per block
  uint8_t *rq;              // 4 per block
  uint8_t *rlen;            // 1 per block: candidate count (0: not refined)
};

// A wave-uniform 64-bit value in scalar registers.
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int rf_scanpos_offset(int cx, int cy) {
  const int a = cx * cy;
  return a == 1 ? 0 : (a == 2 ? 64 : (a == 4 ? (cx == 2 ? 192 : 448) : (a == 16 ? 1216 : 0)));
}

__device__ __forceinline__ void refine_list_body(const RefineArgs &A, int i) {
  int cls = -1;
  if (i < A.xb * A.yb) {
    const uint8_t cell = A.cells[i];
    if (cell & 1) {
      A.rlen[i] = 0;
      const int s = cell >> 1;
      const bool small = s == JXG_DCT || s == JXG_DCT16X8 || s == JXG_DCT8X16 || s == JXG_DCT16X16;
      if ((small || s == JXG_DCT32X32) && A.qf[i] > 1) cls = small ? 0 : 1;
    }
  }
  // one atomic per wave and class (the list order is free: each entry
  // writes its own block's record)
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const uint64_t m = __ballot(cls == c);
    if (!m) continue;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t first = 0;
    if (lane == leader) first = atomicAdd(&A.count[c], (uint32_t)__popcll(m));
    first = __shfl(first, leader);
    if (cls == c) A.list[c][first + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)i;
  }
}

__global__ __launch_bounds__(256) void k_refine_list(RefineArgs A) {
  refine_list_body(A, blockIdx.x * 256 + threadIdx.x);
}

// Buffers of one transform (NMAX coefficients, in-place transform pitch W + 1).
template <int NMAX, int NPITCH>
struct RfLds {
  // source pixels, row-major (pitch W); the 32x32 kernels read the planes
  // instead (L2-resident), which leaves LDS for four workgroups per CU
  float pix[3][NMAX == 1024 ? 1 : NMAX];
  float co[3][NPITCH];  // coefficients after CfL, in-place layout (pitch W + 1)
  float err[3][NMAX];   // quantizer errors, storage order
  float sp[3][NPITCH];  // per channel: the nonzero levels' rate terms, then the spatial error
  uint64_t nzm[3][NMAX / 64];
  uint32_t mx[3];       // per channel: largest scan position of a nonzero level + 1
  float win[3][16];     // per channel and window: SSIM deficit
  float quot[3][64];
  float rate[256];
  float chan_rate[3], chan_dist[3];
  double erased[4];     // per wavefront: the selection quantizer's erased energy
};

// The reconstruction-domain cost of the transform of strategy S whose first
// block is blk, at each of the len field values qs_list (the shape fixed at
// compile time, so that every index is a shift or a constant):
// reconstruction_dist_and_rate then rd_cost(Reconstruction, d, meta, ...)
// (src/ac_strategy.rs:1538,1691).  With selq the 32-family's quantizer is
// first scaled by the selection quantizer (reconstruction_strategy_cost_
// and_base, :774-860).  Every thread returns the costs.
template <bool F, bool BIG, int S>
__device__ __forceinline__ void recon_shape(RfLds<BIG ? 1024 : 256, BIG ? 1056 : 272> &L,
                                            const RefineArgs &A, uint32_t blk,
                                            const int *qs_list, int len, float meta, bool selq,
                                            float *costs) {
  constexpr int NMAX = BIG ? 1024 : 256, T = BIG ? 256 : 64;
  (void)NMAX;
  const int tid = threadIdx.x, lane = tid & 63;
  const jxg_derived *D = A.D;
  const int bx = (int)(blk % (uint32_t)A.xb), by = (int)(blk / (uint32_t)A.xb);
  constexpr int s = S;
  constexpr int cvx = (S == JXG_DCT8X16 || S == JXG_DCT16X16) ? 2 : (S == JXG_DCT32X32 ? 4 : 1);
  constexpr int cvy = (S == JXG_DCT16X8 || S == JXG_DCT16X16) ? 2 : (S == JXG_DCT32X32 ? 4 : 1);
  constexpr int W = cvx * 8, H = cvy * 8, P = W + 1, N = W * H;
  constexpr int cx = cvx > cvy ? cvx : cvy, cy = cvx > cvy ? cvy : cvx, SW = cx * 8, SH = cy * 8;
  constexpr bool rows_first = S == JXG_DCT || S == JXG_DCT8X16;
  constexpr bool trans = S != JXG_DCT8X16;  // storage k = c * H + r (else r * W + c)
  static_assert(N % T == 0, "whole coefficient rows per thread");
  const int kind = qt_kind(s);
  const uint32_t qoff = D->qt_off[kind], qn = D->qt_n[kind];
  const int cls = jxg_area_class(cx, cy);
  const int spoff = rf_scanpos_offset(cx, cy);
  const int px = bx * 8, py = by * 8;
  const size_t plane = (size_t)A.stride * A.ph;
  {
    const float num = D->dequant_numerator;
    for (int e = tid; e < 192; e += T) {
      const int c = e >> 6, a = e & 63;
      L.quot[c][a] = a >= 2 ? num / (float)a : (a == 1 ? 1.0f - D->dequant_one[c] : 0.0f);
    }
    for (int e = tid; e < 256; e += T) L.rate[e] = D->rate_lut[e];
  }
  // cmap_factors (src/ac_strategy.rs:1838-1855): the first block's 64x64
  // tile, clamped to its DC group's tile grid
  float fx, fb;
  {
    const int gx = bx >> 8, gy = by >> 8;
    const int xsize = min(256, A.xb - gx * 256), ysize = min(256, A.yb - gy * 256);
    const int lbx = bx - gx * 256, lby = by - gy * 256;
    const int tiles_x = (A.W + 63) / 64;
    const int tx = min(lbx / 8, (xsize + 7) / 8 - 1), ty = min(lby / 8, (ysize + 7) / 8 - 1);
    const size_t ti = (size_t)(gy * 32 + ty) * tiles_x + gx * 32 + tx;
    fx = fmla<F>((float)A.ytox[ti], D->cfl_inv_color_factor, D->cfl_base_x);
    fb = fmla<F>((float)A.ytob[ti], D->cfl_inv_color_factor, D->cfl_base_b);
  }
  // source pixels (the padded planes replicate the image edge, as
  // prepare_reconstruction and the forward gather do)
  for (int e = tid; e < 3 * N; e += T) {
    const int c = e / N, rem = e - c * N, r = rem / W, col = rem - r * W;
    const float v = A.xyb[c * plane + (size_t)(py + r) * A.stride + px + col];
    if constexpr (!BIG) L.pix[c][rem] = v;
    L.co[c][r * P + col] = v;
  }
  __syncthreads();
  // forward transform of each channel in place (the reference's pass order)
  {
    const float scale = 1.0f / (float)N;
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
      const bool rows = rows_first ? (pass == 0) : (pass == 1);
      const int nl = rows ? 3 * H : 3 * W;
      for (int e = tid; e < nl; e += T) {
        const int c = rows ? e / H : e / W, l = rows ? e - c * H : e - c * W;
        float *base = rows ? &L.co[c][l * P] : &L.co[c][l];
        const int step = rows ? 1 : P, len = rows ? W : H;
        if (len == 8)
          dct_line<F, 8>(base, step, scale, pass == 1);
        else if (len == 16)
          dct_line<F, 16>(base, step, scale, pass == 1);
        else
          dct_line<F, 32>(base, step, scale, pass == 1);
      }
      __syncthreads();
    }
  }
  auto coef_off = [&](int k) {
    const int r = trans ? k % H : k / W, c = trans ? k / H : k % W;
    return r * P + c;
  };
  // CfL with the tile's factors (apply_cfl, src/ac_strategy.rs:1104)
  for (int k = tid; k < N; k += T) {
    const int o = coef_off(k);
    const float y = L.co[1][o];
    L.co[0][o] = fmla<F>(-fx, y, L.co[0][o]);
    L.co[2][o] = fmla<F>(-fb, y, L.co[2][o]);
  }
  __syncthreads();
  const bool full = px + W <= A.W && py + H <= A.H;
  const int vw = min(W, A.W - px), vh = min(H, A.H - py);
  const int nwx = full ? W / 8 : (vw + 7) / 8, nwy = full ? H / 8 : (vh + 7) / 8;
  constexpr int KPT = NMAX / T;  // coefficients per thread and channel (at most)
  for (int j = 0; j < len; j++) {
    float qac = D->scale * (float)qs_list[j];
    if constexpr (S == JXG_DCT32X32) {
      if (selq && D->selection_quantizer) {
        // selection_quantizer::scale (src/ac_strategy/selection_quantizer.rs:
        // 82-135): the luma energy the coding quantizer erases outside the
        // lowest quarter of each axis, in f64 (per-thread terms in storage
        // order, then the wavefronts' tree sums in order)
        const float *inv_y = A.tables + qoff + 4 * qn;
        double energy = 0.0;
#pragma unroll
        for (int i = 0; i < KPT; i++) {
          const int k = tid + T * i;
          const int y = k >> 5, x = k & 31;
          if (x < 8 && y < 8) continue;
          const float l = L.co[1][coef_off(k)];
          if (fabsf(l * inv_y[k] * qac) < 0.5f) energy += (double)l * (double)l;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) energy += __shfl_xor(energy, off, 64);
        if (lane == 0) L.erased[tid >> 6] = energy;
        __syncthreads();
        const float erased = (float)(((L.erased[0] + L.erased[1]) + L.erased[2]) + L.erased[3]);
        const float dd = D->distance;
        const float t = fminf(fmaxf(erased / (2e-4f * dd * dd) - 1.0f, 0.0f), 1.0f);
        qac *= 1.0f + 0.6f * t;
      }
    }
    if (tid < 3) L.mx[tid] = 0u;
    // quantize Y, then X and B against the luma error
    uint32_t mxl[3] = {0u, 0u, 0u};
#pragma unroll
    for (int phase = 0; phase < 2; phase++) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        if ((phase == 0) != (c == 1)) continue;
        const float *inv = A.tables + qoff + (3 + c) * qn;
        const float *mat = A.tables + qoff + c * qn;
        const float qs = qac * D->qm_mult[c];
        const float inv_scale = 1.0f / qs;
        const float num = D->dequant_numerator;
        const float t0 = D->thr[c][cls][0], t1 = D->thr[c][cls][1];
        const float t2 = D->thr[c][cls][2], t3 = D->thr[c][cls][3];
        const float factor = c == 0 ? fx : fb;
#pragma unroll
        for (int i = 0; i < KPT; i++) {
          const int k = tid + T * i;
          if (T * i >= N) continue;  // uniform: N is a multiple of T
          const int srow = k / SW, scol = k - srow * SW;
          float e = 0.0f, r = 0.0f;
          bool nz = false;
          if (!(scol < cx && srow < cy)) {
            const int o = coef_off(k);
            const float coef = c == 1 ? L.co[1][o] : fmla<F>(factor, L.err[1][k], L.co[c][o]);
            const int yfix = srow >= SH / 2 ? 2 : 0;
            const float thr = scol >= SW / 2 ? (yfix ? t3 : t1) : (yfix ? t2 : t0);
            const float scaled = (inv[k] * qs) * coef;
            const float quantized = fabsf(scaled) >= thr ? round_away(scaled) : 0.0f;
            const float aq = fabsf(quantized);
            const float lev =
                aq < 64.0f ? quantized - copysignf(L.quot[c][(int)aq], quantized) : quantized - num / quantized;
            e = (scaled - lev) * (mat[k] * inv_scale);
            if (quantized != 0.0f) {
              nz = true;
              if (aq < 256.0f)
                r = L.rate[(int)aq];
              else {
                const uint32_t ki = aq >= 4294967295.0f ? 0xffffffffu : (uint32_t)aq;
                r = jxg_rate_log2(D, aq);
              }
            }
          }
          L.err[c][k] = e;
          L.sp[c][k] = r;  // the rate term, until the inverse transform
          if (nz) mxl[c] = max(mxl[c], (uint32_t)A.scanpos[spoff + k] + 1u);
          const uint64_t m = __ballot(nz);
          if (lane == 0) L.nzm[c][k >> 6] = m;
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
      if (mxl[c]) atomicMax(&L.mx[c], mxl[c]);
    __syncthreads();
    // the coefficient rate model: one wave per channel walks the nonzero
    // levels in storage order (the order of the sum), taking each term from
    // its lane
    for (int c = tid >> 6; c < 3; c += T / 64) {
      float mag = 0.0f;
      uint32_t nzc = 0;
      for (int ch = 0; ch < N / 64; ch++) {
        uint64_t m = uniform64(L.nzm[c][ch]);
        if (!m) continue;
        const int val = __float_as_int(L.sp[c][ch * 64 + lane]);
        nzc += (uint32_t)__popcll(m);
        while (m) {
          const int b = __builtin_ctzll(m);
          m &= m - 1ull;
          mag += __int_as_float(__builtin_amdgcn_readlane(val, b));
        }
      }
      if (lane == 0) {
        const uint32_t mx = L.mx[c];
        const float nzf = (float)nzc;
        const float header =
            0.4f * (nzc < 256u ? L.rate[nzc] : jxg_rate_log2(D, nzf));
        float visited = 0.0f;
        if (nzc != 0) {
          const int64_t span = (int64_t)mx - (int64_t)(cx * cy + nzc);
          visited = span > 0 ? (float)span : 0.0f;
        }
        L.chan_rate[c] = nzf * 1.6f + 1.0f * mag + header + 0.5f * visited;
      }
    }
    __syncthreads();
    // reconstruct_error with the scalar inverse transforms (src/dct.rs:
    // 1594-1680), the three channels side by side: squares take the storage
    // transposed into place, scaled by N, columns then rows; DCT8X16
    // columns (8) then rows (16); DCT16X8 8-point rows of the transposed
    // storage, then 16-point columns.
    {
      const float nsc = (float)N;
      for (int e = tid; e < 3 * N; e += T) {
        const int c = e / N, k = e - c * N;
        const float v = L.err[c][k] * nsc;
        int r, col;
        if (s == JXG_DCT8X16) {
          r = k / 16;
          col = k % 16;
        } else if (s == JXG_DCT16X8) {
          // storage u * 16 + v -> (row v, column u)
          r = k % 16;
          col = k / 16;
        } else {
          r = k % W;  // out[b * n + a] = coef[a * n + b]: a = k / n, b = k % n
          col = k / W;
        }
        L.sp[c][r * P + col] = v;
      }
      __syncthreads();
      if (s == JXG_DCT16X8) {
        // rows (8 points) first, then columns (16 points)
        for (int e = tid; e < 3 * H; e += T) {
          float *sp = L.sp[e / H];
          const int row = e % H;
          float b[8];
#pragma unroll
          for (int i = 0; i < 8; i++) b[i] = sp[row * P + i];
          inv_dct8(b);
#pragma unroll
          for (int i = 0; i < 8; i++) sp[row * P + i] = b[i];
        }
        __syncthreads();
        for (int e = tid; e < 3 * W; e += T) {
          float *sp = L.sp[e / W];
          const int col = e % W;
          float b[16];
#pragma unroll
          for (int i = 0; i < 16; i++) b[i] = sp[i * P + col];
          inv_dct16(b);
#pragma unroll
          for (int i = 0; i < 16; i++) sp[i * P + col] = b[i];
        }
      } else {
        // columns first, then rows
        for (int e = tid; e < 3 * W; e += T) {
          float *sp = L.sp[e / W];
          const int col = e % W;
          if (H == 8) {
            float b[8];
#pragma unroll
            for (int i = 0; i < 8; i++) b[i] = sp[i * P + col];
            inv_dct8(b);
#pragma unroll
            for (int i = 0; i < 8; i++) sp[i * P + col] = b[i];
          } else if (H == 16) {
            float b[16];
#pragma unroll
            for (int i = 0; i < 16; i++) b[i] = sp[i * P + col];
            inv_dct16(b);
#pragma unroll
            for (int i = 0; i < 16; i++) sp[i * P + col] = b[i];
          } else {
            float b[32];
#pragma unroll
            for (int i = 0; i < 32; i++) b[i] = sp[i * P + col];
            inv_dct32(b);
#pragma unroll
            for (int i = 0; i < 32; i++) sp[i * P + col] = b[i];
          }
        }
        __syncthreads();
        for (int e = tid; e < 3 * H; e += T) {
          float *sp = L.sp[e / H];
          const int row = e % H;
          if (W == 8) {
            float b[8];
#pragma unroll
            for (int i = 0; i < 8; i++) b[i] = sp[row * P + i];
            inv_dct8(b);
#pragma unroll
            for (int i = 0; i < 8; i++) sp[row * P + i] = b[i];
          } else if (W == 16) {
            float b[16];
#pragma unroll
            for (int i = 0; i < 16; i++) b[i] = sp[row * P + i];
            inv_dct16(b);
#pragma unroll
            for (int i = 0; i < 16; i++) sp[row * P + i] = b[i];
          } else {
            float b[32];
#pragma unroll
            for (int i = 0; i < 32; i++) b[i] = sp[row * P + i];
            inv_dct32(b);
#pragma unroll
            for (int i = 0; i < 32; i++) sp[row * P + i] = b[i];
          }
        }
      }
      __syncthreads();
    }
    // SSIM deficit per 8x8 window and channel (ssim_deficit_scalar_kernel,
    // or the visible variant for footprints past the image edge)
    const int nw = nwx * nwy;
    if (tid < 3 * nw) {
      const int c = tid / nw, wi = tid - c * nw;
      const int wx = wi % nwx, wy = wi / nwx, x0 = wx * 8, y0 = wy * 8;
      const int ww = full ? 8 : min(8, vw - x0), wh = full ? 8 : min(8, vh - y0);
      // the source pixels: from the (L2-resident) planes for 32x32, else LDS
      const float *o = BIG ? A.xyb + c * plane + (size_t)py * A.stride + px : L.pix[c];
      const size_t ostride = BIG ? (size_t)A.stride : (size_t)W;
      const float *sp = L.sp[c];
      const float bo = o[(size_t)y0 * ostride + x0], br = bo - sp[y0 * P + x0];
      float so = 0.0f, sr = 0.0f;
      if (full) {
        // whole window: rows of 8 loaded together, summed in row-major order
#pragma unroll
        for (int y = 0; y < 8; y++) {
          float ov[8], rv[8];
#pragma unroll
          for (int x = 0; x < 8; x++) {
            ov[x] = o[(size_t)(y0 + y) * ostride + x0 + x];
            rv[x] = ov[x] - sp[(y0 + y) * P + x0 + x];
          }
#pragma unroll
          for (int x = 0; x < 8; x++) {
            so += ov[x] - bo;
            sr += rv[x] - br;
          }
        }
      } else {
        for (int y = y0; y < y0 + wh; y++)
          for (int x = x0; x < x0 + ww; x++) {
            const float ov = o[(size_t)y * ostride + x], rv = ov - sp[y * P + x];
            so += ov - bo;
            sr += rv - br;
          }
      }
      const float count = (float)(ww * wh);
      const float inv_count = full ? (1.0f / 64.0f) : 1.0f / count;
      const float mo = bo + so * inv_count, mr = br + sr * inv_count;
      float vo = 0.0f, vr = 0.0f, cv = 0.0f;
      if (full) {
#pragma unroll
        for (int y = 0; y < 8; y++) {
          float ov[8], rv[8];
#pragma unroll
          for (int x = 0; x < 8; x++) {
            ov[x] = o[(size_t)(y0 + y) * ostride + x0 + x];
            rv[x] = ov[x] - sp[(y0 + y) * P + x0 + x];
          }
#pragma unroll
          for (int x = 0; x < 8; x++) {
            const float co_ = ov[x] - mo, cr = rv[x] - mr;
            vo = fmla<F>(co_, co_, vo);
            vr = fmla<F>(cr, cr, vr);
            cv = fmla<F>(co_, cr, cv);
          }
        }
      } else {
        for (int y = y0; y < y0 + wh; y++)
          for (int x = x0; x < x0 + ww; x++) {
            const float ov = o[(size_t)y * ostride + x], rv = ov - sp[y * P + x];
            const float co_ = ov - mo, cr = rv - mr;
            vo = fmla<F>(co_, co_, vo);
            vr = fmla<F>(cr, cr, vr);
            cv = fmla<F>(co_, cr, cv);
          }
      }
      vo *= inv_count;
      vr *= inv_count;
      cv *= inv_count;
      const float lum = (2.0f * mo * mr + 1e-4f) / (mo * mo + mr * mr + 1e-4f);
      const float st = (2.0f * cv + 9e-4f) / (vo + vr + 9e-4f);
      L.win[c][wi] = (1.0f - lum * st) * (full ? 64.0f : count);
    }
    __syncthreads();
    if (tid < 3) {
      float deficit = 0.0f;
      for (int w = 0; w < nw; w++) deficit += L.win[tid][w];
      L.chan_dist[tid] = D->channel_weight[tid] * deficit;
    }
    __syncthreads();
    // recon rate Y, X, B; distortion X, Y, B; rd_cost (src/ac_strategy.rs:1688)
    float rate = 0.0f;
    rate += L.chan_rate[1];
    rate += L.chan_rate[0];
    rate += L.chan_rate[2];
    float distortion = 0.0f;
    distortion += L.chan_dist[0];
    distortion += L.chan_dist[1];
    distortion += L.chan_dist[2];
    costs[j] = fmla<F>(D->recon_lambda, rate + meta, distortion);
    __syncthreads();
  }
}

// One listed transform of the refinement: the current value, then one or
// two coarser ones (find_quant_refinements, src/ac_strategy/selection.rs:
// 556-690), priced without the metadata charge.
template <bool F, bool BIG, int S>
__device__ __forceinline__ void refine_shape(RfLds<BIG ? 1024 : 256, BIG ? 1056 : 272> &L,
                                             const RefineArgs &A, uint32_t blk) {
  const int q0 = A.qf[blk];
  int qs_list[4] = {q0, 0, 0, 0};
  int len = 1;
  {
    const int steps = A.D->refine_steps;
    const int cand[3] = {q0 - 1, steps == 2 ? q0 - 2 : q0, q0};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int c = cand[k] < 0 ? 0 : cand[k];
      bool dup = c == 0;
      for (int m = 0; m < len; m++) dup = dup || qs_list[m] == c;
      if (!dup) qs_list[len++] = c;
    }
  }
  float costs[4];
  recon_shape<F, BIG, S>(L, A, blk, qs_list, len, 0.0f, false, costs);
  if (threadIdx.x == 0) {
    for (int j = 0; j < 4; j++) {
      A.rq[(size_t)blk * 4 + j] = (uint8_t)qs_list[j < len ? j : 0];
      A.rcost[(size_t)blk * 4 + j] = costs[j < len ? j : 0];
    }
    A.rlen[blk] = (uint8_t)(len > 1 ? len : 0);
  }
}

template <bool F, bool BIG>
__device__ __forceinline__ void refine_costs_body(RfLds<BIG ? 1024 : 256, BIG ? 1056 : 272> &L,
                                                  const RefineArgs &A, uint32_t item) {
  const uint32_t blk = A.list[BIG ? 1 : 0][item];
  if constexpr (BIG) {
    refine_shape<F, true, JXG_DCT32X32>(L, A, blk);
  } else {
    switch (A.cells[blk] >> 1) {
      case JXG_DCT: refine_shape<F, false, JXG_DCT>(L, A, blk); break;
      case JXG_DCT16X8: refine_shape<F, false, JXG_DCT16X8>(L, A, blk); break;
      case JXG_DCT8X16: refine_shape<F, false, JXG_DCT8X16>(L, A, blk); break;
      default: refine_shape<F, false, JXG_DCT16X16>(L, A, blk); break;
    }
  }
}

// A fixed grid walks the list (its length is known on the device only).
template <bool F, bool BIG>
__global__ __launch_bounds__(BIG ? 256 : 64, BIG ? 4 : 1) void k_refine_costs(RefineArgs A) {
  __shared__ RfLds<BIG ? 1024 : 256, BIG ? 1056 : 272> L;
  const uint32_t n = A.count[BIG ? 1 : 0];
  for (uint32_t item = blockIdx.x; item < n; item += gridDim.x) {
    refine_costs_body<F, BIG>(L, A, item);
    __syncthreads();
  }
}

__global__ void k_zero_u32(uint32_t *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// The SSIM rerank at the Fast tier (d <= 1.709; rerank_large_transforms and
// find_rerank_downgrades, src/ac_strategy/selection.rs:2718-3232,3568-3611),
// between the re-tiling and adjust_quant_field, on the field before the
// adjustment.  Every merged transform (2 or more blocks) is priced in the
// reconstruction domain with the rerank's metadata charge, against its
// DCT8 tiling and, for a 32x32, mixtures with its four aligned 16x16
// children (best_reconstruction_plan, :2636-2716); where the cheaper of
// those beats the merge (times MergeTuning::rerank_margin) the merge is
// split.  At Fast the scope has no rectangles and no fine mosaic, the saved
// child layouts hold no 1x1 leaf but DCT8, no learned prices exist and the
// upgrade passes need d >= 2, so these arms are all there is.
//
//   k_rerank_list    per block: each merge gets its record and its cost
//                    jobs (the merge, its DCT8 children in raster order, a
//                    32x32's 16x16 children), one atomic per wavefront.
//   k_rerank_costs   one workgroup per job, the refinement's cost kernel
//                    with the metadata charge and the selection quantizer.
//   k_rerank_decide  one lane per merge: the plan DP in LDS, the decision,
//                    the split written into the cells.

#define RR_META_R 2.6465739748323758f

struct RrJob {
  uint32_t blk;  // first block of the priced transform
  uint8_t s, q;  // its strategy and the footprint's aggregated field value
  uint16_t pad;
};

struct RrMerge {
  uint32_t blk;
  uint32_t small;  // first small job: [the merge unless 32x32], DCT8 children, 16x16 children
  uint32_t big;    // the 32x32's own job in the big list
  uint32_t s;
};

struct RerankArgs {
  RefineArgs A;  // the planes, maps, tables and derived parameters (qf: before adjustment)
  uint8_t *cells;
  RrMerge *merges;
  RrJob *jobs[2];  // small (<= 16x16), big (32x32)
  float *cost[2];
  uint32_t *count;  // [0] merges, [1] small jobs, [2] big jobs
};

__device__ __forceinline__ uint8_t rr_region_quant(const RerankArgs &R, int bx, int by, int w,
                                                   int h) {
  uint8_t max_q = 1;
  uint32_t sum = 0;
  for (int y = by; y < by + h; y++)
    for (int x = bx; x < bx + w; x++) {
      const uint8_t q = R.A.qf[(size_t)y * R.A.xb + x];
      max_q = q > max_q ? q : max_q;
      sum += q;
    }
  return jxg_aggregate_quant(max_q, sum, w * h, R.A.D->distance);
}

// Exclusive wavefront prefix of v and the wavefront's total.
__device__ __forceinline__ uint32_t wave_excl(uint32_t v, uint32_t &total) {
  const int lane = threadIdx.x & 63;
  uint32_t incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = (uint32_t)__shfl_up((int)incl, d, 64);
    if (lane >= d) incl += o;
  }
  total = (uint32_t)__shfl((int)incl, 63, 64);
  return incl - v;
}

__device__ __forceinline__ void rerank_list_body(const RerankArgs &R, int i) {
  const int xb = R.A.xb, yb = R.A.yb;
  int s = -1, cxs = 1, cys = 1;
  uint32_t nm = 0, ns = 0, nbg = 0;
  if (i < xb * yb) {
    const uint8_t cell = R.cells[i];
    if (cell & 1) {
      s = cell >> 1;
      cxs = cov_x(s);
      cys = cov_y(s);
      // the merged shapes of the Fast layout (the rectangles, DCT64 family
      // and the rest come only from slower tiers)
      const bool fast_merge = s == JXG_DCT16X8 || s == JXG_DCT8X16 || s == JXG_DCT16X16 ||
                              s == JXG_DCT32X32;
      if (fast_merge) {
        nm = 1;
        ns = s == JXG_DCT32X32 ? 20u : 1u + (uint32_t)(cxs * cys);
        nbg = s == JXG_DCT32X32 ? 1u : 0u;
      }
    }
  }
  uint32_t tm, ts, tb;
  const uint32_t em = wave_excl(nm, tm), es = wave_excl(ns, ts), eb = wave_excl(nbg, tb);
  uint32_t bm = 0, bs = 0, bb = 0;
  if ((threadIdx.x & 63) == 63) {
    if (tm) bm = atomicAdd(&R.count[0], tm);
    if (ts) bs = atomicAdd(&R.count[1], ts);
    if (tb) bb = atomicAdd(&R.count[2], tb);
  }
  bm = (uint32_t)__shfl((int)bm, 63, 64);
  bs = (uint32_t)__shfl((int)bs, 63, 64);
  bb = (uint32_t)__shfl((int)bb, 63, 64);
  if (!nm) return;
  const int bx = i % xb, by = i / xb;
  RrMerge m;
  m.blk = (uint32_t)i;
  m.small = bs + es;
  m.big = s == JXG_DCT32X32 ? bb + eb : 0xffffffffu;
  m.s = (uint32_t)s;
  R.merges[bm + em] = m;
  uint32_t k = m.small;
  auto job = [&](int x, int y, int js, int w, int h) {
    RrJob j;
    j.blk = (uint32_t)(y * xb + x);
    j.s = (uint8_t)js;
    j.q = rr_region_quant(R, x, y, w, h);
    j.pad = 0;
    return j;
  };
  if (s == JXG_DCT32X32)
    R.jobs[1][m.big] = job(bx, by, s, 4, 4);
  else
    R.jobs[0][k++] = job(bx, by, s, cxs, cys);
  for (int iy = 0; iy < cys; iy++)
    for (int ix = 0; ix < cxs; ix++) R.jobs[0][k++] = job(bx + ix, by + iy, JXG_DCT, 1, 1);
  if (s == JXG_DCT32X32)
    for (int iy = 0; iy < 4; iy += 2)
      for (int ix = 0; ix < 4; ix += 2) R.jobs[0][k++] = job(bx + ix, by + iy, JXG_DCT16X16, 2, 2);
}

__global__ __launch_bounds__(256) void k_rerank_list(RerankArgs R) {
  rerank_list_body(R, (int)(blockIdx.x * 256 + threadIdx.x));
}

template <bool F, bool BIG>
__device__ __forceinline__ void rerank_cost_body(RfLds<BIG ? 1024 : 256, BIG ? 1056 : 272> &L,
                                                 const RerankArgs &R, uint32_t item) {
  const RrJob j = R.jobs[BIG ? 1 : 0][item];
  const int q = j.q;
  float c[1];
  if constexpr (BIG) {
    recon_shape<F, true, JXG_DCT32X32>(L, R.A, j.blk, &q, 1, RR_META_R, true, c);
  } else {
    switch (j.s) {
      case JXG_DCT: recon_shape<F, false, JXG_DCT>(L, R.A, j.blk, &q, 1, RR_META_R, true, c); break;
      case JXG_DCT16X8: recon_shape<F, false, JXG_DCT16X8>(L, R.A, j.blk, &q, 1, RR_META_R, true, c); break;
      case JXG_DCT8X16: recon_shape<F, false, JXG_DCT8X16>(L, R.A, j.blk, &q, 1, RR_META_R, true, c); break;
      default: recon_shape<F, false, JXG_DCT16X16>(L, R.A, j.blk, &q, 1, RR_META_R, true, c); break;
    }
  }
  if (threadIdx.x == 0) R.cost[BIG ? 1 : 0][item] = c[0];
}

template <bool F, bool BIG>
__global__ __launch_bounds__(BIG ? 256 : 64) void k_rerank_costs(RerankArgs R) {
  __shared__ RfLds<BIG ? 1024 : 256, BIG ? 1056 : 272> L;
  const uint32_t n = R.count[BIG ? 2 : 1];
  for (uint32_t item = blockIdx.x; item < n; item += gridDim.x) {
    rerank_cost_body<F, BIG>(L, R, item);
    __syncthreads();
  }
}

// best_reconstruction_plan's table for one footprint (at most 4x4 blocks):
// per (log2 width, log2 height, origin) the plan's cost and how it was
// formed (0 none, 1 a tile, 2 left + right halves, 3 top + bottom halves).
struct RrPlan {
  float cost[9 * 16];
  uint8_t kind[9 * 16];
  uint8_t tile[9 * 16];
};

__device__ __forceinline__ int rr_index(int lx, int ly, int x, int y) {
  return (ly * 3 + lx) * 16 + y * 4 + x;
}

// The decision for one merge, in one lane (its plan table in LDS).
template <bool F>
__device__ __forceinline__ void rerank_decide_one(RrPlan &P, const RerankArgs &R, uint32_t item) {
  const jxg_derived *D = R.A.D;
  const RrMerge m = R.merges[item];
  const int s = (int)m.s, cxb = cov_x(s), cyb = cov_y(s);
  const int xb = R.A.xb, bx = (int)(m.blk % (uint32_t)xb), by = (int)(m.blk / (uint32_t)xb);
  const float j_big = s == JXG_DCT32X32 ? R.cost[1][m.big] : R.cost[0][m.small];
  const uint32_t c8 = m.small + (s == JXG_DCT32X32 ? 0u : 1u);
  // the tiles: DCT8 children in raster order, then the 16x16 children
  int tx[20], ty[20], tw[20];
  float tc[20];
  int nt = 0;
  float j_dct8 = 0.0f;
  for (int iy = 0; iy < cyb; iy++)
    for (int ix = 0; ix < cxb; ix++) {
      const float c = R.cost[0][c8 + iy * cxb + ix];
      tx[nt] = ix;
      ty[nt] = iy;
      tw[nt] = 1;
      tc[nt++] = c;
      j_dct8 += c;
    }
  if (s == JXG_DCT32X32)
    for (int k = 0; k < 4; k++) {
      tx[nt] = 2 * (k & 1);
      ty[nt] = 2 * (k >> 1);
      tw[nt] = 2;
      tc[nt++] = R.cost[0][m.small + 16 + k];
    }
  const int lw = cxb == 4 ? 2 : cxb - 1, lh = cyb == 4 ? 2 : cyb - 1;
  for (int i = 0; i < 9 * 16; i++) {
    P.cost[i] = INFINITY;
    P.kind[i] = 0;
  }
  for (int i = 0; i < nt; i++) {
    const int l = tw[i] == 2 ? 1 : 0;
    const int slot = rr_index(l, l, tx[i], ty[i]);
    if (tc[i] < P.cost[slot]) {
      P.cost[slot] = tc[i];
      P.kind[slot] = 1;
      P.tile[slot] = (uint8_t)i;
    }
  }
  for (int ly = 0; ly <= lh; ly++) {
    const int h = 1 << ly;
    for (int lx = 0; lx <= lw; lx++) {
      const int w = 1 << lx;
      for (int y = 0; y < cyb; y += h)
        for (int x = 0; x < cxb; x += w) {
          float best = INFINITY;
          int kind = 0;
          if (lx > 0) {
            best = P.cost[rr_index(lx - 1, ly, x, y)] + P.cost[rr_index(lx - 1, ly, x + w / 2, y)];
            kind = 2;
          }
          if (ly > 0) {
            const float split =
                P.cost[rr_index(lx, ly - 1, x, y)] + P.cost[rr_index(lx, ly - 1, x, y + h / 2)];
            if (split < best) {
              best = split;
              kind = 3;
            }
          }
          const int slot = rr_index(lx, ly, x, y);
          if (!(P.cost[slot] < best)) {
            P.cost[slot] = best;
            P.kind[slot] = (uint8_t)kind;
          }
        }
    }
  }
  const float j_child = P.cost[rr_index(lw, lh, 0, 0)];
  const bool child_wins = j_child < j_dct8;
  const float j_alt = child_wins ? j_child : j_dct8;
  if (!(j_alt < j_big * D->rerank_margin)) return;
  // the split: the plan's tiles (expanded from the top entry), or DCT8s
  auto put = [&](int x, int y, int st) {
    const int cw = cov_x(st), ch = cov_y(st);
    for (int iy = 0; iy < ch; iy++)
      for (int ix = 0; ix < cw; ix++)
        R.cells[(size_t)(by + y + iy) * xb + bx + x + ix] = (uint8_t)((st << 1) | (ix == 0 && iy == 0 ? 1 : 0));
  };
  if (!child_wins) {
    for (int iy = 0; iy < cyb; iy++)
      for (int ix = 0; ix < cxb; ix++) put(ix, iy, JXG_DCT);
    return;
  }
  int stack[32];
  int sp = 0;
  stack[sp++] = rr_index(lw, lh, 0, 0);
  while (sp > 0) {
    const int e = stack[--sp];
    const int lv = e >> 4, ly = lv / 3, lx = lv - ly * 3, y = (e >> 2) & 3, x = e & 3;
    const int kind = P.kind[e];
    if (kind == 1) {
      const int t = P.tile[e];
      put(tx[t], ty[t], tw[t] == 2 ? JXG_DCT16X16 : JXG_DCT);
    } else if (kind == 2) {
      stack[sp++] = rr_index(lx - 1, ly, x, y);
      stack[sp++] = rr_index(lx - 1, ly, x + (1 << lx) / 2, y);
    } else if (kind == 3) {
      stack[sp++] = rr_index(lx, ly - 1, x, y);
      stack[sp++] = rr_index(lx, ly - 1, x, y + (1 << ly) / 2);
    }
  }
}

// Four waves per workgroup, lane 0 of each deciding one merge at a time.
template <bool F>
__global__ __launch_bounds__(256) void k_rerank_decide(RerankArgs R) {
  __shared__ RrPlan P[4];
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) != 0) return;
  const uint32_t n = R.count[0];
  for (uint32_t item = blockIdx.x * 4 + wave; item < n; item += gridDim.x * 4)
    rerank_decide_one<F>(P[wave], R, item);
}

// ---------------------------------------------------------------------------
// The QF trellis per DC group

#define TR_T 1024
#define TR_WARM 16  // guess steps before a segment (measured: <= 5 reach the exact vector)

struct TrellisArgs {
  const uint8_t *cells;
  uint8_t *qf;
  int xb, yb;
  const float *rcost;
  const uint8_t *rq, *rlen;
  uint32_t *chain_blk;  // scratch, xb * yb each: the chain's blocks,
  float4 *rec_c;        // per step: candidate costs,
  uint32_t *rec_q;      // candidate values (a byte each),
  uint8_t *rec_len;     // candidate count (1: the fixed value at cost 0)
  uint8_t *chain_q, *chain_fixed, *chain_back;
  const jxg_derived *D;
};

struct TrLds {
  float prices[4][512];
  uint32_t counts[4][512];
  uint32_t totals[4];
  uint32_t scan[TR_T];
  float sc[TR_T][4];    // per segment: start vector (survivor costs)
  float ec[TR_T][4];    // per segment: end vector
  uint8_t elen[TR_T];   // per segment: states at its last step
  uint8_t map[TR_T];    // per segment: last-step state -> state before it (2 bits each)
  uint8_t send[TR_T];   // per segment: state at its last step on the chosen path
  int redo;             // the last verification pass that found a mismatch
};

// STRATEGY_CODE_LUT (src/dc_group_data.rs:73-78)
__constant__ uint8_t STRATEGY_CODE_DEV[24] = {0,  6,  7,  4,  5,  3,  12, 13, 10, 11, 14, 15,
                                              16, 17, 18, 19, 20, 1,  2,  21, 22, 23, 8,  9};
__device__ __forceinline__ int strategy_code(int s) { return STRATEGY_CODE_DEV[s]; }

// qf_symbol (src/ac_strategy/selection.rs:725-733): the writer's context of
// the left value and the packed residual, clamped to the table.
__device__ __forceinline__ int qf_slot(int left, int qf) {
  const int c = left > 11 ? 0 : (left > 5 ? 1 : (left > 3 ? 2 : 3));
  const int64_t d = (int64_t)qf - left;
  const uint32_t pk = (uint32_t)((d << 1) ^ (d >> 63));
  return c * 512 + (int)(pk < 511u ? pk : 511u);
}

// The survivors after one chain step: their values, costs (less the floor)
// and count.
struct TrVec {
  int q[4];
  float c[4];
  int len;
};

// The candidates of one chain step.
struct TrRec {
  float4 c;
  uint32_t q;
  int len;
};

__device__ __forceinline__ TrRec tr_rec(const TrellisArgs &A, size_t base, int i) {
  TrRec r;
  r.c = A.rec_c[base + i];
  r.q = A.rec_q[base + i];
  r.len = A.rec_len[base + i];
  return r;
}

// The survivor values of step i with arbitrary costs (the costs are set by
// the caller).
__device__ __forceinline__ void tr_load(const TrellisArgs &A, size_t base, int i, TrVec &v) {
  const TrRec r = tr_rec(A, base, i);
#pragma unroll
  for (int k = 0; k < 4; k++) v.q[k] = (int)((r.q >> (8 * k)) & 255u);
  v.len = r.len;
}

// One step of the recursion (:750-786) on v; the back pointers (2 bits per
// state) are stored when `keep`.  Step 0 prices against first_left.
template <bool F>
__device__ __forceinline__ void tr_step(const TrLds &L, const TrellisArgs &A, size_t base, int i,
                                        int first_left, float lambda_bits, TrVec &v, bool keep,
                                        const TrRec &rec) {
  int cq[4];
  const float cc[4] = {rec.c.x, rec.c.y, rec.c.z, rec.c.w};
  const int len = rec.len;
#pragma unroll
  for (int k = 0; k < 4; k++) cq[k] = (int)((rec.q >> (8 * k)) & 255u);
  float cost[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  uint32_t bk = 0;
#pragma unroll
  for (int st = 0; st < 4; st++) {
    if (st >= len) continue;
    const int qf = cq[st] - 1;
    if (i == 0) {
      cost[st] = cc[st] + lambda_bits * (&L.prices[0][0])[qf_slot(first_left, qf)];
      continue;
    }
#pragma unroll
    for (int pp = 0; pp < 4; pp++) {
      if (pp >= v.len) continue;
      const float total = v.c[pp] + cc[st] + lambda_bits * (&L.prices[0][0])[qf_slot(v.q[pp] - 1, qf)];
      if (total < cost[st]) {
        cost[st] = total;
        bk = (bk & ~(3u << (2 * st))) | ((uint32_t)pp << (2 * st));
      }
    }
  }
  float fl = INFINITY;
#pragma unroll
  for (int st = 0; st < 4; st++)
    if (st < len) fl = fminf(fl, cost[st]);
#pragma unroll
  for (int st = 0; st < 4; st++) {
    v.q[st] = cq[st];
    v.c[st] = st < len ? cost[st] - fl : INFINITY;
  }
  v.len = len;
  if (keep) A.chain_back[base + i] = (uint8_t)bk;
}

// Steps [from, to) of the recursion, each step's record loaded one step
// ahead.
template <bool F>
__device__ __forceinline__ void tr_run(const TrLds &L, const TrellisArgs &A, size_t base, int from,
                                       int to, int first_left, float lambda_bits, TrVec &v,
                                       bool keep) {
  if (from >= to) return;
  TrRec r = tr_rec(A, base, from);
  for (int i = from; i < to; i++) {
    TrRec nx = r;
    if (i + 1 < to) nx = tr_rec(A, base, i + 1);
    tr_step<F>(L, A, base, i, first_left, lambda_bits, v, keep, r);
    r = nx;
  }
}

template <bool F>
__device__ __forceinline__ void qf_trellis_body(TrLds &L, const TrellisArgs &A, int gx, int gy) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int x0 = gx * 256, y0 = gy * 256;
  const int xs = min(256, A.xb - x0), ys = min(256, A.yb - y0);
  if (xs <= 0 || ys <= 0) return;
  const jxg_derived *D = A.D;
  const size_t base = (size_t)y0 * A.xb + (size_t)x0 * ys;  // this group's scratch
  // the raster chain of first blocks: each thread counts a quarter row
  // (64 cells), a workgroup scan places them in raster order
  const int parts = ys * 4;
  int cnt = 0;
  size_t cell0 = 0;
  int cw = 0;
  if (tid < parts) {
    const int r = tid >> 2, q = tid & 3;
    cw = max(0, min(64, xs - q * 64));
    cell0 = (size_t)(y0 + r) * A.xb + x0 + q * 64;
    for (int x = 0; x < cw; x++) cnt += A.cells[cell0 + x] & 1;
  }
  int incl = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d);
    if (lane >= d) incl += o;
  }
  if (lane == 63) L.scan[wave] = (uint32_t)incl;
  __syncthreads();
  int before = incl - cnt;
  int n = 0;
  for (int w = 0; w < TR_T / 64; w++) {
    const int t = (int)L.scan[w];
    if (w < wave) before += t;
    n += t;
  }
  if (cnt) {
    int o = before;
    for (int x = 0; x < cw; x++) {
      const size_t c = cell0 + x;
      if (A.cells[c] & 1) {
        A.chain_blk[base + o] = (uint32_t)c;
        A.chain_fixed[base + o] = A.qf[c];
        o++;
      }
    }
  }
  if (n == 0) return;
  __syncthreads();
  // per chain step: its candidates (a refined transform's states, else its
  // fixed value at cost 0) in chain order; independent_choice (:695-704)
  for (int i = tid; i < n; i += TR_T) {
    const uint32_t blk = A.chain_blk[base + i];
    const int ln = A.rlen[blk];
    const uint8_t fq = A.chain_fixed[base + i];
    uint8_t q = fq;
    if (ln > 1) {
      const uint32_t qq = *(const uint32_t *)(A.rq + (size_t)blk * 4);
      const float4 c4 = *(const float4 *)(A.rcost + (size_t)blk * 4);
      A.rec_c[base + i] = c4;
      A.rec_q[base + i] = qq;
      A.rec_len[base + i] = (uint8_t)ln;
      const float cc[4] = {c4.x, c4.y, c4.z, c4.w};
      int bi = 0;
      float bc = cc[0];
      for (int k = 1; k < ln; k++) {
        if (cc[k] < bc) {
          bc = cc[k];
          bi = k;
        }
      }
      q = (uint8_t)((qq >> (8 * bi)) & 255u);
    } else {
      A.rec_c[base + i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      A.rec_q[base + i] = (uint32_t)fq * 0x01010101u;
      A.rec_len[base + i] = 1;
    }
    A.chain_q[base + i] = q;
  }
  const int first_left = strategy_code(A.cells[(size_t)y0 * A.xb + x0] >> 1);
  const float lambda_bits = fmla<F>(D->recon_lambda, 1.0f + 0.0f, 0.0f);
  for (int round = 0; round < 2; round++) {
    // qf_residual_prices (:707-723): the histogram of the chain
    for (int e = tid; e < 2048; e += TR_T) (&L.counts[0][0])[e] = 0;
    if (tid < 4) L.totals[tid] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += TR_T) {
      const int left = i == 0 ? first_left : (int)A.chain_q[base + i - 1] - 1;
      const int slot = qf_slot(left, (int)A.chain_q[base + i] - 1);
      atomicAdd(&(&L.counts[0][0])[slot], 1u);
      atomicAdd(&L.totals[slot >> 9], 1u);
    }
    __syncthreads();
    for (int e = tid; e < 2048; e += TR_T) {
      const int c = e >> 9, v = e & 511;
      const float denom = (float)L.totals[c] + 2.0f;
      const float prior = 2.0f * fast_exp2<F>(-fminf((float)v + 1.0f, 100.0f), D->exp2_table);
      (&L.prices[0][0])[e] = -dirty_log2f<F>(((float)L.counts[c][v] + prior) / denom);
    }
    __syncthreads();
    // qf_trellis (:737-800), segmented: lane t owns steps [t*seg, t*seg+seg)
    const int seg = (n + TR_T - 1) / TR_T;
    const int nseg = (n + seg - 1) / seg;
    const int s0 = tid * seg, s1 = min(n, s0 + seg);
    const bool own = tid < nseg;
    TrVec v;
    if (tid == 0) L.redo = 0;
    if (own) {
      // the start vector: exact at step 0, else a guess from TR_WARM steps
      // of the recursion begun at zero costs
      const int w0 = max(0, s0 - TR_WARM);
      if (w0 == 0) {
        v.len = 0;
      } else {
        tr_load(A, base, w0 - 1, v);
        for (int k = 0; k < 4; k++) v.c[k] = 0.0f;
      }
      tr_run<F>(L, A, base, w0, s0, first_left, lambda_bits, v, false);
      for (int k = 0; k < 4; k++) L.sc[tid][k] = v.c[k];
      tr_run<F>(L, A, base, s0, s1, first_left, lambda_bits, v, true);
      for (int k = 0; k < 4; k++) L.ec[tid][k] = v.c[k];
      L.elen[tid] = (uint8_t)v.len;
    }
    // verify: a segment's result stands once its start vector equals its
    // predecessor's end vector bit for bit (step 0 is exact); a mismatch is
    // re-run from the predecessor's end, until none is left
    for (int pass = 1;; pass++) {
      __syncthreads();
      bool redo = false;
      float start[4];
      if (own && tid > 0) {
        const int ln = L.elen[tid - 1];
        for (int k = 0; k < 4; k++) {
          start[k] = L.ec[tid - 1][k];
          if (k < ln && __float_as_uint(start[k]) != __float_as_uint(L.sc[tid][k])) redo = true;
        }
      }
      if (redo) L.redo = pass;
      __syncthreads();
      if (L.redo != pass) break;
      if (redo) {
        tr_load(A, base, s0 - 1, v);
        for (int k = 0; k < 4; k++) {
          v.c[k] = start[k];
          L.sc[tid][k] = start[k];
        }
        tr_run<F>(L, A, base, s0, s1, first_left, lambda_bits, v, true);
        for (int k = 0; k < 4; k++) L.ec[tid][k] = v.c[k];
      }
    }
    // backtrack: per segment the map from its last step's state to the
    // state before its first step, composed from the end by one lane
    uint32_t map = 0;
    if (own) {
      int st[4] = {0, 1, 2, 3};
      for (int i = s1 - 1; i >= s0; i--) {
        const uint32_t bk = A.chain_back[base + i];
#pragma unroll
        for (int k = 0; k < 4; k++) st[k] = (int)((bk >> (2 * st[k])) & 3u);
      }
      map = (uint32_t)(st[0] | st[1] << 2 | st[2] << 4 | st[3] << 6);
      L.map[tid] = (uint8_t)map;
    }
    __syncthreads();
    // the state at each segment's last step: the suffix compositions
    // h_t = map_{t+1} o ... o map_last (a doubling scan of 2-bit state maps)
    // applied to the end state, the first minimum (every cost is >= +0
    // after the floor)
    if (own) L.send[tid] = tid + 1 < nseg ? L.map[tid + 1] : (uint8_t)0xE4;  // 0xE4: identity
    __syncthreads();
    for (int d = 1; d < nseg; d <<= 1) {
      uint32_t h = 0;
      const bool upd = own && tid + d < nseg;
      if (upd) {
        const uint32_t f = L.send[tid], g = L.send[tid + d];
#pragma unroll
        for (int k = 0; k < 4; k++) h |= ((f >> (2 * ((g >> (2 * k)) & 3u))) & 3u) << (2 * k);
      }
      __syncthreads();
      if (upd) L.send[tid] = (uint8_t)h;
      __syncthreads();
    }
    {
      const int last = nseg - 1, ln = L.elen[last];
      int sb = 0;
      for (int k = 1; k < ln; k++)
        if (L.ec[last][k] < L.ec[last][sb]) sb = k;
      if (own) L.send[tid] = (uint8_t)((L.send[tid] >> (2 * sb)) & 3u);
    }
    __syncthreads();
    if (own) {
      int sstate = L.send[tid];
      for (int i = s1 - 1; i >= s0; i--) {
        const uint8_t q = (uint8_t)((A.rec_q[base + i] >> (8 * sstate)) & 255u);
        A.chain_q[base + i] = q;
        sstate = (int)((A.chain_back[base + i] >> (2 * sstate)) & 3u);
      }
    }
    __syncthreads();
  }
  // the chosen values over each transform whose value changed
  for (int i = tid; i < n; i += TR_T) {
    const uint8_t q = A.chain_q[base + i], old = A.chain_fixed[base + i];
    if (q == old) continue;
    const uint32_t blk = A.chain_blk[base + i];
    const int s = A.cells[blk] >> 1, cw = cov_x(s), ch = cov_y(s);
    const int bx = (int)(blk % (uint32_t)A.xb), by = (int)(blk / (uint32_t)A.xb);
    for (int y = 0; y < ch; y++)
      for (int x = 0; x < cw; x++) A.qf[(size_t)(by + y) * A.xb + bx + x] = q;
  }
}

template <bool F>
__global__ __launch_bounds__(TR_T) void k_qf_trellis(TrellisArgs A) {
  __shared__ TrLds L;
  qf_trellis_body<F>(L, A, blockIdx.x, blockIdx.y);
}


struct RfPtrs {
  float *rcost;
  uint8_t *rq, *rlen;
  uint32_t *list0, *list1, *count, *chain_blk;
  float4 *rec_c;
  uint32_t *rec_q;
  uint8_t *chain_q, *chain_fixed, *chain_back, *rec_len;
};

__host__ __device__ inline uint8_t *align16(uint8_t *p) {
  return (uint8_t *)(((uintptr_t)p + 15) & ~(uintptr_t)15);
}

// The refinement scratch of nb blocks (jxg_refine_scratch_bytes): per block
// 4 candidate costs, values and a count; two transform lists and their
// counters; per chain step its block, candidates (costs, values, count),
// chosen and fixed value and back pointers.
__host__ __device__ inline RfPtrs rf_ptrs(uint8_t *base, size_t nb) {
  RfPtrs r;
  r.rcost = (float *)base;
  r.rq = base + 16 * nb;
  r.rlen = r.rq + 4 * nb;
  r.list0 = (uint32_t *)align16(r.rlen + nb);
  r.list1 = r.list0 + nb;
  r.count = (uint32_t *)align16((uint8_t *)(r.list1 + nb));
  r.chain_blk = r.count + 4;
  r.rec_c = (float4 *)align16((uint8_t *)(r.chain_blk + nb));
  r.rec_q = (uint32_t *)(r.rec_c + nb);
  r.chain_q = (uint8_t *)(r.rec_q + nb);
  r.chain_fixed = r.chain_q + nb;
  r.chain_back = r.chain_fixed + nb;
  r.rec_len = r.chain_back + nb;
  return r;
}

__host__ __device__ inline void rf_grids(size_t nb, unsigned *g0, unsigned *g1) {
  *g0 = (unsigned)(nb < 16384 ? nb : 16384);
  *g1 = (unsigned)(nb / 16 + 1 < 2048 ? nb / 16 + 1 : 2048);
}

inline RefineArgs refine_args(const float *xyb, int stride, int ph, int W, int H, const uint8_t *cells,
                       const uint8_t *qf, const int8_t *ytox, const int8_t *ytob,
                       const float *tables, const jxg_derived *D, const uint16_t *scanpos,
                       const RfPtrs &r) {
  RefineArgs A;
  A.xyb = xyb;
  A.stride = stride;
  A.ph = ph;
  A.W = W;
  A.H = H;
  A.xb = (W + 7) / 8;
  A.yb = (H + 7) / 8;
  A.cells = cells;
  A.qf = qf;
  A.ytox = ytox;
  A.ytob = ytob;
  A.tables = tables;
  A.D = D;
  A.scanpos = scanpos;
  A.list[0] = r.list0;
  A.list[1] = r.list1;
  A.count = r.count;
  A.rcost = r.rcost;
  A.rq = r.rq;
  A.rlen = r.rlen;
  return A;
}

__device__ inline RefineArgs refine_args_dev(const BatchImg &I, const uint16_t *scanpos) {
  const RfPtrs r = rf_ptrs(I.rf, (size_t)I.xb * I.yb);
  RefineArgs A;
  A.xyb = I.xyb;
  A.stride = I.stride;
  A.ph = I.ph;
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
  A.scanpos = scanpos;
  A.list[0] = r.list0;
  A.list[1] = r.list1;
  A.count = r.count;
  A.rcost = r.rcost;
  A.rq = r.rq;
  A.rlen = r.rlen;
  return A;
}

__host__ __device__ inline TrellisArgs trellis_args(const uint8_t *cells, uint8_t *qf, int xb, int yb, const jxg_derived *D,
                         const RfPtrs &r) {
  TrellisArgs T;
  T.cells = cells;
  T.qf = qf;
  T.xb = xb;
  T.yb = yb;
  T.rcost = r.rcost;
  T.rq = r.rq;
  T.rlen = r.rlen;
  T.chain_blk = r.chain_blk;
  T.rec_c = r.rec_c;
  T.rec_q = r.rec_q;
  T.rec_len = r.rec_len;
  T.chain_q = r.chain_q;
  T.chain_fixed = r.chain_fixed;
  T.chain_back = r.chain_back;
  T.D = D;
  return T;
}

// ---- batch forms: the same bodies behind the flat index map ----

__global__ void k_refine_zero_b(const BatchImg *__restrict__ img, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const BatchImg &I = img[i];
    if (!I.D->refine_steps) continue;
    uint32_t *c = rf_ptrs(I.rf, (size_t)I.xb * I.yb).count;
    c[0] = 0;
    c[1] = 0;
  }
}

__global__ __launch_bounds__(256) void k_refine_list_b(const BatchImg *__restrict__ img,
                                                       const uint32_t *__restrict__ pfx, int n) {
  const int bi = batch_image(pfx, n, blockIdx.x);
  const BatchImg &I = img[bi];
  if (!I.D->refine_steps) return;
  const RefineArgs A = refine_args_dev(I, nullptr);
  refine_list_body(A, (int)((blockIdx.x - pfx[bi]) * 256 + threadIdx.x));
}

template <bool F, bool BIG>
__global__ __launch_bounds__(BIG ? 256 : 64) void k_refine_costs_b(const BatchImg *__restrict__ img,
                                                                   const uint32_t *__restrict__ pfx,
                                                                   int n,
                                                                   const uint16_t *__restrict__ scanpos) {
  __shared__ RfLds<BIG ? 1024 : 256, BIG ? 1056 : 272> L;
  const int bi = batch_image(pfx, n, blockIdx.x);
  const BatchImg &I = img[bi];
  if (!I.D->refine_steps) return;
  const RefineArgs A = refine_args_dev(I, scanpos);
  const uint32_t first = blockIdx.x - pfx[bi], grid = pfx[bi + 1] - pfx[bi];
  const uint32_t cnt = A.count[BIG ? 1 : 0];
  for (uint32_t item = first; item < cnt; item += grid) {
    refine_costs_body<F, BIG>(L, A, item);
    __syncthreads();
  }
}

template <bool F>
__global__ __launch_bounds__(TR_T) void k_qf_trellis_b(const BatchImg *__restrict__ img,
                                                       const uint32_t *__restrict__ pfx, int n) {
  const int bi = batch_image(pfx, n, blockIdx.x);
  const BatchImg &I = img[bi];
  if (!I.D->refine_steps) return;
  const RfPtrs r = rf_ptrs(I.rf, (size_t)I.xb * I.yb);
  const TrellisArgs T = trellis_args(I.cells, I.qf, I.xb, I.yb, I.D, r);
  const uint32_t bl = blockIdx.x - pfx[bi];
  const int gw = (I.xb + 255) / 256;
  __shared__ TrLds L;
  qf_trellis_body<F>(L, T, (int)(bl % (uint32_t)gw), (int)(bl / (uint32_t)gw));
}

struct RrPtrs {
  RrMerge *merges;
  RrJob *jobs0, *jobs1;
  float *cost0, *cost1;
  uint32_t *count;
};

// The rerank's scratch inside the refinement scratch of nb blocks (the
// refinement runs after it): merges <= nb / 2, small jobs <= 1.5 nb (a pair
// has 3 jobs), big jobs <= nb / 16.
__host__ __device__ inline RrPtrs rr_ptrs(uint8_t *base, size_t nb) {
  RrPtrs r;
  r.merges = (RrMerge *)align16(base);
  r.jobs0 = (RrJob *)(r.merges + nb / 2 + 1);
  r.jobs1 = r.jobs0 + (3 * nb) / 2 + 1;
  r.cost0 = (float *)(r.jobs1 + nb / 16 + 1);
  r.cost1 = r.cost0 + (3 * nb) / 2 + 1;
  r.count = (uint32_t *)align16((uint8_t *)(r.cost1 + nb / 16 + 1));
  return r;
}

__host__ __device__ inline void rr_grids(size_t nb, unsigned *g_list, unsigned *g_small,
                                         unsigned *g_big, unsigned *g_decide) {
  *g_list = (unsigned)((nb + 255) / 256);
  const size_t ns = (3 * nb) / 2 + 1, nbg = nb / 16 + 1, nm = nb / 8 + 1;
  *g_small = (unsigned)(ns < 16384 ? ns : 16384);
  *g_big = (unsigned)(nbg < 2048 ? nbg : 2048);
  *g_decide = (unsigned)(nm < 4096 ? nm : 4096);
}

__host__ __device__ inline RerankArgs rerank_args(const RefineArgs &A, uint8_t *cells,
                                                  const RrPtrs &r) {
  RerankArgs R;
  R.A = A;
  R.cells = cells;
  R.merges = r.merges;
  R.jobs[0] = r.jobs0;
  R.jobs[1] = r.jobs1;
  R.cost[0] = r.cost0;
  R.cost[1] = r.cost1;
  R.count = r.count;
  return R;
}

__device__ inline RerankArgs rerank_args_dev(const BatchImg &I, const uint16_t *scanpos) {
  return rerank_args(refine_args_dev(I, scanpos), I.cells, rr_ptrs(I.rf, (size_t)I.xb * I.yb));
}

__global__ void k_rerank_zero_b(const BatchImg *__restrict__ img, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const BatchImg &I = img[i];
    if (!I.D->rerank) continue;
    uint32_t *c = rr_ptrs(I.rf, (size_t)I.xb * I.yb).count;
    c[0] = c[1] = c[2] = 0;
  }
}

__global__ __launch_bounds__(256) void k_rerank_list_b(const BatchImg *__restrict__ img,
                                                       const uint32_t *__restrict__ pfx, int n) {
  const int bi = batch_image(pfx, n, blockIdx.x);
  const BatchImg &I = img[bi];
  if (!I.D->rerank) return;
  rerank_list_body(rerank_args_dev(I, nullptr), (int)((blockIdx.x - pfx[bi]) * 256 + threadIdx.x));
}

template <bool F, bool BIG>
__global__ __launch_bounds__(BIG ? 256 : 64) void k_rerank_costs_b(const BatchImg *__restrict__ img,
                                                                   const uint32_t *__restrict__ pfx,
                                                                   int n,
                                                                   const uint16_t *__restrict__ scanpos) {
  __shared__ RfLds<BIG ? 1024 : 256, BIG ? 1056 : 272> L;
  const int bi = batch_image(pfx, n, blockIdx.x);
  const BatchImg &I = img[bi];
  if (!I.D->rerank) return;
  const RerankArgs R = rerank_args_dev(I, scanpos);
  const uint32_t first = blockIdx.x - pfx[bi], grid = pfx[bi + 1] - pfx[bi];
  const uint32_t cnt = R.count[BIG ? 2 : 1];
  for (uint32_t item = first; item < cnt; item += grid) {
    rerank_cost_body<F, BIG>(L, R, item);
    __syncthreads();
  }
}

template <bool F>
__global__ __launch_bounds__(256) void k_rerank_decide_b(const BatchImg *__restrict__ img,
                                                         const uint32_t *__restrict__ pfx, int n) {
  __shared__ RrPlan P[4];
  const int bi = batch_image(pfx, n, blockIdx.x);
  const BatchImg &I = img[bi];
  if (!I.D->rerank) return;
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) != 0) return;
  const RerankArgs R = rerank_args_dev(I, nullptr);
  const uint32_t first = blockIdx.x - pfx[bi], grid = pfx[bi + 1] - pfx[bi];
  const uint32_t cnt = R.count[0];
  for (uint32_t item = first * 4 + wave; item < cnt; item += grid * 4)
    rerank_decide_one<F>(P[wave], R, item);
}

}  // namespace

// Host side: the SSIM rerank (Fast, d <= 1.709) after the re-tiling and
// before adjust_quant_field, on st; qf is the field before the adjustment.
JXG_HIDDEN int rerank_stages(jxg_ctx *ctx, hipStream_t st, const float *xyb, uint32_t stride,
                             uint32_t ph, uint32_t W, uint32_t H, const uint8_t *qf,
                             const int8_t *ytox, const int8_t *ytob, const float *tables,
                             const jxg_frame_params *p, uint8_t *cells) {
  const int xb = (int)((W + 7) / 8), yb = (int)((H + 7) / 8);
  const size_t nb = (size_t)xb * yb;
  int rc;
  if ((rc = ensure(&ctx->d_rf, &ctx->rf_cap, jxg_refine_scratch_bytes(nb)))) return rc;
  const RrPtrs r = rr_ptrs(ctx->d_rf, nb);
  const RfPtrs rp = rf_ptrs(ctx->d_rf, nb);
  const RefineArgs A = refine_args(xyb, (int)stride, (int)ph, (int)W, (int)H, cells, qf, ytox, ytob,
                                   tables, ctx->d_derived, ctx->d_scanpos, rp);
  const RerankArgs R = rerank_args(A, cells, r);
  unsigned gl, gs, gb, gd;
  rr_grids(nb, &gl, &gs, &gb, &gd);
  prof_begin(ctx, JXG_PROF_RERANK, st);
  k_zero_u32<<<1, 64, 0, st>>>(r.count, 3);
  k_rerank_list<<<gl, 256, 0, st>>>(R);
  LAUNCH_CHECK("k_rerank_list");
  if (p->fused) {
    k_rerank_costs<true, false><<<gs, 64, 0, st>>>(R);
    k_rerank_costs<true, true><<<gb, 256, 0, st>>>(R);
    k_rerank_decide<true><<<gd, 256, 0, st>>>(R);
  } else {
    k_rerank_costs<false, false><<<gs, 64, 0, st>>>(R);
    k_rerank_costs<false, true><<<gb, 256, 0, st>>>(R);
    k_rerank_decide<false><<<gd, 256, 0, st>>>(R);
  }
  LAUNCH_CHECK("k_rerank");
  prof_end(ctx, JXG_PROF_RERANK, st);
  return JXG_OK;
}

// Batch: the same stages over every image (images whose tier does not
// rerank have no workgroups).
JXG_HIDDEN int batch_rerank(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                            const uint32_t *d_pfx, int n, const uint32_t total[], int fused) {
  const uint32_t *p = d_pfx;
  const int m = n + 1;
  if (!total[BS_RRDECIDE]) return JXG_OK;
  k_rerank_zero_b<<<1, 256, 0, st>>>(d_img, n);
  k_rerank_list_b<<<total[BS_RRLIST], 256, 0, st>>>(d_img, p + BS_RRLIST * m, n);
  if (fused) {
    k_rerank_costs_b<true, false><<<total[BS_RRCOST_S], 64, 0, st>>>(d_img, p + BS_RRCOST_S * m, n, ctx->d_scanpos);
    k_rerank_costs_b<true, true><<<total[BS_RRCOST_B], 256, 0, st>>>(d_img, p + BS_RRCOST_B * m, n, ctx->d_scanpos);
    k_rerank_decide_b<true><<<total[BS_RRDECIDE], 256, 0, st>>>(d_img, p + BS_RRDECIDE * m, n);
  } else {
    k_rerank_costs_b<false, false><<<total[BS_RRCOST_S], 64, 0, st>>>(d_img, p + BS_RRCOST_S * m, n, ctx->d_scanpos);
    k_rerank_costs_b<false, true><<<total[BS_RRCOST_B], 256, 0, st>>>(d_img, p + BS_RRCOST_B * m, n, ctx->d_scanpos);
    k_rerank_decide_b<false><<<total[BS_RRDECIDE], 256, 0, st>>>(d_img, p + BS_RRDECIDE * m, n);
  }
  LAUNCH_CHECK("batch rerank");
  return JXG_OK;
}

// Workgroups of the rerank stages of one image (the batch grids).
JXG_HIDDEN void rerank_grids(uint32_t W, uint32_t H, uint32_t *g_list, uint32_t *g_small,
                             uint32_t *g_big, uint32_t *g_decide) {
  const size_t nb = (size_t)((W + 7) / 8) * ((H + 7) / 8);
  unsigned a, b, c, d;
  rr_grids(nb, &a, &b, &c, &d);
  *g_list = a;
  *g_small = b;
  *g_big = c;
  *g_decide = d;
}

// Host side: the refinement after adjust_quant_field (single image), on st.
JXG_HIDDEN int refine_stages(jxg_ctx *ctx, hipStream_t st, const float *xyb, uint32_t stride,
                             uint32_t ph, uint32_t W, uint32_t H, uint8_t *qf, const int8_t *ytox,
                             const int8_t *ytob, const float *tables, const jxg_frame_params *p,
                             const uint8_t *cells) {
  const int xb = (int)((W + 7) / 8), yb = (int)((H + 7) / 8);
  const size_t nb = (size_t)xb * yb;
  int rc;
  if ((rc = ensure(&ctx->d_rf, &ctx->rf_cap, jxg_refine_scratch_bytes(nb)))) return rc;
  const RfPtrs r = rf_ptrs(ctx->d_rf, nb);
  const RefineArgs A = refine_args(xyb, (int)stride, (int)ph, (int)W, (int)H, cells, qf, ytox, ytob,
                                   tables, ctx->d_derived, ctx->d_scanpos, r);
  unsigned g0, g1;
  rf_grids(nb, &g0, &g1);
  prof_begin(ctx, JXG_PROF_REFINE, st);
  k_zero_u32<<<1, 64, 0, st>>>(r.count, 2);
  k_refine_list<<<(unsigned)((nb + 255) / 256), 256, 0, st>>>(A);
  LAUNCH_CHECK("k_refine_list");
  if (p->fused) {
    k_refine_costs<true, false><<<g0, 64, 0, st>>>(A);
    k_refine_costs<true, true><<<g1, 256, 0, st>>>(A);
  } else {
    k_refine_costs<false, false><<<g0, 64, 0, st>>>(A);
    k_refine_costs<false, true><<<g1, 256, 0, st>>>(A);
  }
  LAUNCH_CHECK("k_refine_costs");
  const TrellisArgs T = trellis_args(cells, qf, xb, yb, ctx->d_derived, r);
  const dim3 g((xb + 255) / 256, (yb + 255) / 256);
  if (p->fused)
    k_qf_trellis<true><<<g, TR_T, 0, st>>>(T);
  else
    k_qf_trellis<false><<<g, TR_T, 0, st>>>(T);
  LAUNCH_CHECK("k_qf_trellis");
  prof_end(ctx, JXG_PROF_REFINE, st);
  return JXG_OK;
}

// Batch: the same stages over every image (images whose tier does not
// refine have no workgroups).
JXG_HIDDEN int batch_refine(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                            const uint32_t *d_pfx, int n, const uint32_t total[], int fused) {
  const uint32_t *p = d_pfx;
  const int m = n + 1;
  if (!total[BS_TRELLIS]) return JXG_OK;
  k_refine_zero_b<<<1, 256, 0, st>>>(d_img, n);
  k_refine_list_b<<<total[BS_RLIST], 256, 0, st>>>(d_img, p + BS_RLIST * m, n);
  if (fused) {
    k_refine_costs_b<true, false><<<total[BS_RCOST_S], 64, 0, st>>>(d_img, p + BS_RCOST_S * m, n, ctx->d_scanpos);
    k_refine_costs_b<true, true><<<total[BS_RCOST_B], 256, 0, st>>>(d_img, p + BS_RCOST_B * m, n, ctx->d_scanpos);
    k_qf_trellis_b<true><<<total[BS_TRELLIS], TR_T, 0, st>>>(d_img, p + BS_TRELLIS * m, n);
  } else {
    k_refine_costs_b<false, false><<<total[BS_RCOST_S], 64, 0, st>>>(d_img, p + BS_RCOST_S * m, n, ctx->d_scanpos);
    k_refine_costs_b<false, true><<<total[BS_RCOST_B], 256, 0, st>>>(d_img, p + BS_RCOST_B * m, n, ctx->d_scanpos);
    k_qf_trellis_b<false><<<total[BS_TRELLIS], TR_T, 0, st>>>(d_img, p + BS_TRELLIS * m, n);
  }
  LAUNCH_CHECK("batch refine");
  return JXG_OK;
}

// Workgroups of the refinement stages of one image (the batch grids).
JXG_HIDDEN void refine_grids(uint32_t W, uint32_t H, uint32_t *g_list, uint32_t *g0, uint32_t *g1,
                             uint32_t *g_trellis) {
  const size_t xb = (W + 7) / 8, yb = (H + 7) / 8, nb = xb * yb;
  unsigned a, b;
  rf_grids(nb, &a, &b);
  *g_list = (uint32_t)((nb + 255) / 256);
  *g0 = a;
  *g1 = b;
  *g_trellis = (uint32_t)(((xb + 255) / 256) * ((yb + 255) / 256));
}
