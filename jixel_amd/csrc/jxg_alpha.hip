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
//   k_alpha_stats     the plane's invisible count, minimum and maximum.
//   k_fill_bands      fill_invisible (src/invisible.rs:34-146).  Each invisible
//                     sample reads its already filled left and upper-right
//                     neighbours, so sample (x, y) can run once (x - 1, y) and
//                     (x + 1, y - 1) are done: the anti-diagonals x + 2y are the
//                     parallel steps.  One wavefront per 64-row band and
//                     channel, lane = row, the bands pipelined through a
//                     hand-over row in global memory (each band waits only on
//                     the band above).  Each sample is computed once with the
//                     reference's operations in its order, so the result is
//                     bit-identical.  Critical path: about W + 2 H steps.
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
// [5] set when the fill's wait for a band above ran out (never expected)
enum { ST_NEAR = 0, ST_CHANGED = 1, ST_MIN = 2, ST_MAX = 3, ST_ZERO = 4, ST_FAIL = 5, ST_WORDS = 8 };

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

// The plane's invisible count and alpha range, 16 samples per thread.
__global__ __launch_bounds__(256) void k_alpha_stats(const uint8_t *__restrict__ a, size_t n,
                                                     uint32_t *__restrict__ st) {
  uint32_t zeros = 0, lo = 255, hi = 0;
  // 16-byte words from the first aligned address; the bytes before it and
  // after the last whole word go to block 0
  const size_t head = min(n, (size_t)((16u - ((uintptr_t)a & 15u)) & 15u));
  const size_t n16 = (n - head) / 16, rest = n - head - 16 * n16;
  const uint4 *body = reinterpret_cast<const uint4 *>(a + head);
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (size_t)gridDim.x * 256) {
    const uint4 q = body[i];
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const uint32_t v = (w[k >> 2] >> (8 * (k & 3))) & 0xFFu;
      zeros += v == 0;
      lo = min(lo, v);
      hi = max(hi, v);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < head + rest) {
    const size_t at = threadIdx.x < head ? threadIdx.x : head + 16 * n16 + (threadIdx.x - head);
    const uint32_t v = a[at];
    zeros += v == 0;
    lo = min(lo, v);
    hi = max(hi, v);
  }
  zeros = wave_sum(zeros);
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, o));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, o));
  }
  // one set of atomics per block (the three words are shared by the whole
  // grid, so their atomics serialise: keep them few)
  __shared__ uint32_t part[3][4];
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[0][wv] = zeros;
    part[1][wv] = lo;
    part[2][wv] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    zeros = part[0][0] + part[0][1] + part[0][2] + part[0][3];
    lo = min(min(part[1][0], part[1][1]), min(part[1][2], part[1][3]));
    hi = max(max(part[2][0], part[2][1]), max(part[2][2], part[2][3]));
    if (zeros) atomicAdd(&st[ST_ZERO], zeros);
    atomicMin(&st[ST_MIN], lo);
    atomicMax(&st[ST_MAX], hi);
  }
}

constexpr uint32_t FB_EMPTY = 0xFFFFFFFFu;  // hand-over slot not written yet (a NaN the fill never makes)
constexpr uint32_t FB_SPIN_LIMIT = 1u << 24;  // bounded wait for the band above
constexpr int FB_T = 16;   // steps per chunk (and slots per LDS block)
constexpr int FB_R = 64;   // ring slots per row (four blocks)
constexpr int FB_M = 12;   // slots 0 .. 11 mirrored at 64 .. 75: a read of up to 13 slots never wraps
constexpr int FB_P = 77;   // row pitch of the sample and visibility rings (odd: rows spread over the banks)
constexpr int FB_OP = 65;  // row pitch of the output ring
constexpr int FB_G = 3;    // 16-byte words of a step's operand pack

typedef _Float16 fb_h2 __attribute__((ext_vector_type(2)));

// Word fill (a kernel, so that a captured stream holds kernel nodes only).
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    p[i] = v;
}

// One block of the band's skewed LDS rings: row r of the band keeps column x
// at slot x + 2 r, so the step t of row r reads slot t of its row (and nearby
// slots of the rows next to it).  A block holds slots s0 .. s0 + 15 of the 65
// sample rows (band rows and the row below) and of the 66 visibility rows
// (the row above, the band rows, the row below); lanes take 16 consecutive
// slots of four rows per load.  Loads go to registers first (fb_fetch) and
// to LDS two chunks later (fb_store).  Every load is unconditional, from
// clamped addresses; fb_store keeps a sample at a clamped position (never a
// neighbour of a stored step) and stores visibility as 2 (visible) or 0
// (invisible, or outside the image), the reference's neighbour weight.
struct FbBlock {
  float v[17];
  uint32_t a[17];
};

// The rows a lane of the fetching wave loads, set up once per band: lane
// (q, j) (q = lane / 16, j = lane % 16) loads sample row min(4 i + q, 64) and
// visibility row min(4 i + q, 65) - 1 at slot s0 + j for i = 0 .. 16.  The
// loads are buffer loads from the band's rows (at most 66 rows, so 32-bit
// offsets always reach): an offset past the rows reads 0.  A sample column
// is clamped to W - 1 (past the row lies padding, which may hold anything);
// a negative column reads the previous row's tail or 0.  Either value is
// finite and only ever met with weight 0.  An alpha column outside the row
// reads a neighbouring row or 0, and fb_store gives that position weight 0.
struct FbRows {
  __amdgpu_buffer_rsrc_t vbuf, abuf;
  uint32_t voff[17], vlim[17];  // sample byte offset at slot 0 (wrapping), and at column W - 1
  uint32_t aoff[17];            // alpha offset at slot 0 (wrapping)
  int y0, H;
  bool interior_rows;           // every visibility row is inside the image
};

// a pointer the compiler may keep in scalar registers (a buffer load's
// descriptor must be wave-uniform)
__device__ __forceinline__ void *fb_uniform(const void *p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v),
                 hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (void *)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ void fb_rows(FbRows &R, const float *plane, int stride,
                                        const uint8_t *alpha, int W, int H, int y0, int sub) {
  const int j = sub & 15, quad = sub >> 4;
  const int v_first = y0, v_last = min(y0 + 64, H - 1);
  const int a_first = max(y0 - 1, 0), a_last = min(y0 + 64, H - 1);
  R.vbuf = __builtin_amdgcn_make_buffer_rsrc(fb_uniform(plane + (size_t)v_first * stride), 0,
                                             __builtin_amdgcn_readfirstlane((v_last - v_first + 1) * stride * 4),
                                             0x00020000);
  R.abuf = __builtin_amdgcn_make_buffer_rsrc(fb_uniform(alpha + (size_t)a_first * W), 0,
                                             __builtin_amdgcn_readfirstlane((a_last - a_first + 1) * W),
                                             0x00020000);
  R.interior_rows = y0 >= 1 && y0 + 64 < H;
  R.y0 = y0;
  R.H = H;
#pragma unroll
  for (int i = 0; i < 17; i++) {
    const int r = min(4 * i + quad, 64), rv = min(4 * i + quad, 65) - 1;
    const uint32_t vrow = (uint32_t)((min(y0 + r, H - 1) - v_first) * stride) * 4u;
    R.voff[i] = vrow + 4u * (uint32_t)(j - 2 * r);
    R.vlim[i] = vrow + 4u * (uint32_t)(W - 1);
    R.aoff[i] = (uint32_t)((min(max(y0 + rv, 0), H - 1) - a_first) * W) + (uint32_t)(j - 2 * rv);
  }
}

__device__ __forceinline__ void fb_fetch(FbBlock &blk, const FbRows &R, int s0) {
#pragma unroll
  for (int i = 0; i < 17; i++) {
    const uint32_t vo = min(R.voff[i] + 4u * (uint32_t)s0, R.vlim[i]);
    blk.v[i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(R.vbuf, (int)vo, 0, 0));
    blk.a[i] = __builtin_amdgcn_raw_buffer_load_b8(R.abuf, (int)(R.aoff[i] + (uint32_t)s0), 0, 0);
  }
}

__device__ __forceinline__ void fb_store(const FbBlock &blk, const FbRows &R, float *tv, float *tw,
                                         int W, int s0, int sub) {
  const int slot = (s0 + (sub & 15)) & (FB_R - 1), quad = sub >> 4;
  float w[17];
  // inside the image for every lane and row: only alpha decides
  if (R.interior_rows && s0 >= 2 * 64 && s0 + 17 < W) {
#pragma unroll
    for (int i = 0; i < 17; i++) w[i] = blk.a[i] != 0 ? 2.0f : 0.0f;
  } else {
#pragma unroll
    for (int i = 0; i < 17; i++) {
      const int rv = min(4 * i + quad, 65) - 1, yv = R.y0 + rv;
      const uint32_t lim = (yv >= 0 && yv < R.H) ? (uint32_t)W : 0u;
      w[i] = ((uint32_t)(s0 + (sub & 15) - 2 * rv) < lim && blk.a[i] != 0) ? 2.0f : 0.0f;
    }
  }
#pragma unroll
  for (int i = 0; i < 17; i++) {
    const int r = min(4 * i + quad, 64), ra = min(4 * i + quad, 65);
    tv[r * FB_P + slot] = blk.v[i];
    tw[ra * FB_P + slot] = w[i];  // row ra - 1 at index ra
  }
  // the block at ring slots 0 .. 15 also goes to the mirror slots 64 .. 75
  if ((s0 & (FB_R - 1)) == 0 && (sub & 15) < FB_M) {
#pragma unroll
    for (int i = 0; i < 17; i++) {
      const int r = min(4 * i + quad, 64), ra = min(4 * i + quad, 65);
      tv[r * FB_P + slot + FB_R] = blk.v[i];
      tw[ra * FB_P + slot + FB_R] = w[i];
    }
  }
}

// fill_invisible as a pipeline of 64-row bands.  Workgroup 3 b + c fills
// band b of channel c; step t handles column t - 2 r of band row r, so the
// upper-right neighbour of a step is row r - 1's output of the step before
// and the upper neighbour the upper-right of the step before.  The eight
// waves of the workgroup (two per SIMD) split a chunk of 16 steps:
//
//   the chain wave (wave 0) runs the steps: lane r = row r, one DPP wave
//     shift per step for the upper-right value, and per step only the
//     operations that depend on the previous step, from a 48-byte operand
//     pack.  The other wave on its SIMD only keeps the barriers, so the
//     chain has the SIMD to itself;
//   four builder waves build the packs of the next chunk (4 steps each)
//     from the sample and visibility rings;
//   the store wave writes the outputs of the previous chunk back to the
//     plane, and the band's last row to the hand-over row;
//   the fetch wave fetches the rings' blocks (three chunks ahead) and loads
//     the hand-over row of the band above (loads only, so the compiler waits
//     for exactly the loads it needs).
//
// The roles follow the SIMDs the waves run on (read from the hardware at
// the start); the result does not depend on them.
//
// One barrier per chunk.  The hand-over row of band b - 1 (its last row,
// published with agent-scope stores a chunk after the step that made it) is
// read with agent-scope loads two chunks ahead and again while a slot is
// still empty; a band waits only on a lower workgroup index, which is
// dispatched before it.
//
// The sum follows the reference's order (invisible.rs:83-140):
//   sum = left c_L (c_L = 1 or 2 for a left neighbour, visible counting twice)
//       + up-right c_UR (1 with a row above and a right neighbour)
//       + R (2 right when visible) + up-right w_UR + DR (2 lower-right)
//       + up w_U + D (2 lower),
// one rounding per term as the reference's adds: the multipliers are 0, 1 or
// 2, so each product is exact, and a fused multiply-add (fp16 multiplier,
// exact) or an add of the exact pre-doubled neighbour rounds the sum once; an
// absent term adds a zero, which leaves the sum (never -0: its first
// operation adds +0) unchanged.  The quotient sum / weight (weight 2 .. 13;
// weight 0 or 1 leaves the sum) is the correctly rounded one: in FP64,
// q = s r, e = fma(-q, w, s), q' = fma(e, r, q) with r = 1 / w rounded, then
// rounded to float, equals IEEE float division for every finite float sum
// but -0 and every such weight (checked exhaustively); r = 0 for weight 0.
// The pack's mask then keeps the original sample where the step is visible
// or outside the image.
struct FbPack {
  uint32_t cl_cur, wur_wu;  // fp16 pairs: (c_L, c_UR), (w_UR, w_U)
  float R, DR, D, own;
  uint32_t keep, pad;       // keep: all ones where the original sample stays
  double r, w;              // 1 / weight (0 for weight 0), weight
};
static_assert(sizeof(FbPack) == 16 * FB_G, "pack layout");

__global__ __launch_bounds__(512) void k_fill_bands(float *__restrict__ xyb, int stride,
                                                    int padded_h, int W, int H,
                                                    const uint8_t *__restrict__ alpha,
                                                    uint32_t *__restrict__ xchg,
                                                    uint32_t *__restrict__ st) {
  if (st[ST_ZERO] == 0) return;  // every sample visible: nothing to fill
  __shared__ float tv[65 * FB_P];              // original samples
  __shared__ float tw[66 * FB_P];              // 2 * visibility, 0 outside the image
  __shared__ float tout[64 * FB_OP];           // outputs
  __shared__ uint4 pk[2 * FB_T * FB_G * 64];   // operand packs of two chunks
  __shared__ float4 uprow4[FB_R / 4];          // hand-over samples of the row above
  __shared__ double rtab[16];                  // 1 / w rounded, 0 for w = 0
  float *uprow = reinterpret_cast<float *>(uprow4);
  const int b = blockIdx.x / 3, c = blockIdx.x % 3;
  const int nbands = (H + 63) / 64;
  float *plane = xyb + (size_t)c * stride * padded_h;
  const uint32_t *xin = b > 0 ? xchg + ((size_t)c * nbands + (b - 1)) * W : nullptr;
  uint32_t *xout = xchg + ((size_t)c * nbands + b) * W;
  const bool publish = b + 1 < nbands;  // a band below reads this band's last row
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int y0 = b * 64;
  // steps t = -1 .. W + 126: chunk k covers t = k T - 1 .. k T + T - 2, whose
  // lane-0 upper-right columns k T .. k T + T - 1 are hand-over block k;
  // whole rounds of four chunks (the fetch wave's register blocks rotate by four)
  const int rounds = (W + 128 + 4 * FB_T - 1) / (4 * FB_T);

  // ---- wave 3 state: register blocks, hand-over loads
  uint32_t spins = 0;
  // (band 0 loads its own hand-over row, unused, so that every round issues
  // the same loads and the compiler waits for exactly the ones it needs)
  const uint32_t *xsrc = xin ? xin : xout;
  auto hand_load = [&](int k) __attribute__((always_inline)) -> uint32_t {
    const int col = min(k * FB_T + (lane & 15), W - 1);
    return __hip_atomic_load(xsrc + col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  auto hand_store = [&](int k, uint32_t v) __attribute__((always_inline)) {
    const int col = k * FB_T + (lane & 15);
    if (lane < 16) {
      if (xin) {
        while (v == FB_EMPTY && col < W && spins < FB_SPIN_LIMIT) {
          __builtin_amdgcn_s_sleep(2);
          v = __hip_atomic_load(xin + col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          spins++;
        }
      }
      // past the row (or with no row above) the slot is never a neighbour,
      // but its weight is 0 and 0 * NaN would not be: keep it finite
      uprow[col & (FB_R - 1)] = (xin && col < W) ? __uint_as_float(v) : 0.0f;
    }
  };
  // ---- the store wave: the outputs of chunk k (slots k T - 1 .. k T + T - 2)
  // of rows 4 i + q (lane = 16 q + j, i = 0 .. 15) to the plane, and the
  // last row to the hand-over row.  Visible samples come back unchanged (bit
  // for bit), so every sample inside the image is stored.
  float *wrow[16];
  uint32_t wlim[16];
  int wcol[16];
  auto wb_setup = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int r = 4 * i + (lane >> 4);
      wrow[i] = plane + (size_t)min(y0 + r, H - 1) * stride;
      wlim[i] = y0 + r < H ? (uint32_t)W : 0u;
      wcol[i] = (lane & 15) - 1 - 2 * r;  // column at chunk 0
    }
  };
  auto write_back = [&](int k) __attribute__((always_inline)) {
    const int slot = (k * FB_T - 1 + (lane & 15)) & (FB_R - 1), quad = lane >> 4;
    float v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = tout[(4 * i + quad) * FB_OP + slot];  // unconditional
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int col = k * FB_T + wcol[i];
      if ((uint32_t)col < wlim[i]) wrow[i][col] = v[i];
    }
    if (publish && quad == 3) {  // row 63
      const int col = k * FB_T + wcol[15];
      if ((uint32_t)col < wlim[15])
        __hip_atomic_store(xout + col, __float_as_uint(v[15]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
  };

  // ---- builder h (0 .. 3): the packs of chunk k (steps 4 h .. 4 h + 3 of it)
  const int row = lane, y = y0 + row;
  const bool has_up = y > 0, row_ok = y < H;
  const float cur_row = has_up ? 1.0f : 0.0f;
  auto build = [&](int k, int h) __attribute__((always_inline)) {
    const int tb = k * FB_T - 1 + 4 * h;
    uint4 *dst = pk + (size_t)((k & 1) * FB_T + 4 * h) * FB_G * 64 + lane;
    // windows: own row slots tb - 1 .. tb + 4, the row above tb - 2 .. tb + 2,
    // the row below tb + 2 .. tb + 6 (ring rows: tv r, r + 1; tw r, r + 1, r + 2)
    const float *vo = tv + row * FB_P + (tb & (FB_R - 1));
    const float *vd = tv + (row + 1) * FB_P + ((tb + 2) & (FB_R - 1));
    const float *wo = tw + (row + 1) * FB_P + ((tb - 1) & (FB_R - 1));
    const float *wu = tw + row * FB_P + ((tb - 2) & (FB_R - 1));
    const float *wd = tw + (row + 2) * FB_P + ((tb + 2) & (FB_R - 1));
    float o[5], d[5], vis[6], up[5], dn[5];
#pragma unroll
    for (int j = 0; j < 5; j++) {
      o[j] = vo[j];
      d[j] = vd[j];
      up[j] = wu[j];
      dn[j] = wd[j];
    }
#pragma unroll
    for (int j = 0; j < 6; j++) vis[j] = wo[j];
    // the weights first, then their reciprocals (table reads issued
    // together), then the packs
    float wt[4];
    uint32_t keep[4];
    auto item = [&](int i, bool interior) __attribute__((always_inline)) {
      // slot t - 1 + j of the own row is vis[i + j], slot t + j is o[i + j]
      float cl, cur;
      bool store;
      if (interior) {  // 0 < x < W - 1 for every row of the band
        cl = fmaf(vis[i], 0.5f, 1.0f);  // 1, or 2 when visible
        cur = cur_row;
        store = row_ok && vis[i + 1] == 0.0f;
      } else {
        const int x = tb + i - 2 * row;
        cl = x > 0 ? fmaf(vis[i], 0.5f, 1.0f) : 0.0f;
        cur = x + 1 < W ? cur_row : 0.0f;
        store = row_ok && (uint32_t)x < (uint32_t)W && vis[i + 1] == 0.0f;
      }
      wt[i] = ((cl + cur) + (vis[i + 2] + up[i + 1])) + ((dn[i + 1] + up[i]) + dn[i]);
      keep[i] = store ? 0u : ~0u;
      const uint32_t c0 = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(cl, cur));
      const uint32_t c1 = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(up[i + 1], up[i]));
      dst[(size_t)(i * FB_G) * 64] = make_uint4(c0, c1, __float_as_uint(o[i + 1] * vis[i + 2]),
                                                __float_as_uint(d[i + 1] * dn[i + 1]));
    };
    if (tb >= 2 * 63 + 1 && tb + 3 <= W - 2) {
#pragma unroll
      for (int i = 0; i < 4; i++) item(i, true);
    } else {
#pragma unroll
      for (int i = 0; i < 4; i++) item(i, false);
    }
    double rd[4];
#pragma unroll
    for (int i = 0; i < 4; i++) rd[i] = rtab[(int)wt[i]];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      dst[(size_t)(i * FB_G + 1) * 64] =
          make_uint4(__float_as_uint(d[i] * dn[i]), __float_as_uint(o[i]), keep[i], 0u);
      const double w = (double)wt[i];
      dst[(size_t)(i * FB_G + 2) * 64] =
          make_uint4((uint32_t)__double_as_longlong(rd[i]), (uint32_t)(__double_as_longlong(rd[i]) >> 32),
                     (uint32_t)__double_as_longlong(w), (uint32_t)(__double_as_longlong(w) >> 32));
    }
  };

  // ---- wave 0: the steps of chunk k
  float left = 0.0f, upv = 0.0f, prev = 0.0f;
  float *tor = tout + lane * FB_OP;
  auto run = [&](int k) __attribute__((always_inline)) {
    const uint4 *src = pk + (size_t)((k & 1) * FB_T) * FB_G * 64 + lane;
    // the chunk's hand-over samples first: the first step waits for them
    float4 u4[FB_T / 4];
#pragma unroll
    for (int m = 0; m < FB_T / 4; m++) u4[m] = uprow4[((k * FB_T + 4 * m) & (FB_R - 1)) >> 2];
    uint4 a[2][4][FB_G];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int g = 0; g < FB_G; g++) a[0][q][g] = src[(size_t)(q * FB_G + g) * 64];
#pragma unroll
    for (int m = 0; m < FB_T / 4; m++) {
      if (m + 1 < FB_T / 4) {
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
          for (int g = 0; g < FB_G; g++)
            a[(m + 1) & 1][q][g] = src[(size_t)((4 * (m + 1) + q) * FB_G + g) * 64];
      }
      // issue the next group's reads before this group's steps (a step has
      // too few instructions to hide an LDS read issued just before it)
      __builtin_amdgcn_sched_barrier(0);
      const float upr[4] = {u4[m].x, u4[m].y, u4[m].z, u4[m].w};
      float outq[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        FbPack p;
        __builtin_memcpy(&p, &a[m & 1][q][0], sizeof(p));
        const fb_h2 c2 = __builtin_bit_cast(fb_h2, p.cl_cur), w2 = __builtin_bit_cast(fb_h2, p.wur_wu);
        // (x + 1, y - 1), filled: row r - 1's output of the step before; lane 0
        // keeps the hand-over sample
        const float ur = __int_as_float(__builtin_amdgcn_update_dpp(
            __float_as_int(upr[q]), __float_as_int(prev), 0x138, 0xf, 0xf, false));
        float s = fmaf(left, (float)c2.x, 0.0f);
        s = fmaf(ur, (float)c2.y, s);
        s = s + p.R;
        s = fmaf(ur, (float)w2.x, s);
        s = s + p.DR;
        s = fmaf(upv, (float)w2.y, s);
        s = s + p.D;
        const double sd = (double)s, q0 = sd * p.r;
        const float qf = (float)fma(fma(-q0, p.w, sd), p.r, q0);
        const float v = __uint_as_float((__float_as_uint(p.own) & p.keep) |
                                        (__float_as_uint(qf) & ~p.keep));
        outq[q] = v;
        left = v;
        prev = v;
        upv = ur;
      }
#pragma unroll
      for (int q = 0; q < 4; q++) tor[(k * FB_T - 1 + 4 * m + q) & (FB_R - 1)] = outq[q];
    }
  };

  // Each role runs its own loop (a join of the roles' paths would copy the
  // fetch wave's registers in flight and wait for its loads); every wave
  // passes the same barriers: one for the roles, two in the prologue, one per
  // chunk.
  //
  // prologue: blocks -1 .. 2 in the rings, blocks 3 .. 5 in registers,
  // hand-over block 0 stored and blocks 1 and 2 in flight; the packs of
  // chunk 0.  Iteration k: the chain runs chunk k, the builders build chunk
  // k + 1, the store wave writes back chunk k - 1, the fetch wave stores
  // hand-over block k + 1 and ring block k + 3 and fetches ring block k + 6
  // and hand-over block k + 3.
  //
  // roles: the chain is wave 0, the first other wave on its SIMD keeps the
  // barriers only, the rest take the builder, fetch and store roles in order
  __shared__ int simd_of[8];
  if (lane == 0) simd_of[wave] = (__builtin_amdgcn_s_getreg((31 << 11) | 4) >> 4) & 3;  // HW_ID
  __syncthreads();
  enum { R_CHAIN, R_IDLE, R_BUILD, R_FETCH, R_STORE };
  int role = R_CHAIN, h = 0;
  if (wave != 0) {
    int partner = 0, rank = 0;
    for (int w = 1; w < 8; w++)
      if (!partner && simd_of[w] == simd_of[0]) partner = w;
    for (int w = 1; w < wave; w++) rank += w != partner;
    if (wave == partner) role = R_IDLE;
    else if (rank < 4) { role = R_BUILD; h = rank; }
    else if (rank == 4) role = R_FETCH;
    else if (rank == 5) role = R_STORE;
    else role = R_IDLE;
  }
  role = __builtin_amdgcn_readfirstlane(role);
  h = __builtin_amdgcn_readfirstlane(h);
  if (role == R_FETCH) {
    FbRows R;
    fb_rows(R, plane, stride, alpha, W, H, y0, lane);
    for (int b0 = -1; b0 <= 2; b0++) {
      FbBlock t;
      fb_fetch(t, R, b0 * FB_T);
      fb_store(t, R, tv, tw, W, b0 * FB_T, lane);
    }
    FbBlock blk[4];  // ring block j in blk[j % 4], fetched three chunks before it is stored
    uint32_t hv[4];  // hand-over block j in hv[j % 4], loaded two chunks before it is stored
    fb_fetch(blk[3], R, 3 * FB_T);
    fb_fetch(blk[0], R, 4 * FB_T);
    fb_fetch(blk[1], R, 5 * FB_T);
    hand_store(0, hand_load(0));
    hv[1] = hand_load(1);
    hv[2] = hand_load(2);
    hv[3] = hv[0] = 0u;
    __syncthreads();
    __syncthreads();
    // k = k4 + u with k4 a multiple of 4, so every register index is static
    auto io = [&](int k4, int u) __attribute__((always_inline)) {
      const int k = k4 + u;
      hand_store(k + 1, hv[(u + 1) & 3]);
      fb_store(blk[(u + 3) & 3], R, tv, tw, W, (k + 3) * FB_T, lane);
      fb_fetch(blk[(u + 2) & 3], R, (k + 6) * FB_T);
      hv[(u + 3) & 3] = hand_load(k + 3);
    };
    // the first round ahead of the loop: the loop is then entered with the
    // loads of a round in flight, as from its own end, so the compiler needs
    // no wait for them at its head
    auto round = [&](int k4) __attribute__((always_inline)) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        io(k4, u);
        __syncthreads();
      }
    };
    round(0);
    for (int k = 4; k < 4 * rounds; k += 4) round(k);
    if (spins >= FB_SPIN_LIMIT) atomicOr(&st[ST_FAIL], 1u);
  } else if (role == R_CHAIN) {
    __syncthreads();
    __syncthreads();
    for (int k = 0; k < 4 * rounds; k++) {
      run(k);
      __syncthreads();
    }
  } else if (role == R_BUILD) {
    if (h == 0 && lane < 16) rtab[lane] = lane ? 1.0 / (double)lane : 0.0;
    __syncthreads();
    build(0, h);
    __syncthreads();
    for (int k = 0; k < 4 * rounds; k++) {
      build(k + 1, h);
      __syncthreads();
    }
  } else if (role == R_STORE) {
    wb_setup();
    __syncthreads();
    __syncthreads();
    for (int k = 0; k < 4 * rounds; k++) {
      if (k > 0) write_back(k - 1);
      __syncthreads();
    }
    write_back(4 * rounds - 1);
  } else {
    __syncthreads();
    __syncthreads();
    for (int k = 0; k < 4 * rounds; k++) __syncthreads();
  }
}

// Edge padding of the three planes from the (filled) edge samples: the
// columns W .. stride - 1 of rows 0 .. H - 1, then every column of rows
// H .. padded_h - 1 (i = flat index over those two regions).
__global__ __launch_bounds__(256) void k_repad(float *__restrict__ xyb, int stride, int padded_h,
                                               int W, int H, const uint32_t *__restrict__ st) {
  if (st[ST_ZERO] == 0) return;
  const size_t side = (size_t)(stride - W) * H, bottom = (size_t)stride * (padded_h - H);
  float *plane = xyb + (size_t)blockIdx.y * stride * padded_h;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < side + bottom;
       i += (size_t)gridDim.x * 256) {
    int x, y;
    if (i < side) {
      y = (int)(i / (size_t)(stride - W));
      x = W + (int)(i % (size_t)(stride - W));
    } else {
      y = H + (int)((i - side) / (size_t)stride);
      x = (int)((i - side) % (size_t)stride);
    }
    plane[(size_t)y * stride + x] = plane[(size_t)min(y, H - 1) * stride + min(x, W - 1)];
  }
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

int alpha_stats(jxg_ctx *ctx, const uint8_t *a, uint32_t W, uint32_t H) {
  const size_t n = (size_t)W * H;
  const unsigned g = (unsigned)min((n / 16 + 255) / 256 + 1, (size_t)256);
  k_alpha_stats<<<g, 256, 0, ctx->stream>>>(a, n, ctx->d_astat);
  LAUNCH_CHECK("k_alpha_stats");
  return JXG_OK;
}

int fill_planes(jxg_ctx *ctx, float *xyb, uint32_t stride, uint32_t ph, uint32_t W, uint32_t H,
                const uint8_t *alpha) {
  const uint32_t nbands = (H + 63) / 64;
  const size_t slots = (size_t)3 * nbands * W;
  int rc;
  if ((rc = ensure(&ctx->d_fxchg, &ctx->fxchg_cap, slots))) return rc;
  const unsigned fg = (unsigned)min((slots + 1023) / 1024, (size_t)2048);
  k_fill_u32<<<fg, 256, 0, ctx->stream>>>(ctx->d_fxchg, slots, FB_EMPTY);
  LAUNCH_CHECK("k_fill_u32");
  prof_begin(ctx, JXG_PROF_FILL, ctx->stream);
  k_fill_bands<<<3 * nbands, 512, 0, ctx->stream>>>(xyb, (int)stride, (int)ph, (int)W, (int)H,
                                                    alpha, ctx->d_fxchg, ctx->d_astat);
  LAUNCH_CHECK("k_fill_bands");
  prof_end(ctx, JXG_PROF_FILL, ctx->stream);
  const size_t pad = (size_t)(stride - W) * H + (size_t)stride * (ph - H);
  if (pad) {
    const unsigned rg = (unsigned)min((pad + 255) / 256, (size_t)1024);
    k_repad<<<dim3(rg, 3), 256, 0, ctx->stream>>>(xyb, (int)stride, (int)ph, (int)W, (int)H,
                                                  ctx->d_astat);
    LAUNCH_CHECK("k_repad");
  }
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
  if ((rc = alpha_stats(ctx, ctx->d_alpha, W, H))) return rc;
  if ((rc = fill_planes(ctx, xyb, stride, ph, W, H, ctx->d_alpha))) return rc;
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
  if ((rc = alpha_stats(ctx, d_alpha, width, height))) return rc;
  return fill_planes(ctx, d_xyb, stride, padded_h, width, height, d_alpha);
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
