// jxg_epf.hip - the six-term EPF luma proxy (LumaProxy,
// src/epf_proxy.rs:151-245; box_filter.rs; row_kernels.rs).
//
// A proxy holds the source luma and, for its half- and quarter-resolution
// pyramids, the 5x5 mirrored box mean and variance.  compare() scores a
// candidate luma plane: the structure pool of both scales and the
// error / excess / fourth-power pools of the adjacent-pixel gradients.
//
// Every per-pixel value follows the reference's f32 operation order: the
// box sums are the running sums of original_box5 (box_filter.rs:466-488;
// the streaming kernels are pinned to it bit for bit).  A running sum is a
// serial recurrence along its line, so the passes are laid out for short
// serial steps and coalesced loads rather than for many threads:
//   * one wavefront keeps ONE running sum (the mean, the square or the
//     reference cross term) for 64 lines, a few instructions per step: the
//     line's terms sit in a register window, every sample is loaded once,
//     four chunks of 16 samples ahead of the recurrence;
//   * the vertical pass (lanes = columns of the row-major plane) writes its
//     sums x-major through an LDS transpose, so the horizontal pass (lanes =
//     rows) reads and writes 256-byte rows too;
//   * both scales go through each launch together;
//   * the structure pool is then an independent per-pixel map over the box
//     planes, fused with the gradient pools in one launch.
// The pooled sums are accumulated in f64 by fixed trees (wavefront
// shuffles, the workgroup's four wavefronts in order, then one workgroup
// over the partials in a fixed order): deterministic and more accurate than
// the reference's single f32 running sum.
#include <hip/hip_runtime.h>

#include "jxg_ctx.h"

namespace {

__device__ __forceinline__ int mirror(int i, int n) {
  if (i >= 0 && i < n) return i;
  int m = i % (2 * n);
  if (m < 0) m += 2 * n;
  return m < n ? m : 2 * n - m - 1;
}

// half / half_y (row_kernels.rs:60-70)
__global__ void k_epf_half(const float *__restrict__ src, int sstride, int w, int h, int bias,
                           float *__restrict__ dst) {
  const int ow = w / 2, oh = h / 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ow * oh) return;
  const int y = i / ow, x = i - y * ow;
  const float *r0 = src + (size_t)(2 * y) * sstride, *r1 = src + (size_t)(2 * y + 1) * sstride;
  float t00 = r0[2 * x], t10 = r1[2 * x], t01 = r0[2 * x + 1], t11 = r1[2 * x + 1];
  if (bias) {
    t00 += 0.01f;
    t10 += 0.01f;
    t01 += 0.01f;
    t11 += 0.01f;
  }
  dst[i] = 0.25f * (t00 + t10 + t01 + t11);
}

#define EPF_CH 16                // steps per chunk
#define EPF_NA 4                 // chunks of loads in flight ahead of the recurrence
#define EPF_WIN (EPF_CH + 5)     // window terms a chunk of steps reads
#define EPF_LINES 64             // lines per workgroup (exactly one wavefront)
#define EPF_TS (EPF_LINES + 1)   // LDS transpose tile stride (conflict-free both ways)

// One scale's share of a pass.  Vertical: test / ref are row-major w x h,
// out[q] x-major (element (x, y) at x * hp + y, the pitch hp being h rounded
// up to 64).  Horizontal: in[q] and out[q] x-major.  groups = workgroups
// (64 lines) per quantity.
struct PassJob {
  const float *test, *ref;
  const float *in[3];
  float *out[3];
  int w, h, hp, groups;
};

struct PassArgs {
  PassJob job[2];  // half and quarter scale
  int blocks0;     // workgroups of job 0 (nq * groups)
};

__device__ __forceinline__ float *pick(float *const p[3], int q) {
  return q == 0 ? p[0] : q == 1 ? p[1] : p[2];
}
__device__ __forceinline__ const float *pick(const float *const p[3], int q) {
  return q == 0 ? p[0] : q == 1 ? p[1] : p[2];
}

// The running sums of a line of n samples s: with P[k] = s[mirror(k - 2)],
// the window sum starts from -0.0 and adds P[0..4] in order; step y outputs
// sum * 0.2 and then (if y + 1 < n) adds (P[y + 5] - P[y]) - the order of
// box_filter.rs:118-150 and :466-488.  A line's terms live in a register
// window win[k] = P[c + k], k < EPF_WIN, for the chunk of steps starting at
// c; each sample is loaded once, EPF_NA chunks ahead of its step.
//
// Both passes run on workgroups of exactly one wavefront.  LDS operations
// of a wavefront are performed in order, so the LDS transpose only needs
// the compiler to keep program order, which the wavefront-scope fences give.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// --- vertical pass (box_filter.rs:118-150) --------------------------------
// Q = 0: sum of test; Q = 1: sum of test^2; Q = 2: sum of ref * test.  The
// term of a sample is rounded to f32 before it enters the sums, exactly as
// sum += a * a - b * b rounds it.
template <int Q>
struct VRaw {
  float t[EPF_CH];
  float r[Q == 2 ? EPF_CH : 1];
};

template <int Q>
__device__ __forceinline__ float v_term(float t, float r) {
  if constexpr (Q == 0)
    return t;
  else if constexpr (Q == 1)
    return t * t;
  else
    return r * t;
}

// Index of P[k] for the fetch of chunk c's new terms P[c + 5 .. c + 20]:
// the fetch starts at min(c, n - 1) + 5 (a chunk past the end fetches
// in-range samples it never uses), so the sample index i = k - 2 lies in
// [3, n + 17] and one reflection at the far end covers it when n >= 18.
// Shorter lines take the general mirror.  Fetches are issued
// unconditionally and without branches, so the loads of a chunk stay in
// flight while the steps of earlier chunks run.
#define EPF_SHORT 24

template <bool SHORT>
__device__ __forceinline__ int fetch_index(int i, int n) {
  if constexpr (SHORT)
    return mirror(i, n);
  else
    return i >= n ? 2 * n - i - 1 : i;
}

template <int Q, bool SHORT>
__device__ __forceinline__ void v_fetch(VRaw<Q> &d, const float *__restrict__ test,
                                        const float *__restrict__ ref, int w, int h, int x,
                                        int c) {
  const int r0 = min(c, h - 1) + 3;
#pragma unroll
  for (int j = 0; j < EPF_CH; j++) {
    const size_t i = (size_t)fetch_index<SHORT>(r0 + j, h) * w + x;
    d.t[j] = test[i];
    if constexpr (Q == 2) d.r[j] = ref[i];
  }
}

template <int Q, bool SHORT>
__device__ __forceinline__ void v_pass(const PassJob &J, int g, float *__restrict__ S) {
  const int lane = threadIdx.x, w = J.w, h = J.h, hp = J.hp, x0 = g * EPF_LINES;
  const int x = min(x0 + lane, w - 1);  // lanes past the edge repeat the last column
  const float *__restrict__ test = J.test;
  const float *__restrict__ ref = J.ref;
  float *__restrict__ out = J.out[Q];
  float win[EPF_WIN];
#pragma unroll
  for (int k = 0; k < EPF_WIN; k++) {
    const size_t i = (size_t)mirror(k - 2, h) * w + x;
    win[k] = v_term<Q>(test[i], Q == 2 ? ref[i] : 0.0f);
  }
  float acc = -0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) acc += win[k];
  // slot b holds the new terms of the chunks c with c / 16 = b (mod EPF_NA)
  VRaw<Q> raw[EPF_NA];
#pragma unroll
  for (int b = 1; b < EPF_NA; b++) v_fetch<Q, SHORT>(raw[b], test, ref, w, h, x, EPF_CH * b);
  for (int y0 = 0; y0 < h; y0 += EPF_CH * EPF_NA) {
#pragma unroll
    for (int b = 0; b < EPF_NA; b++) {
      const int c = y0 + EPF_CH * b;
      if (c >= h) break;
      v_fetch<Q, SHORT>(raw[b], test, ref, w, h, x, c + EPF_CH * EPF_NA);
      float *srow = S + lane * EPF_TS + (c & (EPF_LINES - 1));
      if (c + EPF_CH < h) {
#pragma unroll
        for (int j = 0; j < EPF_CH; j++) {
          srow[j] = acc * 0.2f;
          acc += win[j + 5] - win[j];
        }
      } else {
#pragma unroll
        for (int j = 0; j < EPF_CH; j++) {
          if (c + j >= h) break;
          srow[j] = acc * 0.2f;
          if (c + j + 1 < h) acc += win[j + 5] - win[j];
        }
      }
      if (((c + EPF_CH) & (EPF_LINES - 1)) == 0 || c + EPF_CH >= h) {
        // Rows yb .. yb + 63 go out x-major in 16 stores of 4 columns x 64
        // rows, 16 bytes a lane.  The pitch is a multiple of 64, so every
        // tile is whole and aligned (rows past h land in the pitch padding,
        // which nothing reads); columns past the edge store the last
        // column's identical values again.  No branch and no loop here: a
        // loop would make the compiler wait for every load in flight.
        const int yb = c & ~(EPF_LINES - 1);
        const int q4 = (lane & 15) * 4, cl = lane >> 4;
        wave_sync();
#pragma unroll
        for (int k = 0; k < EPF_LINES / 4; k++) {
          const int col = 4 * k + cl;
          const float *sp = S + col * EPF_TS + q4;
          const float4 v4 = make_float4(sp[0], sp[1], sp[2], sp[3]);
          *reinterpret_cast<float4 *>(out + (size_t)min(x0 + col, w - 1) * hp + yb + q4) = v4;
        }
        wave_sync();
      }
      const VRaw<Q> &nx = raw[(b + 1) % EPF_NA];
#pragma unroll
      for (int k = 0; k < 5; k++) win[k] = win[k + EPF_CH];
#pragma unroll
      for (int j = 0; j < EPF_CH; j++) win[5 + j] = v_term<Q>(nx.t[j], Q == 2 ? nx.r[j] : 0.0f);
    }
  }
}

template <int Q>
__device__ __forceinline__ void v_pass_any(const PassJob &J, int g, float *__restrict__ S) {
  if (J.h < EPF_SHORT)
    v_pass<Q, true>(J, g, S);
  else
    v_pass<Q, false>(J, g, S);
}

__global__ __launch_bounds__(EPF_LINES) void k_epf_vpass(PassArgs A) {
  __shared__ float S[EPF_LINES * EPF_TS];
  int b = blockIdx.x;
  const bool second = b >= A.blocks0;
  const PassJob J = second ? A.job[1] : A.job[0];
  if (second) b -= A.blocks0;
  const int q = b / J.groups, g = b - q * J.groups;
  if (q == 0)
    v_pass_any<0>(J, g, S);
  else if (q == 1)
    v_pass_any<1>(J, g, S);
  else
    v_pass_any<2>(J, g, S);
}

// --- horizontal pass: the same recurrence along x on the x-major sums ------
template <bool SHORT>
__device__ __forceinline__ void h_fetch(float (&d)[EPF_CH], const float *__restrict__ v, int w,
                                        int hp, int y, int c) {
  const int c0 = min(c, w - 1) + 3;
#pragma unroll
  for (int j = 0; j < EPF_CH; j++) d[j] = v[(size_t)fetch_index<SHORT>(c0 + j, w) * hp + y];
}

// Lanes past the last row run the last row again and store into the pitch
// padding (y < hp), which nothing reads, so no lane is masked.
template <bool SHORT>
__device__ __forceinline__ void h_pass(const float *__restrict__ v, float *__restrict__ out, int w,
                                       int h, int hp, int y) {
  const int yc = min(y, h - 1);
  float win[EPF_WIN];
#pragma unroll
  for (int k = 0; k < EPF_WIN; k++) win[k] = v[(size_t)mirror(k - 2, w) * hp + yc];
  float acc = -0.0f;
#pragma unroll
  for (int k = 0; k < 5; k++) acc += win[k];
  float raw[EPF_NA][EPF_CH];
#pragma unroll
  for (int s = 1; s < EPF_NA; s++) h_fetch<SHORT>(raw[s], v, w, hp, yc, EPF_CH * s);
  for (int x0 = 0; x0 < w; x0 += EPF_CH * EPF_NA) {
#pragma unroll
    for (int s = 0; s < EPF_NA; s++) {
      const int c = x0 + EPF_CH * s;
      if (c >= w) break;
      h_fetch<SHORT>(raw[s], v, w, hp, yc, c + EPF_CH * EPF_NA);
      float *o = out + (size_t)c * hp + y;
      if (c + EPF_CH < w) {
#pragma unroll
        for (int j = 0; j < EPF_CH; j++) {
          o[(size_t)j * hp] = acc * 0.2f;
          acc += win[j + 5] - win[j];
        }
      } else {
#pragma unroll
        for (int j = 0; j < EPF_CH; j++) {
          if (c + j >= w) break;
          o[(size_t)j * hp] = acc * 0.2f;
          if (c + j + 1 < w) acc += win[j + 5] - win[j];
        }
      }
      const float *nx = raw[(s + 1) % EPF_NA];
#pragma unroll
      for (int k = 0; k < 5; k++) win[k] = win[k + EPF_CH];
#pragma unroll
      for (int j = 0; j < EPF_CH; j++) win[5 + j] = nx[j];
    }
  }
}

__global__ __launch_bounds__(EPF_LINES) void k_epf_hpass(PassArgs A) {
  int b = blockIdx.x;
  const bool second = b >= A.blocks0;
  const PassJob J = second ? A.job[1] : A.job[0];
  if (second) b -= A.blocks0;
  const int q = b / J.groups, g = b - q * J.groups;
  const int y = g * EPF_LINES + (int)threadIdx.x;
  if (J.w < EPF_SHORT)
    h_pass<true>(pick(J.in, q), pick(J.out, q), J.w, J.h, J.hp, y);
  else
    h_pass<false>(pick(J.in, q), pick(J.out, q), J.w, J.h, J.hp, y);
}

// Scale::new's variance (row_kernels.rs:80-84): max(box(v^2) - box(v)^2, 0).
__global__ void k_epf_var(const float *__restrict__ mean, const float *__restrict__ sq, size_t n,
                          float *__restrict__ var) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float m = mean[i], d = sq[i] - m * m;
  var[i] = d > 0.0f ? d : 0.0f;
}

// --- pools -------------------------------------------------------------------
#define EPF_PPX 8  // pixels per thread in the pool kernel

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

struct PoolArgs {
  const float *bm[2], *bsq[2], *bcr[2], *mean[2], *var[2];  // x-major box planes
  size_t n[2];        // w * hp
  int sh[2], shp[2];  // h and hp of the scale
  int blocks[2];
  double *part[2];  // two partials per workgroup
  const float *src, *cand;
  int cstride, w, h;
  double *pg;  // three partials per workgroup
};

// Workgroups [0, blocks[0]) pool the half scale, the next blocks[1] the
// quarter scale (Scale::compare's pool, row_kernels.rs:86-103), the rest the
// gradients of the full-size planes (src/epf_proxy.rs:218-245,
// row_kernels.rs:104-114): every pixel evaluates its horizontal pair
// (x, x + 1) and vertical pair (y, y + 1) in f32 as the reference does.
__global__ __launch_bounds__(256) void k_epf_pools(PoolArgs A) {
  __shared__ double ws[4][3];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int b = blockIdx.x, s = 2;
  if (b < A.blocks[0]) {
    s = 0;
  } else if (b < A.blocks[0] + A.blocks[1]) {
    s = 1;
    b -= A.blocks[0];
  } else {
    b -= A.blocks[0] + A.blocks[1];
  }
  double t0 = 0.0, t1 = 0.0, t2 = 0.0;
  if (s < 2) {
    const float *bm = A.bm[s], *bsq = A.bsq[s], *bcr = A.bcr[s], *mean = A.mean[s],
                *rvar = A.var[s];
    const size_t n = A.n[s];
    const int h = A.sh[s], hp = A.shp[s];
    // a wavefront's 64 elements lie in one column (hp is a multiple of 64)
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    for (int k = 0; k < EPF_PPX; k++) {
      const size_t b0 = ((size_t)b * EPF_PPX + k) * 256 + (size_t)wv * 64;
      if (b0 >= n) break;
      const size_t col = b0 / (size_t)hp;
      if ((int)(b0 - col * hp) + lane >= h) continue;  // pitch padding
      const size_t i = b0 + lane;
      const float a = mean[i], rv = rvar[i], bb = bm[i], sq = bsq[i], cr = bcr[i];
      float var = sq - bb * bb;
      if (!(var > 0.0f)) var = 0.0f;
      const float cov = (rv == 0.0f || var == 0.0f) ? 0.0f : cr - a * bb;
      float v = 1.0f - (1.0f - (a - bb) * (a - bb)) * (2.0f * cov + 0.0009f) / (rv + var + 0.0009f);
      if (!(v > 0.0f)) v = 0.0f;
      t0 += (double)v;
      t1 += (double)(v * v * v * v);
    }
  } else {
    auto pair = [&](float s0, float s1, float c0, float c1) {
      const float a = s1 - s0, d = c1 - c0;
      float x1 = fabsf(d) - fabsf(a);
      if (!(x1 > 0.0f)) x1 = 0.0f;
      float loss = fabsf(a) - fabsf(d);
      if (!(loss > 0.0f)) loss = 0.0f;
      t0 += (double)fabsf(d - a);
      t1 += (double)x1;
      t2 += (double)(loss * loss * loss * loss);
    };
    const int w = A.w, h = A.h, cs = A.cstride;
    for (int k = 0; k < EPF_PPX; k++) {
      const size_t i = ((size_t)b * EPF_PPX + k) * 256 + tid;
      if (i >= (size_t)w * h) break;
      const int y = (int)(i / (size_t)w), x = (int)(i - (size_t)y * w);
      const float *sr = A.src + (size_t)y * w, *cr = A.cand + (size_t)y * cs;
      if (x + 1 < w) pair(sr[x], sr[x + 1], cr[x], cr[x + 1]);
      if (y + 1 < h) pair(sr[x], sr[x + w], cr[x], cr[x + cs]);
    }
  }
  t0 = wave_sum_d(t0);
  t1 = wave_sum_d(t1);
  t2 = wave_sum_d(t2);
  if (lane == 0) {
    ws[wave][0] = t0;
    ws[wave][1] = t1;
    ws[wave][2] = t2;
  }
  __syncthreads();
  const double r = tid < 3 ? ((ws[0][tid] + ws[1][tid]) + ws[2][tid]) + ws[3][tid] : 0.0;
  if (s < 2) {
    if (tid < 2) A.part[s][2 * (size_t)b + tid] = r;
  } else if (tid < 3) {
    A.pg[3 * (size_t)b + tid] = r;
  }
}

// The six terms and the weighted score (src/epf_proxy.rs:186-214).  The
// seven pooled sums are reduced together in a fixed order: each thread sums
// a contiguous run of partials, then one tree over the 256 threads.
__global__ __launch_bounds__(256) void k_epf_final(const double *__restrict__ ph, int nbh, int hn,
                                                   const double *__restrict__ pq, int nbq, int qn,
                                                   const double *__restrict__ pg, int nbg,
                                                   float ng, float *__restrict__ out) {
  __shared__ double sh[7][256];
  const int tid = threadIdx.x;
  double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int per_h = (nbh + 255) / 256, per_q = (nbq + 255) / 256, per_g = (nbg + 255) / 256;
  for (int k = tid * per_h; k < min(nbh, (tid + 1) * per_h); k++) {
    acc[0] += ph[2 * (size_t)k];
    acc[1] += ph[2 * (size_t)k + 1];
  }
  for (int k = tid * per_q; k < min(nbq, (tid + 1) * per_q); k++) {
    acc[2] += pq[2 * (size_t)k];
    acc[3] += pq[2 * (size_t)k + 1];
  }
  for (int k = tid * per_g; k < min(nbg, (tid + 1) * per_g); k++) {
    acc[4] += pg[3 * (size_t)k];
    acc[5] += pg[3 * (size_t)k + 1];
    acc[6] += pg[3 * (size_t)k + 2];
  }
  for (int i = 0; i < 7; i++) sh[i][tid] = acc[i];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o)
      for (int i = 0; i < 7; i++) sh[i][tid] += sh[i][tid + o];
    __syncthreads();
  }
  if (tid != 0) return;
  double s[7];
  for (int i = 0; i < 7; i++) s[i] = sh[i][0];
  float t[6];
  t[0] = (float)s[2] / (float)qn;
  t[1] = (float)s[0] / (float)hn;
  t[2] = sqrtf(sqrtf((float)s[1] / (float)hn));
  t[3] = (float)s[4] / ng;
  t[4] = (float)s[5] / ng;
  t[5] = sqrtf(sqrtf((float)s[6] / ng));
  const float W[6] = {109.4163228369586f, 4.685925102200196f, 2.491781270704224f,
                      2.167275675983448f, 4.55227651921194f,  2.847714678987977f};
  float score = -0.0f;
  for (int i = 0; i < 6; i++) {
    out[1 + i] = t[i];
    score += t[i] * W[i];
  }
  out[0] = score;
}

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }
int pool_blocks(size_t n) { return (int)((n + 256 * EPF_PPX - 1) / (256 * EPF_PPX)); }

// One pyramid level of the proxy.
struct EpfScale {
  int w, h, hp;       // hp: pitch of the x-major planes, h rounded up to 64
  float *v;           // source values, row-major
  float *cand;        // candidate values, row-major
  float *vt[3];       // vertical sums, x-major
  float *bt[3];       // box values of the candidate, x-major
  float *mean, *var;  // box mean and variance of the source, x-major
  size_t n() const { return (size_t)w * h; }
  size_t nx() const { return (size_t)w * hp; }
};

}  // namespace

struct jxg_epf_proxy {
  jxg_ctx *ctx;
  int w, h;
  float *src;       // source luma, w x h
  EpfScale s[2];    // half and quarter scale
  float *arena;     // every plane above
  double *part;     // pool partials
  int nb[3];        // pool workgroups: half, quarter, gradients
  float *d_out, *h_out;  // score + 6 terms
};

static void epf_free(jxg_epf_proxy *p) {
  if (!p) return;
  (void)hipFree(p->arena);
  (void)hipFree(p->part);
  (void)hipFree(p->d_out);
  (void)hipHostFree(p->h_out);
  delete p;
}

// Vertical (compare: test = candidate, ref = source; setup: test = source,
// mean and square only) or horizontal pass arguments for both scales.
static PassArgs pass_args(const jxg_epf_proxy *p, bool vertical, bool compare, unsigned *blocks) {
  PassArgs a{};
  const int nq = compare ? 3 : 2;
  for (int k = 0; k < 2; k++) {
    const EpfScale &S = p->s[k];
    PassJob &J = a.job[k];
    J.w = S.w;
    J.h = S.h;
    J.hp = S.hp;
    J.groups = ((vertical ? S.w : S.h) + EPF_LINES - 1) / EPF_LINES;
    for (int q = 0; q < 3; q++) {
      J.in[q] = S.vt[q];
      J.out[q] = vertical ? S.vt[q] : S.bt[q];
    }
    if (vertical) {
      J.test = compare ? S.cand : S.v;
      J.ref = compare ? S.v : nullptr;
    } else if (!compare) {
      J.out[0] = S.mean;
    }
  }
  a.blocks0 = nq * a.job[0].groups;
  *blocks = (unsigned)(a.blocks0 + nq * a.job[1].groups);
  return a;
}

extern "C" {

int jxg_epf_proxy_new(jxg_ctx *ctx, const float *d_y, uint32_t row_stride, uint32_t width,
                      uint32_t height, jxg_epf_proxy **out) {
  if (!ctx || !d_y || !out) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (width < 4 || height < 4 || width > JXG_MAX_DIM || height > JXG_MAX_DIM)
    return set_err(JXG_ERR_INVALID_ARG, "EPF proxy needs 4 <= width, height <= %u (got %ux%u)",
                   JXG_MAX_DIM, width, height);
  if (row_stride < width) return set_err(JXG_ERR_INVALID_ARG, "row stride %u < width", row_stride);
  HIP_TRY(hipSetDevice(ctx->device));
  jxg_epf_proxy *p = new jxg_epf_proxy();
  p->ctx = ctx;
  p->w = (int)width;
  p->h = (int)height;
  p->s[0].w = p->w / 2;
  p->s[0].h = p->h / 2;
  p->s[1].w = p->s[0].w / 2;
  p->s[1].h = p->s[0].h / 2;
  const size_t n = (size_t)p->w * p->h;
  for (int k = 0; k < 2; k++) p->s[k].hp = (p->s[k].h + EPF_LINES - 1) & ~(EPF_LINES - 1);
  // the source plane, then per scale 2 row-major and 8 x-major planes, each
  // starting on a 256-byte boundary
  auto up = [](size_t v) { return (v + 63) & ~(size_t)63; };
  size_t floats = up(n);
  for (int k = 0; k < 2; k++) floats += 2 * up(p->s[k].n()) + 8 * p->s[k].nx();
  p->nb[0] = pool_blocks(p->s[0].nx());
  p->nb[1] = pool_blocks(p->s[1].nx());
  p->nb[2] = pool_blocks(n);
  const size_t parts = 2 * (size_t)p->nb[0] + 2 * (size_t)p->nb[1] + 3 * (size_t)p->nb[2];
  bool ok = hipMalloc((void **)&p->arena, floats * 4) == hipSuccess &&
            hipMalloc((void **)&p->part, parts * 8) == hipSuccess &&
            hipMalloc((void **)&p->d_out, 8 * 4) == hipSuccess &&
            hipHostMalloc((void **)&p->h_out, 8 * 4, hipHostMallocDefault) == hipSuccess;
  if (!ok) {
    epf_free(p);
    return set_err(JXG_ERR_ALLOC, "EPF proxy buffers for %ux%u", width, height);
  }
  float *f = p->arena;
  p->src = f;
  f += up(n);
  for (int k = 0; k < 2; k++) {
    EpfScale &S = p->s[k];
    S.v = f;
    f += up(S.n());
    S.cand = f;
    f += up(S.n());
    float **planes[8] = {&S.vt[0], &S.vt[1], &S.vt[2], &S.bt[0],
                         &S.bt[1], &S.bt[2], &S.mean, &S.var};
    for (float **pl : planes) {
      *pl = f;
      f += S.nx();
    }
  }
  hipStream_t st = ctx->stream;
  HIP_TRY(hipMemcpy2DAsync(p->src, width * 4, d_y, (size_t)row_stride * 4, width * 4, height,
                           hipMemcpyDeviceToDevice, st));
  const EpfScale &H = p->s[0], &Q = p->s[1];
  k_epf_half<<<blocks_for(H.n()), 256, 0, st>>>(p->src, p->w, p->w, p->h, 1, H.v);
  k_epf_half<<<blocks_for(Q.n()), 256, 0, st>>>(H.v, H.w, H.w, H.h, 0, Q.v);
  unsigned nb = 0;
  const PassArgs va = pass_args(p, true, false, &nb);
  k_epf_vpass<<<nb, EPF_LINES, 0, st>>>(va);
  const PassArgs ha = pass_args(p, false, false, &nb);
  k_epf_hpass<<<nb, EPF_LINES, 0, st>>>(ha);
  for (int k = 0; k < 2; k++) {
    const EpfScale &S = p->s[k];
    k_epf_var<<<blocks_for(S.nx()), 256, 0, st>>>(S.mean, S.bt[1], S.nx(), S.var);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    epf_free(p);
    return set_err(JXG_ERR_DEVICE, "EPF proxy setup: %s", hipGetErrorString(e));
  }
  *out = p;
  return JXG_OK;
}

void jxg_epf_proxy_destroy(jxg_epf_proxy *p) {
  if (!p) return;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  epf_free(p);
}

int jxg_epf_proxy_compare(jxg_epf_proxy *p, const float *d_y, uint32_t row_stride, float *score,
                          float terms[6]) {
  if (!p || !d_y || !score) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  if (row_stride < (uint32_t)p->w) return set_err(JXG_ERR_INVALID_ARG, "row stride %u < width", row_stride);
  jxg_ctx *ctx = p->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const EpfScale &H = p->s[0], &Q = p->s[1];
  k_epf_half<<<blocks_for(H.n()), 256, 0, st>>>(d_y, (int)row_stride, p->w, p->h, 1, H.cand);
  k_epf_half<<<blocks_for(Q.n()), 256, 0, st>>>(H.cand, H.w, H.w, H.h, 0, Q.cand);
  unsigned nb = 0;
  const PassArgs va = pass_args(p, true, true, &nb);
  k_epf_vpass<<<nb, EPF_LINES, 0, st>>>(va);
  const PassArgs ha = pass_args(p, false, true, &nb);
  k_epf_hpass<<<nb, EPF_LINES, 0, st>>>(ha);
  double *ph = p->part, *pq = ph + 2 * (size_t)p->nb[0], *pg = pq + 2 * (size_t)p->nb[1];
  PoolArgs pa{};
  for (int k = 0; k < 2; k++) {
    const EpfScale &S = p->s[k];
    pa.bm[k] = S.bt[0];
    pa.bsq[k] = S.bt[1];
    pa.bcr[k] = S.bt[2];
    pa.mean[k] = S.mean;
    pa.var[k] = S.var;
    pa.n[k] = S.nx();
    pa.sh[k] = S.h;
    pa.shp[k] = S.hp;
    pa.blocks[k] = p->nb[k];
  }
  pa.part[0] = ph;
  pa.part[1] = pq;
  pa.src = p->src;
  pa.cand = d_y;
  pa.cstride = (int)row_stride;
  pa.w = p->w;
  pa.h = p->h;
  pa.pg = pg;
  k_epf_pools<<<p->nb[0] + p->nb[1] + p->nb[2], 256, 0, st>>>(pa);
  const float ng = (float)((size_t)p->h * (p->w - 1) + (size_t)(p->h - 1) * p->w);
  k_epf_final<<<1, 256, 0, st>>>(ph, p->nb[0], (int)H.n(), pq, p->nb[1], (int)Q.n(), pg,
                                 p->nb[2], ng, p->d_out);
  LAUNCH_CHECK("EPF proxy compare");
  HIP_TRY(hipMemcpyAsync(p->h_out, p->d_out, 7 * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  *score = p->h_out[0];
  if (terms)
    for (int i = 0; i < 6; i++) terms[i] = p->h_out[1 + i];
  return JXG_OK;
}

}  // extern "C"
