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
// the streaming kernels are pinned to it bit for bit), kept in order by one
// thread per column (vertical pass) and one thread per row (horizontal
// pass).  The pooled sums over all pixels are accumulated in f64 (per row in
// pixel order, then over rows in order), which is deterministic and more
// accurate than the reference's single f32 running sum.
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

// Vertical running sums of test, test^2 and (ref * test) per column, as
// box_filter.rs:118-150 (sums from -0.0, output sum * 0.2 before the step).
__global__ void k_epf_box_v(const float *__restrict__ test, const float *__restrict__ ref, int w,
                            int h, float *__restrict__ vm, float *__restrict__ vsq,
                            float *__restrict__ vcr) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= w) return;
  float m = -0.0f, sq = -0.0f, cr = -0.0f;
  for (int y = -2; y <= 2; y++) {
    const size_t i = (size_t)mirror(y, h) * w + x;
    const float v = test[i];
    m += v;
    sq += v * v;
    if (ref) cr += ref[i] * v;
  }
  for (int y = 0; y < h; y++) {
    const size_t o = (size_t)y * w + x;
    vm[o] = m * 0.2f;
    vsq[o] = sq * 0.2f;
    if (ref) vcr[o] = cr * 0.2f;
    if (y + 1 < h) {
      const size_t ia = (size_t)mirror(y + 3, h) * w + x, ib = (size_t)mirror(y - 2, h) * w + x;
      const float a = test[ia], b = test[ib];
      m += a - b;
      sq += a * a - b * b;
      if (ref) cr += ref[ia] * a - ref[ib] * b;
    }
  }
}

// Horizontal running sums of a row of vertical sums; returns nothing, the
// caller consumes (x, value) in order.
struct HSum {
  const float *row;
  int w;
  float sum;
  __device__ __forceinline__ void init() {
    sum = -0.0f;
    for (int x = -2; x <= 2; x++) sum += row[mirror(x, w)];
  }
  __device__ __forceinline__ float value_then_step(int x) {
    const float v = sum * 0.2f;
    if (x + 1 < w) sum += row[mirror(x + 3, w)] - row[mirror(x - 2, w)];
    return v;
  }
};

// Scale::new (src/epf_proxy.rs:75-97): box mean and variance.
__global__ void k_epf_box_h_new(const float *__restrict__ vm, const float *__restrict__ vsq, int w,
                                int h, float *__restrict__ mean, float *__restrict__ var) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= h) return;
  HSum sm{vm + (size_t)y * w, w, 0.0f}, ss{vsq + (size_t)y * w, w, 0.0f};
  sm.init();
  ss.init();
  for (int x = 0; x < w; x++) {
    const float m = sm.value_then_step(x), s = ss.value_then_step(x);
    const float d = s - m * m;  // variance (row_kernels.rs:80-84)
    mean[(size_t)y * w + x] = m;
    var[(size_t)y * w + x] = d > 0.0f ? d : 0.0f;
  }
}

// Scale::compare's pool (row_kernels.rs:86-103) per row, in f64.
__global__ void k_epf_box_h_pool(const float *__restrict__ vm, const float *__restrict__ vsq,
                                 const float *__restrict__ vcr, const float *__restrict__ rmean,
                                 const float *__restrict__ rvar, int w, int h,
                                 double *__restrict__ part) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= h) return;
  HSum sm{vm + (size_t)y * w, w, 0.0f}, ss{vsq + (size_t)y * w, w, 0.0f},
      sc{vcr + (size_t)y * w, w, 0.0f};
  sm.init();
  ss.init();
  sc.init();
  double p0 = 0.0, p1 = 0.0;
  for (int x = 0; x < w; x++) {
    const float b = sm.value_then_step(x), sq = ss.value_then_step(x), cr = sc.value_then_step(x);
    const size_t i = (size_t)y * w + x;
    const float a = rmean[i], rv = rvar[i];
    float var = sq - b * b;
    if (!(var > 0.0f)) var = 0.0f;
    const float cov = (rv == 0.0f || var == 0.0f) ? 0.0f : cr - a * b;
    float s = 1.0f - (1.0f - (a - b) * (a - b)) * (2.0f * cov + 0.0009f) / (rv + var + 0.0009f);
    if (!(s > 0.0f)) s = 0.0f;
    p0 += (double)s;
    p1 += (double)(s * s * s * s);
  }
  part[2 * y] = p0;
  part[2 * y + 1] = p1;
}

// gradients (src/epf_proxy.rs:218-245, row_kernels.rs:104-114) per row in
// f64: the horizontal pairs of row y and the vertical pairs (y, y + 1).
__global__ void k_epf_grad(const float *__restrict__ src, const float *__restrict__ cand,
                           int cstride, int w, int h, double *__restrict__ part) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= h) return;
  double e = 0.0, ex = 0.0, l4 = 0.0;
  auto pair = [&](float s0, float s1, float c0, float c1) {
    const float a = s1 - s0, b = c1 - c0;
    float x1 = fabsf(b) - fabsf(a);
    if (!(x1 > 0.0f)) x1 = 0.0f;
    float loss = fabsf(a) - fabsf(b);
    if (!(loss > 0.0f)) loss = 0.0f;
    e += (double)fabsf(b - a);
    ex += (double)x1;
    l4 += (double)(loss * loss * loss * loss);
  };
  const float *sr = src + (size_t)y * w, *cr = cand + (size_t)y * cstride;
  for (int x = 0; x + 1 < w; x++) pair(sr[x], sr[x + 1], cr[x], cr[x + 1]);
  if (y + 1 < h)
    for (int x = 0; x < w; x++) pair(sr[x], sr[x + w], cr[x], cr[x + cstride]);
  part[3 * y] = e;
  part[3 * y + 1] = ex;
  part[3 * y + 2] = l4;
}

// The six terms and the weighted score (src/epf_proxy.rs:186-214).
__global__ void k_epf_final(const double *__restrict__ ph, int hh, int hn,
                            const double *__restrict__ pq, int qh, int qn,
                            const double *__restrict__ pg, int h, float ng,
                            float *__restrict__ out) {
  if (threadIdx.x != 0) return;
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int y = 0; y < hh; y++) {
    s[0] += ph[2 * y];
    s[1] += ph[2 * y + 1];
  }
  for (int y = 0; y < qh; y++) {
    s[2] += pq[2 * y];
    s[3] += pq[2 * y + 1];
  }
  for (int y = 0; y < h; y++) {
    s[4] += pg[3 * y];
    s[5] += pg[3 * y + 1];
    s[6] += pg[3 * y + 2];
  }
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

unsigned blocks_for(int n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct jxg_epf_proxy {
  jxg_ctx *ctx;
  int w, h, hw, hh, qw, qh;
  float *src;                   // source luma, w x h
  float *hv, *hm, *hvar;        // half scale
  float *qv, *qm, *qvar;        // quarter scale
  float *cand_h, *cand_q;       // candidate pyramid
  float *vm, *vsq, *vcr;        // vertical sums (half-scale size)
  double *part;                 // row partials
  float *d_out, *h_out;         // score + 6 terms
};

static void epf_free(jxg_epf_proxy *p) {
  if (!p) return;
  float *bufs[] = {p->src, p->hv, p->hm, p->hvar, p->qv, p->qm, p->qvar, p->cand_h, p->cand_q,
                   p->vm, p->vsq, p->vcr, p->d_out};
  for (float *b : bufs) (void)hipFree(b);
  (void)hipFree(p->part);
  (void)hipHostFree(p->h_out);
  delete p;
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
  p->hw = p->w / 2;
  p->hh = p->h / 2;
  p->qw = p->hw / 2;
  p->qh = p->hh / 2;
  const size_t n = (size_t)p->w * p->h, hn = (size_t)p->hw * p->hh, qn = (size_t)p->qw * p->qh;
  const size_t rows = (size_t)(p->h > 1 ? p->h : 1);
  bool ok = hipMalloc((void **)&p->src, n * 4) == hipSuccess &&
            hipMalloc((void **)&p->hv, hn * 4) == hipSuccess &&
            hipMalloc((void **)&p->hm, hn * 4) == hipSuccess &&
            hipMalloc((void **)&p->hvar, hn * 4) == hipSuccess &&
            hipMalloc((void **)&p->qv, qn * 4) == hipSuccess &&
            hipMalloc((void **)&p->qm, qn * 4) == hipSuccess &&
            hipMalloc((void **)&p->qvar, qn * 4) == hipSuccess &&
            hipMalloc((void **)&p->cand_h, hn * 4) == hipSuccess &&
            hipMalloc((void **)&p->cand_q, qn * 4) == hipSuccess &&
            hipMalloc((void **)&p->vm, hn * 4) == hipSuccess &&
            hipMalloc((void **)&p->vsq, hn * 4) == hipSuccess &&
            hipMalloc((void **)&p->vcr, hn * 4) == hipSuccess &&
            hipMalloc((void **)&p->part, (2 * (size_t)p->hh + 2 * (size_t)p->qh + 3 * rows) * 8) ==
                hipSuccess &&
            hipMalloc((void **)&p->d_out, 8 * 4) == hipSuccess &&
            hipHostMalloc((void **)&p->h_out, 8 * 4, hipHostMallocDefault) == hipSuccess;
  if (!ok) {
    epf_free(p);
    return set_err(JXG_ERR_ALLOC, "EPF proxy buffers for %ux%u", width, height);
  }
  hipStream_t st = ctx->stream;
  HIP_TRY(hipMemcpy2DAsync(p->src, width * 4, d_y, (size_t)row_stride * 4, width * 4, height,
                           hipMemcpyDeviceToDevice, st));
  k_epf_half<<<blocks_for((int)hn), 256, 0, st>>>(p->src, p->w, p->w, p->h, 1, p->hv);
  k_epf_half<<<blocks_for((int)qn), 256, 0, st>>>(p->hv, p->hw, p->hw, p->hh, 0, p->qv);
  k_epf_box_v<<<blocks_for(p->hw), 256, 0, st>>>(p->hv, nullptr, p->hw, p->hh, p->vm, p->vsq,
                                                  nullptr);
  k_epf_box_h_new<<<blocks_for(p->hh), 256, 0, st>>>(p->vm, p->vsq, p->hw, p->hh, p->hm, p->hvar);
  k_epf_box_v<<<blocks_for(p->qw), 256, 0, st>>>(p->qv, nullptr, p->qw, p->qh, p->vm, p->vsq,
                                                  nullptr);
  k_epf_box_h_new<<<blocks_for(p->qh), 256, 0, st>>>(p->vm, p->vsq, p->qw, p->qh, p->qm, p->qvar);
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
  const int hn = p->hw * p->hh, qn = p->qw * p->qh;
  double *ph = p->part, *pq = ph + 2 * p->hh, *pg = pq + 2 * p->qh;
  k_epf_half<<<blocks_for(hn), 256, 0, st>>>(d_y, (int)row_stride, p->w, p->h, 1, p->cand_h);
  k_epf_half<<<blocks_for(qn), 256, 0, st>>>(p->cand_h, p->hw, p->hw, p->hh, 0, p->cand_q);
  k_epf_box_v<<<blocks_for(p->hw), 256, 0, st>>>(p->cand_h, p->hv, p->hw, p->hh, p->vm, p->vsq,
                                                  p->vcr);
  k_epf_box_h_pool<<<blocks_for(p->hh), 256, 0, st>>>(p->vm, p->vsq, p->vcr, p->hm, p->hvar, p->hw,
                                                       p->hh, ph);
  k_epf_box_v<<<blocks_for(p->qw), 256, 0, st>>>(p->cand_q, p->qv, p->qw, p->qh, p->vm, p->vsq,
                                                  p->vcr);
  k_epf_box_h_pool<<<blocks_for(p->qh), 256, 0, st>>>(p->vm, p->vsq, p->vcr, p->qm, p->qvar, p->qw,
                                                       p->qh, pq);
  k_epf_grad<<<blocks_for(p->h), 256, 0, st>>>(p->src, d_y, (int)row_stride, p->w, p->h, pg);
  const float ng = (float)((size_t)p->h * (p->w - 1) + (size_t)(p->h - 1) * p->w);
  k_epf_final<<<1, 64, 0, st>>>(ph, p->hh, hn, pq, p->qh, qn, pg, p->h, ng, p->d_out);
  LAUNCH_CHECK("EPF proxy compare");
  HIP_TRY(hipMemcpyAsync(p->h_out, p->d_out, 7 * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  *score = p->h_out[0];
  if (terms)
    for (int i = 0; i < 6; i++) terms[i] = p->h_out[1 + i];
  return JXG_OK;
}

}  // extern "C"
