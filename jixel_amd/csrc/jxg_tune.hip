// jxg_tune.hip - gradient tuning of the per-block quant field: a hand-written
// adjoint through DCT -> soft quantization -> IDCT -> channel-weighted
// squared error plus a log-rate proxy, and Adam on the log of the field.
//
// The objective (this build's own; the reference has no differentiable path):
// for every transform T of the layout with theta_T the mean field of its
// blocks, per channel c and AC coefficient k (the lowest cov_y x cov_x
// frequencies, the DC image, are left exact)
//   s_k     = 1 / (inv_c[k] * scale * theta_T * qm_c)   the quantizer's step
//   Chat_k  = C_k + s_k u_k,  u_k uniform in [-1/2, 1/2) (seeded hash per step)
//   Xhat    = C_h^T Chat C_w                              exact inverse transform
//   L       = sum_c w_c |Xhat - X|^2 (pixels inside the image)
//             + lambda sum_k log2(1 + |Chat_k / s_k|)
// C_n[k][j] = c_k cos(pi k (2j + 1) / 2n) (c_0 = 1, c_k = sqrt 2), so the
// forward transform C = C_h X C_w^T / (w h) is the front end's scaled DCT-II
// and C_n C_n^T = n I.  The adjoint of the inverse transform is the forward
// product without the 1/(w h) scale: dL/dChat = C_h R C_w^T with
// R = 2 w_c (Xhat - X) masked to the image.  dChat/dtheta = -u s / theta and
// dq/dtheta = (q - u) / theta for q = Chat / s.
//
// Kernels (one workgroup per 32x32 region, the transforms whose footprint
// lies inside it; 256 threads, one 8x8 block per wavefront and step):
//   k_tune_coeffs  C of every transform, once (FP32, dense products with the
//                  cosine matrices in LDS)
//   k_tune_step    per step: noise, Chat, Xhat, residual, adjoint, and per
//                  transform dL/dtheta_T (written / n_T to its blocks) and
//                  the region's loss; wave-shuffle trees and block-order sums,
//                  so results do not depend on scheduling
//   k_tune_adam    Adam on phi = ln theta per block
//   k_tune_round   the integer field: round(theta) clamped to 1..255
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <vector>

#include "jxg_ctx.h"
#include "jxg_device.h"

using namespace jxg;

namespace {

constexpr int TP = 33;  // LDS pitch of the 32x32 region planes

__device__ __host__ __forceinline__ uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// uniform in [-1/2, 1/2) of (seed, step, channel, first block, storage index)
__device__ __forceinline__ float tune_noise(uint32_t seed, uint32_t step, int c, uint32_t first,
                                            int k) {
  uint32_t h = lowbias32(seed ^ 0x9e3779b9U);
  h = lowbias32(h ^ step);
  h = lowbias32(h ^ (uint32_t)(c * 0x10000 + k));
  h = lowbias32(h ^ first);
  return (float)(h >> 8) * (1.0f / 16777216.0f) - 0.5f;
}

__device__ __forceinline__ bool tune_strategy(int s) {
  return s == JXG_DCT || s == JXG_DCT16X8 || s == JXG_DCT8X16 || s == JXG_DCT16X16 ||
         s == JXG_DCT32X32 || s == JXG_DCT32X16 || s == JXG_DCT16X32 || s == JXG_DCT32X8 ||
         s == JXG_DCT8X32;
}

struct TuneArgs {
  const float *xyb;
  int stride, ph, W, H, xb, yb;
  const uint8_t *cells;
  const int32_t *tmap;  // per block: first block index of its transform in a region, or -1
  const float *tables;
  const jxg_derived *D;
  const float *cmat;  // C_8 (64), C_16 (256), C_32 (1024)
  float *coef;        // 3 planes like xyb: C of each transform at its footprint
  float lambda;
  uint32_t seed;
};

struct RegionLds {
  float a[32 * TP], b[32 * TP], x[32 * TP];
  float cm[64 + 256 + 1024];
  // per block of the region: transform origin (px, region-relative), size, strategy
  int8_t ox[16], oy[16], tw[16], th[16], st[16];
  float theta[16];  // the transform's mean field, at every covered block
  float bsum[3][16];  // per block sums: gradient, distortion, rate
};

__device__ __forceinline__ const float *cm_of(const RegionLds &L, int n) {
  return n == 8 ? L.cm : (n == 16 ? L.cm + 64 : L.cm + 320);
}

// Per block of the region: its transform's geometry (if the transform's
// first block is in the region and the footprint fits), or st = -1.
__device__ void region_setup(RegionLds &L, const TuneArgs &A, int rbx, int rby,
                             const float *theta /* may be null */) {
  const int tid = threadIdx.x;
  for (int e = tid; e < 64 + 256 + 1024; e += blockDim.x) L.cm[e] = A.cmat[e];
  if (tid < 16) {
    const int lx = tid & 3, ly = tid >> 2, bx = rbx + lx, by = rby + ly;
    int8_t st = -1;
    if (bx < A.xb && by < A.yb) {
      const int32_t f = A.tmap[(size_t)by * A.xb + bx];
      if (f >= 0) {
        const int fx = f % A.xb, fy = f / A.xb;
        const int s = A.cells[f] >> 1;
        L.ox[tid] = (int8_t)((fx - rbx) * 8);
        L.oy[tid] = (int8_t)((fy - rby) * 8);
        L.tw[tid] = (int8_t)(cov_x(s) * 8);
        L.th[tid] = (int8_t)(cov_y(s) * 8);
        st = (int8_t)s;
        // mean of the transform's blocks in block raster order
        float sum = 0.0f;
        if (theta)
          for (int iy = 0; iy < cov_y(s); iy++)
            for (int ix = 0; ix < cov_x(s); ix++) sum += theta[(size_t)(fy + iy) * A.xb + fx + ix];
        L.theta[tid] = sum / (float)(cov_x(s) * cov_y(s));
      }
    }
    L.st[tid] = st;
  }
  __syncthreads();
}

// Thread e -> region pixel: block b = e >> 6 of the 4x4 blocks, row and
// column inside it.
__device__ __forceinline__ void pix_of(int e, int *b, int *py, int *px) {
  *b = e >> 6;
  *py = (*b >> 2) * 8 + ((e >> 3) & 7);
  *px = (*b & 3) * 8 + (e & 7);
}

// dst[y][x] = sum_j M[f][j] src[...]: one separable pass inside each
// transform's footprint.  ROW: along x (f = x - ox over the transform's
// width), else along y.  TRANS: use M[j][f] (the transposed matrix).
template <bool ROW, bool TRANS>
__device__ void pass(RegionLds &L, const float *src, float *dst, bool scale) {
  for (int e = threadIdx.x; e < 1024; e += blockDim.x) {
    int b, py, px;
    pix_of(e, &b, &py, &px);
    if (L.st[b] < 0) continue;
    const int ox = L.ox[b], oy = L.oy[b], w = L.tw[b], h = L.th[b];
    const int n = ROW ? w : h;
    const float *M = cm_of(L, n);
    const int f = ROW ? px - ox : py - oy;
    float acc = 0.0f;
    for (int j = 0; j < n; j++) {
      const float m = TRANS ? M[j * n + f] : M[f * n + j];
      const float v = ROW ? src[py * TP + ox + j] : src[(oy + j) * TP + px];
      acc += m * v;
    }
    dst[py * TP + px] = scale ? acc * (1.0f / (float)n) : acc;
  }
  __syncthreads();
}

__device__ __forceinline__ void load_region(float *dst, const float *plane, int stride, int px0,
                                            int py0) {
  for (int e = threadIdx.x; e < 1024; e += blockDim.x) {
    const int r = e >> 5, c = e & 31;
    dst[r * TP + c] = plane[(size_t)(py0 + r) * stride + px0 + c];
  }
}

__global__ __launch_bounds__(256) void k_tune_coeffs(TuneArgs A) {
  __shared__ RegionLds L;
  const int rbx = blockIdx.x * 4, rby = blockIdx.y * 4;
  region_setup(L, A, rbx, rby, nullptr);
  const size_t plane = (size_t)A.stride * A.ph;
  for (int c = 0; c < 3; c++) {
    load_region(L.x, A.xyb + c * plane, A.stride, rbx * 8, rby * 8);
    __syncthreads();
    pass<true, false>(L, L.x, L.a, true);   // along x: C_w X^T / w
    pass<false, false>(L, L.a, L.b, true);  // along y: C_h . / h
    for (int e = threadIdx.x; e < 1024; e += blockDim.x) {
      const int r = e >> 5, col = e & 31;
      A.coef[c * plane + (size_t)(rby * 8 + r) * A.stride + rbx * 8 + col] = L.b[r * TP + col];
    }
    __syncthreads();
  }
}

__device__ __forceinline__ float wave_sum_f(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__global__ __launch_bounds__(256) void k_tune_step(TuneArgs A, const float *theta, uint32_t step,
                                                   float *grad, float *region_loss) {
  __shared__ RegionLds L;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int rbx = blockIdx.x * 4, rby = blockIdx.y * 4;
  region_setup(L, A, rbx, rby, theta);
  const size_t plane = (size_t)A.stride * A.ph;
  const jxg_derived *D = A.D;
  float tg[4] = {0, 0, 0, 0}, td[4] = {0, 0, 0, 0}, tr[4] = {0, 0, 0, 0};
  for (int c = 0; c < 3; c++) {
    load_region(L.x, A.xyb + c * plane, A.stride, rbx * 8, rby * 8);
    load_region(L.a, A.coef + c * plane, A.stride, rbx * 8, rby * 8);
    __syncthreads();
    // noise, soft levels and rate terms; L.b = Chat
    float un_r[4], st_r[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int e = tid + 256 * j;
      int b, py, px;
      pix_of(e, &b, &py, &px);
      un_r[j] = 0.0f;
      st_r[j] = 0.0f;
      float chat = L.a[py * TP + px];
      if (L.st[b] >= 0) {
        const int s = L.st[b], w = L.tw[b], h = L.th[b];
        const int v = py - L.oy[b], u = px - L.ox[b];
        if (!(v < h / 8 && u < w / 8)) {
          const int k = storage_index(s, w, h, v, u);
          const int kind = qt_kind(s);
          const float iv = A.tables[D->qt_off[kind] + (3 + c) * D->qt_n[kind] + k];
          const float qs = D->scale * L.theta[b] * D->qm_mult[c];
          const float stp = 1.0f / (iv * qs);
          const int fx = rbx + (L.ox[b] >> 3), fy = rby + (L.oy[b] >> 3);
          const float un = tune_noise(A.seed, step, c, (uint32_t)(fy * A.xb + fx), k);
          chat = chat + stp * un;
          const float q = chat / stp;
          const float aq = fabsf(q);
          tr[j] += A.lambda * log2f(1.0f + aq);
          // lambda d/dtheta log2(1 + |q|)
          tg[j] += A.lambda * (copysignf(1.0f, q) / ((1.0f + aq) * 0.6931471805599453f)) *
                   (q - un) / L.theta[b];
          un_r[j] = un;
          st_r[j] = stp;
        }
      }
      L.b[py * TP + px] = chat;
    }
    __syncthreads();
    // Xhat = C_h^T Chat C_w: along x (transposed), then along y (transposed)
    pass<true, true>(L, L.b, L.a, false);
    pass<false, true>(L, L.a, L.b, false);
    // residual R = 2 w_c (Xhat - X) inside the image; distortion
    const float wc = D->channel_weight[c];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int e = tid + 256 * j;
      int b, py, px;
      pix_of(e, &b, &py, &px);
      float r = 0.0f;
      if (L.st[b] >= 0 && rbx * 8 + px < A.W && rby * 8 + py < A.H) {
        const float d = L.b[py * TP + px] - L.x[py * TP + px];
        td[j] += wc * d * d;
        r = 2.0f * wc * d;
      }
      L.a[py * TP + px] = r;
    }
    __syncthreads();
    // dL/dChat = C_h R C_w^T: along x, then along y (no scale)
    pass<true, false>(L, L.a, L.b, false);
    pass<false, false>(L, L.b, L.a, false);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int e = tid + 256 * j;
      int b, py, px;
      pix_of(e, &b, &py, &px);
      if (L.st[b] >= 0 && st_r[j] != 0.0f)
        tg[j] += L.a[py * TP + px] * (-un_r[j] * st_r[j] / L.theta[b]);
    }
    __syncthreads();
  }
  // per block sums (one block per wavefront and j), then per transform in
  // block raster order
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const float g = wave_sum_f(tg[j]), d = wave_sum_f(td[j]), r = wave_sum_f(tr[j]);
    if (lane == 0) {
      const int b = (tid + 256 * j) >> 6;
      L.bsum[0][b] = g;
      L.bsum[1][b] = d;
      L.bsum[2][b] = r;
    }
  }
  (void)wave;
  __syncthreads();
  if (tid < 16) {
    const int b = tid;
    if (L.st[b] >= 0 && L.ox[b] == (b & 3) * 8 && L.oy[b] == (b >> 2) * 8) {
      // the transform's first block: sum its blocks in raster order
      const int s = L.st[b], cx = cov_x(s), cy = cov_y(s);
      float g = 0.0f;
      for (int iy = 0; iy < cy; iy++)
        for (int ix = 0; ix < cx; ix++) g += L.bsum[0][b + iy * 4 + ix];
      const float gb = g / (float)(cx * cy);
      for (int iy = 0; iy < cy; iy++)
        for (int ix = 0; ix < cx; ix++)
          grad[(size_t)(rby + (b >> 2) + iy) * A.xb + rbx + (b & 3) + ix] = gb;
    }
  }
  if (tid == 0) {
    float loss = 0.0f;
    for (int b = 0; b < 16; b++)
      if (L.st[b] >= 0) loss += L.bsum[1][b] + L.bsum[2][b];
    region_loss[blockIdx.y * gridDim.x + blockIdx.x] = loss;
  }
}

// Block -> first block of its transform when the transform takes part and
// its footprint lies inside one 32x32 region, else -1; gradients of the
// left-out blocks are zero.
__global__ void k_tune_map(int xb, int yb, int32_t *tmap, float *grad) {
  const int bx = blockIdx.x * 256 + threadIdx.x, by = blockIdx.y;
  if (bx >= xb) return;
  const size_t i = (size_t)by * xb + bx;
  tmap[i] = -1;
  grad[i] = 0.0f;
}

__global__ void k_tune_map_fill(const uint8_t *cells, int xb, int yb, int32_t *tmap,
                                uint32_t *excluded) {
  const int bx = blockIdx.x * 256 + threadIdx.x, by = blockIdx.y;
  if (bx >= xb) return;
  const uint8_t cell = cells[(size_t)by * xb + bx];
  if (!(cell & 1)) return;
  const int s = cell >> 1;
  bool ok = s < JXG_NUM_STRATEGIES && tune_strategy(s);
  if (ok) {
    const int cx = cov_x(s), cy = cov_y(s);
    ok = bx + cx <= xb && by + cy <= yb && (bx & 3) + cx <= 4 && (by & 3) + cy <= 4;
    if (ok)
      for (int iy = 0; iy < cy; iy++)
        for (int ix = 0; ix < cx; ix++) tmap[(size_t)(by + iy) * xb + bx + ix] = by * xb + bx;
  }
  if (!ok) atomicAdd(excluded, 1u);
}

// Adam on phi = ln theta (bias-corrected), theta kept in [1, 255].
__global__ void k_tune_adam(float *theta, const float *grad, float *m, float *v,
                            const int32_t *tmap, int n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n || tmap[i] < 0) return;
  const float th = theta[i];
  const float g = th * grad[i];  // dL/dphi
  const float mi = b1 * m[i] + (1.0f - b1) * g;
  const float vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float phi = logf(th) - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  theta[i] = fminf(fmaxf(expf(phi), 1.0f), 255.0f);
}

__global__ void k_tune_init(const uint8_t *qf, float *theta, float *m, float *v, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  theta[i] = (float)qf[i];
  m[i] = 0.0f;
  v[i] = 0.0f;
}

// The integer field: every block of a transform gets round(theta_T).
__global__ void k_tune_round(const float *theta, const int32_t *tmap, const uint8_t *cells,
                             const uint8_t *qf_in, int xb, int yb, uint8_t *qf_out) {
  const int bx = blockIdx.x * 256 + threadIdx.x, by = blockIdx.y;
  if (bx >= xb) return;
  const size_t i = (size_t)by * xb + bx;
  const int32_t f = tmap[i];
  if (f < 0) {
    qf_out[i] = qf_in[i];
    return;
  }
  const int fx = f % xb, fy = f / xb, s = cells[f] >> 1;
  float sum = 0.0f;
  for (int iy = 0; iy < cov_y(s); iy++)
    for (int ix = 0; ix < cov_x(s); ix++) sum += theta[(size_t)(fy + iy) * xb + fx + ix];
  const float t = sum / (float)(cov_x(s) * cov_y(s));
  qf_out[i] = (uint8_t)fminf(fmaxf(rintf(t), 1.0f), 255.0f);
}

// Fixed-order sum of the region losses (one workgroup).
__global__ void k_tune_loss_sum(const float *parts, int n, float *out) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += (double)parts[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = (float)s[0];
}

}  // namespace

struct jxg_qf_tuner {
  jxg_ctx *ctx = nullptr;
  TuneArgs A;
  float *d_cmat = nullptr, *d_coef = nullptr, *d_theta = nullptr, *d_grad = nullptr;
  float *d_m = nullptr, *d_v = nullptr, *d_parts = nullptr, *d_loss = nullptr;
  jxg_derived *d_der = nullptr;  // a private copy: the context's may change
  int32_t *d_tmap = nullptr;
  uint32_t *d_excl = nullptr;
  uint32_t max_steps = 0;
  int rx = 0, ry = 0;
};

static void tuner_free(jxg_qf_tuner *t) {
  if (!t) return;
  (void)hipSetDevice(t->ctx->device);
  (void)hipStreamSynchronize(t->ctx->stream);
  (void)hipFree(t->d_cmat);
  (void)hipFree(t->d_coef);
  (void)hipFree(t->d_theta);
  (void)hipFree(t->d_grad);
  (void)hipFree(t->d_m);
  (void)hipFree(t->d_v);
  (void)hipFree(t->d_parts);
  (void)hipFree(t->d_loss);
  (void)hipFree(t->d_der);
  (void)hipFree(t->d_tmap);
  (void)hipFree(t->d_excl);
  delete t;
}

// One loss / gradient evaluation at `theta` into `grad`; the loss goes to
// d_loss[slot].
static int tuner_eval(jxg_qf_tuner *t, const float *theta, uint32_t step, float *grad,
                      uint32_t slot) {
  jxg_ctx *ctx = t->ctx;
  k_tune_step<<<dim3(t->rx, t->ry), 256, 0, ctx->stream>>>(t->A, theta, step, grad, t->d_parts);
  k_tune_loss_sum<<<1, 256, 0, ctx->stream>>>(t->d_parts, t->rx * t->ry, t->d_loss + slot);
  LAUNCH_CHECK("k_tune_step");
  return JXG_OK;
}

extern "C" {

int jxg_qf_tuner_create(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                        uint32_t width, uint32_t height, const uint8_t *d_cells,
                        const jxg_frame_params *params, float lambda, uint32_t seed,
                        uint32_t max_steps, jxg_qf_tuner **out) {
  if (!ctx || !d_xyb || !d_cells || !params || !out)
    return set_err(JXG_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  int rc = check_image(width, height);
  if (rc) return rc;
  if ((rc = check_planes(width, height, stride, padded_h))) return rc;
  if (!(lambda >= 0.0f)) return set_err(JXG_ERR_INVALID_ARG, "lambda %g", (double)lambda);
  if (max_steps > (1u << 20)) return set_err(JXG_ERR_INVALID_ARG, "max_steps %u", max_steps);
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = upload_derived(ctx, params))) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  jxg_qf_tuner *t = new jxg_qf_tuner;
  t->ctx = ctx;
  t->max_steps = max_steps;
  const int xb = (int)((width + 7) / 8), yb = (int)((height + 7) / 8);
  const size_t nb = (size_t)xb * yb;
  t->rx = (xb + 3) / 4;
  t->ry = (yb + 3) / 4;
  const size_t plane = (size_t)stride * padded_h;
  auto fail = [&](int code) {
    tuner_free(t);
    return code;
  };
  if (hipMalloc((void **)&t->d_cmat, (64 + 256 + 1024) * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&t->d_coef, 3 * plane * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&t->d_theta, nb * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&t->d_grad, nb * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&t->d_m, nb * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&t->d_v, nb * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&t->d_parts, (size_t)t->rx * t->ry * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&t->d_loss, (size_t)(max_steps + 2) * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&t->d_der, sizeof(jxg_derived)) != hipSuccess ||
      hipMalloc((void **)&t->d_tmap, nb * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void **)&t->d_excl, sizeof(uint32_t)) != hipSuccess)
    return fail(set_err(JXG_ERR_ALLOC, "tuner allocation failed"));
  {
    std::vector<float> cm(64 + 256 + 1024);
    int off = 0;
    for (int n = 8; n <= 32; n *= 2) {
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++) {
          const double ck = k == 0 ? 1.0 : sqrt(2.0);
          cm[off + k * n + j] = (float)(ck * cos(M_PI * k * (2 * j + 1) / (2.0 * n)));
        }
      off += n * n;
    }
    if (hipMemcpy(t->d_cmat, cm.data(), cm.size() * sizeof(float), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(t->d_der, ctx->h_derived, sizeof(jxg_derived), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemset(t->d_excl, 0, sizeof(uint32_t)) != hipSuccess)
      return fail(set_err(JXG_ERR_DEVICE, "tuner upload failed"));
  }
  TuneArgs &A = t->A;
  A.xyb = d_xyb;
  A.stride = (int)stride;
  A.ph = (int)padded_h;
  A.W = (int)width;
  A.H = (int)height;
  A.xb = xb;
  A.yb = yb;
  A.cells = d_cells;
  A.tmap = t->d_tmap;
  {
    const float *selq;
    const int trc = tables_for(ctx, params, &A.tables, &selq);
    if (trc) return fail(trc);
  }
  A.D = t->d_der;
  A.cmat = t->d_cmat;
  A.coef = t->d_coef;
  A.lambda = lambda;
  A.seed = seed;
  const dim3 gb((xb + 255) / 256, yb);
  k_tune_map<<<gb, 256, 0, ctx->stream>>>(xb, yb, t->d_tmap, t->d_grad);
  k_tune_map_fill<<<gb, 256, 0, ctx->stream>>>(d_cells, xb, yb, t->d_tmap, t->d_excl);
  k_tune_coeffs<<<dim3(t->rx, t->ry), 256, 0, ctx->stream>>>(A);
  if (hipGetLastError() != hipSuccess) return fail(set_err(JXG_ERR_DEVICE, "tuner launch failed"));
  *out = t;
  return JXG_OK;
}

void jxg_qf_tuner_destroy(jxg_qf_tuner *t) { tuner_free(t); }

int jxg_qf_tuner_excluded(jxg_qf_tuner *t, uint32_t *count) {
  if (!t || !count) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(t->ctx->device));
  HIP_TRY(hipStreamSynchronize(t->ctx->stream));
  HIP_TRY(hipMemcpy(count, t->d_excl, sizeof(uint32_t), hipMemcpyDeviceToHost));
  return JXG_OK;
}

int jxg_qf_tuner_loss_grad(jxg_qf_tuner *t, const float *d_theta, uint32_t step, float *d_grad,
                           float *loss) {
  if (!t || !d_theta || !d_grad) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(t->ctx->device));
  const int xb = t->A.xb, yb = t->A.yb;
  // zero where no transform takes part
  k_tune_map<<<dim3((xb + 255) / 256, yb), 256, 0, t->ctx->stream>>>(xb, yb, t->d_tmap, d_grad);
  k_tune_map_fill<<<dim3((xb + 255) / 256, yb), 256, 0, t->ctx->stream>>>(t->A.cells, xb, yb,
                                                                          t->d_tmap, t->d_excl);
  int rc = tuner_eval(t, d_theta, step, d_grad, t->max_steps + 1);
  if (rc) return rc;
  if (loss) {
    HIP_TRY(hipMemcpyAsync(loss, t->d_loss + t->max_steps + 1, sizeof(float),
                           hipMemcpyDeviceToHost, t->ctx->stream));
    HIP_TRY(hipStreamSynchronize(t->ctx->stream));
  }
  return JXG_OK;
}

int jxg_qf_tuner_run(jxg_qf_tuner *t, const uint8_t *d_qf_in, uint32_t steps, float lr,
                     uint8_t *d_qf_out, float *d_theta_out, float *loss_trace) {
  if (!t || !d_qf_in || !d_qf_out) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  if (steps > t->max_steps)
    return set_err(JXG_ERR_INVALID_ARG, "steps %u above the tuner's %u", steps, t->max_steps);
  if (!(lr > 0.0f)) return set_err(JXG_ERR_INVALID_ARG, "lr %g", (double)lr);
  HIP_TRY(hipSetDevice(t->ctx->device));
  jxg_ctx *ctx = t->ctx;
  const int xb = t->A.xb, yb = t->A.yb, nb = xb * yb;
  const dim3 gb((xb + 255) / 256, yb);
  const unsigned g1 = (unsigned)((nb + 255) / 256);
  k_tune_init<<<g1, 256, 0, ctx->stream>>>(d_qf_in, t->d_theta, t->d_m, t->d_v, nb);
  k_tune_map<<<gb, 256, 0, ctx->stream>>>(xb, yb, t->d_tmap, t->d_grad);
  k_tune_map_fill<<<gb, 256, 0, ctx->stream>>>(t->A.cells, xb, yb, t->d_tmap, t->d_excl);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float c1 = 1.0f, c2 = 1.0f;
  int rc;
  for (uint32_t s = 0; s < steps; s++) {
    if ((rc = tuner_eval(t, t->d_theta, s, t->d_grad, s))) return rc;
    c1 *= b1;
    c2 *= b2;
    k_tune_adam<<<g1, 256, 0, ctx->stream>>>(t->d_theta, t->d_grad, t->d_m, t->d_v, t->d_tmap, nb,
                                            lr, b1, b2, eps, 1.0f - c1, 1.0f - c2);
  }
  // the loss at the final field, with the last step's noise
  if ((rc = tuner_eval(t, t->d_theta, steps ? steps - 1 : 0, t->d_grad, steps))) return rc;
  k_tune_round<<<gb, 256, 0, ctx->stream>>>(t->d_theta, t->d_tmap, t->A.cells, d_qf_in, xb, yb,
                                           d_qf_out);
  if (d_theta_out)
    HIP_TRY(hipMemcpyAsync(d_theta_out, t->d_theta, (size_t)nb * sizeof(float),
                           hipMemcpyDeviceToDevice, ctx->stream));
  LAUNCH_CHECK("k_tune_round");
  if (loss_trace) {
    HIP_TRY(hipMemcpyAsync(loss_trace, t->d_loss, (size_t)(steps + 1) * sizeof(float),
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  return JXG_OK;
}

}  // extern "C"
