// jxg_batch.hip - the batched front end: descriptors, scratch arena, the
// stage chain as one HIP graph, and the tile-shape autotuner.
//
// The stage kernels are the single-image stage bodies behind a flat-index ->
// (image, block) map (jxg_batch.h), so a batch of N images costs eight
// launches instead of 8N and small images share the device with large ones.
// Images are laid out in the map largest first, so the big workloads start
// early and the small ones fill the tail of each launch.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "jxg_batch.h"
#include "jxg_ctx.h"

struct jxg_batch {
  jxg_ctx *ctx = nullptr;
  int n = 0;
  int fused = 1;
  int xyb_order = JXG_XYB_SCALAR;
  std::vector<jxg_batch_image> images;
  std::vector<int> order;  // images in map order (largest first)
  std::vector<uint32_t> geo;  // per image: stride, ph
  std::vector<BatchImg> host;  // descriptors in map order (gw set per upload)
  BatchImg *d_img = nullptr;
  uint32_t *d_pfx = nullptr;
  jxg_derived *d_der = nullptr;
  float *d_arena = nullptr;
  float *d_mlut = nullptr;
  uint32_t *d_xt = nullptr, *d_xt_img = nullptr;
  uint32_t xt_cap = 0;
  uint32_t *d_counters = nullptr;  // [0] queue count, [1] bad, [4..6] status
  uint32_t total[BS_N] = {};
  std::vector<uint8_t> stage_on;  // per image: re-tiling, quant-field refinement, rerank
  int aq_strip = 32, cm_rows = 16, xt_grid = 256;
  hipStream_t stream = nullptr;  // the graph's stream
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void free_batch(jxg_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  (void)hipFree(b->d_img);
  (void)hipFree(b->d_pfx);
  (void)hipFree(b->d_der);
  (void)hipFree(b->d_arena);
  (void)hipFree(b->d_mlut);
  (void)hipFree(b->d_xt);
  (void)hipFree(b->d_xt_img);
  (void)hipFree(b->d_counters);
  delete b;
}

size_t round64(size_t v) { return (v + 63) & ~(size_t)63; }

// Descriptors and workgroup prefixes for the current tile shapes; uploads
// both.  Arena pointers were fixed at creation.
int upload_layout(jxg_batch *b, const std::vector<BatchImg> &host_img) {
  const int n = b->n;
  std::vector<BatchImg> img(host_img);
  std::vector<uint32_t> pfx((size_t)BS_N * (n + 1), 0);
  uint64_t tot[BS_N] = {};
  for (int k = 0; k < n; k++) {
    const int i = b->order[k];
    const jxg_batch_image &im = b->images[i];
    uint32_t gw[BS_N], gh[BS_N];
    batch_grids(im.width, im.height, b->geo[2 * i], b->geo[2 * i + 1], b->aq_strip, b->cm_rows, gw,
                gh);
    // stages this image's tier skips get no workgroups
    if (!b->stage_on[3 * i]) gw[BS_PIECE] = gw[BS_RETILE] = gw[BS_ADJUST] = 0;
    if (!b->stage_on[3 * i + 1])
      gw[BS_RLIST] = gw[BS_RCOST_S] = gw[BS_RCOST_B] = gw[BS_TRELLIS] = 0;
    if (!b->stage_on[3 * i + 2])
      gw[BS_RRLIST] = gw[BS_RRCOST_S] = gw[BS_RRCOST_B] = gw[BS_RRDECIDE] = 0;
    for (int s = 0; s < BS_N; s++) {
      img[k].gw[s] = (int)gw[s];
      pfx[(size_t)s * (n + 1) + k] = (uint32_t)tot[s];
      tot[s] += (uint64_t)gw[s] * gh[s];
    }
  }
  for (int s = 0; s < BS_N; s++) {
    if (tot[s] >= (1ull << 31)) return set_err(JXG_ERR_TOO_LARGE, "batch too large for one launch");
    pfx[(size_t)s * (n + 1) + n] = (uint32_t)tot[s];
    b->total[s] = (uint32_t)tot[s];
  }
  HIP_TRY(hipMemcpy(b->d_img, img.data(), n * sizeof(BatchImg), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b->d_pfx, pfx.data(), pfx.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  return JXG_OK;
}

int enqueue(jxg_batch *b, hipStream_t st) {
  jxg_ctx *ctx = b->ctx;
  int rc;
  if ((rc = batch_stages_pre(ctx, st, b->d_img, b->d_pfx, b->n, b->total, b->fused, b->d_mlut,
                             b->aq_strip, b->xyb_order)))
    return rc;
  if ((rc = batch_cmap(ctx, st, b->d_img, b->d_pfx, b->n, b->total, b->fused, b->cm_rows))) return rc;
  return batch_stages_post(ctx, st, b->d_img, b->d_pfx, b->n, b->total, b->fused, b->d_xt,
                           b->d_xt_img, b->xt_cap, b->d_counters, b->xt_grid);
}

}  // namespace

extern "C" {

int jxg_batch_create(jxg_ctx *ctx, uint32_t n, const jxg_batch_image *images, jxg_batch **out) {
  if (!ctx || !images || !out || n == 0) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  const int fused = images[0].params.fused;
  for (uint32_t i = 0; i < n; i++) {
    const jxg_batch_image &im = images[i];
    int rc = check_image(im.width, im.height);
    if (rc) return rc;
    if (im.channels != 1 && im.channels != 3)
      return set_err(JXG_ERR_INVALID_ARG, "image %u: batches take 1 or 3 channels", i);
    if (!im.d_rgb || !im.d_ytox || !im.d_ytob || !im.d_qf || !im.d_cells)
      return set_err(JXG_ERR_INVALID_ARG, "image %u: null buffer", i);
    for (int c = 0; c < 3; c++)
      if (!im.d_levels[c] || !im.d_coeffs[c] || !im.d_dc[c])
        return set_err(JXG_ERR_INVALID_ARG, "image %u: null output", i);
    if (im.row_bytes < (size_t)im.width * im.channels)
      return set_err(JXG_ERR_INVALID_ARG, "image %u: row_bytes too small", i);
    if (im.params.fused != fused)
      return set_err(JXG_ERR_INVALID_ARG, "image %u: every image of a batch shares params.fused", i);
    if (im.params.xyb_order != images[0].params.xyb_order ||
        memcmp(im.params.xyb_fwd, images[0].params.xyb_fwd, sizeof(im.params.xyb_fwd)) != 0)
      return set_err(JXG_ERR_INVALID_ARG,
                     "image %u: every image of a batch shares params.xyb_order and xyb_fwd", i);
    if (im.params.xyb_order != JXG_XYB_SCALAR && im.params.xyb_order != JXG_XYB_AVX2)
      return set_err(JXG_ERR_INVALID_ARG, "image %u: xyb_order %d", i, im.params.xyb_order);
    if ((rc = check_speed(&im.params))) return rc;
    if (im.params.flags & JXG_FLAG_CHROMA_HF)
      return set_err(JXG_ERR_INVALID_ARG, "image %u: chroma-HF protection is single-image only", i);
  }
  HIP_TRY(hipSetDevice(ctx->device));
  jxg_batch *b = new jxg_batch;
  b->ctx = ctx;
  b->n = (int)n;
  b->fused = fused ? 1 : 0;
  b->xyb_order = images[0].params.xyb_order;
  b->images.assign(images, images + n);
  b->order.resize(n);
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) {
    return (uint64_t)images[a].width * images[a].height > (uint64_t)images[c].width * images[c].height;
  });
  auto fail = [&](int rc) {
    free_batch(b);
    return rc;
  };
  // derived parameters, arena layout
  std::vector<jxg_derived> der(n);
  b->geo.resize(2 * n);
  b->stage_on.resize(3 * n);
  size_t arena = 0;
  std::vector<size_t> off_xyb(n), off_pre(n), off_part(n), off_pc(n), off_rf(n);
  uint64_t queue = 0;
  for (uint32_t i = 0; i < n; i++) {
    const jxg_batch_image &im = images[i];
    const int rc = jxg_derive(&im.params, &der[i]);
    if (rc) return fail(set_err(rc, "image %u: invalid frame parameters", i));
    der[i].rate_big = ctx->d_rate_big;
    b->stage_on[3 * i] = der[i].retile != 0;
    b->stage_on[3 * i + 1] = der[i].refine_steps != 0;
    b->stage_on[3 * i + 2] = der[i].rerank != 0;
    uint32_t s, ph;
    jxg_plane_geometry(im.width, im.height, &s, &ph);
    b->geo[2 * i] = s;
    b->geo[2 * i + 1] = ph;
    const size_t xb = (im.width + 7) / 8, yb = (im.height + 7) / 8;
    const size_t tiles = (size_t)((im.width + 63) / 64) * ((im.height + 63) / 64);
    off_xyb[i] = arena;
    arena += round64((size_t)3 * s * ph);
    off_pre[i] = arena;
    arena += round64(4 * xb * yb);
    off_part[i] = arena;
    arena += round64(tiles * 64 * 4);
    off_pc[i] = arena;  // re-tiling piece costs (4 planes) and 32x32 costs
    arena += round64(5 * xb * yb);
    off_rf[i] = arena;  // quant-field refinement scratch (bytes, in floats)
    arena += round64((jxg_refine_scratch_bytes(xb * yb) + 3) / 4);
    queue += xb * yb / 2 + 1;
  }
  if (queue >= (1ull << 32)) return fail(set_err(JXG_ERR_TOO_LARGE, "batch too large"));
  b->xt_cap = (uint32_t)queue;
  if (hipMalloc((void **)&b->d_img, n * sizeof(BatchImg)) != hipSuccess ||
      hipMalloc((void **)&b->d_pfx, (size_t)BS_N * (n + 1) * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void **)&b->d_der, n * sizeof(jxg_derived)) != hipSuccess ||
      hipMalloc((void **)&b->d_arena, arena * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&b->d_mlut, 265 * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&b->d_xt, (size_t)b->xt_cap * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void **)&b->d_xt_img, (size_t)b->xt_cap * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc((void **)&b->d_counters, 8 * sizeof(uint32_t)) != hipSuccess)
    return fail(set_err(JXG_ERR_ALLOC, "batch allocation failed"));
  if (hipMemcpy(b->d_der, der.data(), n * sizeof(jxg_derived), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(b->d_counters, 0, 8 * sizeof(uint32_t)) != hipSuccess)
    return fail(set_err(JXG_ERR_DEVICE, "batch upload failed"));
  {
    // the batch's XybMatrix::fwd and the sRGB transfer table
    float mlut[265];
    memcpy(mlut, images[0].params.xyb_fwd, 9 * sizeof(float));
    jxg_srgb_lut(mlut + 9);
    if (hipMemcpy(b->d_mlut, mlut, sizeof(mlut), hipMemcpyHostToDevice) != hipSuccess)
      return fail(set_err(JXG_ERR_DEVICE, "batch upload failed"));
  }
  std::vector<BatchImg> *host = &b->host;
  host->resize(n);
  for (uint32_t k = 0; k < n; k++) {
    const int i = b->order[k];
    const jxg_batch_image &im = images[i];
    BatchImg &d = (*host)[k];
    memset(&d, 0, sizeof(d));
    d.rgb = im.d_rgb;
    d.row_bytes = im.row_bytes;
    d.ch = im.channels;
    d.W = (int)im.width;
    d.H = (int)im.height;
    d.stride = (int)b->geo[2 * i];
    d.ph = (int)b->geo[2 * i + 1];
    d.xb = (d.W + 7) / 8;
    d.yb = (d.H + 7) / 8;
    d.tiles_x = (d.W + 63) / 64;
    d.tiles_y = (d.H + 63) / 64;
    d.pre_w = 2 * d.xb;
    d.pre_h = 2 * d.yb;
    d.xyb = b->d_arena + off_xyb[i];
    d.pre = b->d_arena + off_pre[i];
    d.part = b->d_arena + off_part[i];
    d.pc = b->d_arena + off_pc[i];
    d.c32 = d.pc + (size_t)4 * d.xb * d.yb;
    d.rf = (uint8_t *)(b->d_arena + off_rf[i]);
    d.qf = im.d_qf;
    d.cells = im.d_cells;
    d.ytox = im.d_ytox;
    d.ytob = im.d_ytob;
    for (int c = 0; c < 3; c++) {
      d.lv[c] = im.d_levels[c];
      d.co[c] = im.d_coeffs[c];
      d.dc[c] = im.d_dc[c];
    }
    d.D = b->d_der + i;
    {
      const int trc = tables_for(ctx, &im.params, &d.tables, &d.selq);
      if (trc) return fail(trc);
    }
    d.distance = im.params.distance;
    d.inv_scale = der[i].inv_scale;
    d.cfl_base_x = im.params.cfl_base_x;
    d.cfl_base_b = im.params.cfl_base_b;
    d.cfl_color_factor = im.params.cfl_color_factor;
  }
  int rc = upload_layout(b, *host);
  if (rc) return fail(rc);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess)
    return fail(set_err(JXG_ERR_DEVICE, "batch stream creation failed"));
  *out = b;
  return JXG_OK;
}

void jxg_batch_destroy(jxg_batch *b) { free_batch(b); }

int jxg_batch_run(jxg_batch *b) {
  if (!b) return set_err(JXG_ERR_INVALID_ARG, "null batch");
  HIP_TRY(hipSetDevice(b->ctx->device));
  return enqueue(b, b->ctx->stream);
}

int jxg_batch_launch(jxg_batch *b) {
  if (!b) return set_err(JXG_ERR_INVALID_ARG, "null batch");
  HIP_TRY(hipSetDevice(b->ctx->device));
  if (!b->exec) {
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(b, b->stream);
    const hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return set_err(JXG_ERR_DEVICE, "graph capture: %s", hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      b->exec = nullptr;
      return set_err(JXG_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(ei));
    }
  }
  // ordered after the context stream's prior work and before its later work
  HIP_TRY(hipEventRecord(b->ev_in, b->ctx->stream));
  HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_in, 0));
  HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  HIP_TRY(hipEventRecord(b->ev_out, b->stream));
  HIP_TRY(hipStreamWaitEvent(b->ctx->stream, b->ev_out, 0));
  return JXG_OK;
}

int jxg_batch_status(jxg_batch *b, uint32_t out[2]) {
  if (!b || !out) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(b->ctx->device));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  uint32_t c[8];
  HIP_TRY(hipMemcpy(c, b->d_counters, sizeof(c), hipMemcpyDeviceToHost));
  out[0] = c[4];
  out[1] = c[5];
  return JXG_OK;
}

int jxg_batch_config(jxg_batch *b, int out[3]) {
  if (!b || !out) return set_err(JXG_ERR_INVALID_ARG, "null argument");
  out[0] = b->aq_strip;
  out[1] = b->cm_rows;
  out[2] = b->xt_grid;
  return JXG_OK;
}

int jxg_batch_autotune(jxg_batch *b, int reps, char *report, size_t report_len) {
  if (!b || reps < 1) return set_err(JXG_ERR_INVALID_ARG, "bad argument");
  HIP_TRY(hipSetDevice(b->ctx->device));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  static const int strips[3] = {16, 32, 64}, rows[2] = {16, 32}, grids[3] = {64, 256, 1024};
  hipEvent_t t0, t1;
  HIP_TRY(hipEventCreate(&t0));
  HIP_TRY(hipEventCreate(&t1));
  float best = INFINITY;
  int best_cfg[3] = {b->aq_strip, b->cm_rows, b->xt_grid};
  size_t used = 0;
  if (report && report_len) report[0] = 0;
  int rc = JXG_OK;
  for (int si = 0; si < 3 && !rc; si++)
    for (int ri = 0; ri < 2 && !rc; ri++)
      for (int gi = 0; gi < 3 && !rc; gi++) {
        b->aq_strip = strips[si];
        b->cm_rows = rows[ri];
        b->xt_grid = grids[gi];
        if ((rc = upload_layout(b, b->host))) break;
        if (b->exec) {
          (void)hipGraphExecDestroy(b->exec);
          b->exec = nullptr;
        }
        if ((rc = jxg_batch_launch(b))) break;  // capture + warm-up
        if (hipStreamSynchronize(b->stream) != hipSuccess) {
          rc = set_err(JXG_ERR_DEVICE, "autotune warm-up failed");
          break;
        }
        (void)hipEventRecord(t0, b->stream);
        for (int r = 0; r < reps && !rc; r++) rc = jxg_batch_launch(b);
        (void)hipEventRecord(t1, b->stream);
        if (rc || hipEventSynchronize(t1) != hipSuccess) {
          rc = rc ? rc : set_err(JXG_ERR_DEVICE, "autotune timing failed");
          break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        ms /= (float)reps;
        if (report && used + 1 < report_len)
          used += (size_t)snprintf(report + used, report_len - used,
                                   "aq_strip=%d cm_rows=%d xt_grid=%d %.4f ms\n", strips[si],
                                   rows[ri], grids[gi], ms);
        if (ms < best) {
          best = ms;
          best_cfg[0] = strips[si];
          best_cfg[1] = rows[ri];
          best_cfg[2] = grids[gi];
        }
        if (used >= report_len) used = report_len ? report_len - 1 : 0;
      }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  b->aq_strip = best_cfg[0];
  b->cm_rows = best_cfg[1];
  b->xt_grid = best_cfg[2];
  const int rc2 = upload_layout(b, b->host);
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  return rc ? rc : rc2;
}

}  // extern "C"

