// jxg_ctx.h - host-side context and error plumbing shared by the kernel
// translation units of libjxg.
#ifndef JXG_CTX_H_
#define JXG_CTX_H_

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "jxg.h"
#include "jxg_internal.h"

#define JXG_HIDDEN __attribute__((visibility("hidden")))

// Records a message for jxg_last_error() on the calling thread; returns code.
JXG_HIDDEN int set_err(int code, const char *fmt, ...);

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return set_err(e_ == hipErrorOutOfMemory ? JXG_ERR_ALLOC : JXG_ERR_DEVICE, "%s: %s", \
                     #expr, hipGetErrorString(e_));                                      \
  } while (0)

#define LAUNCH_CHECK(name)                                                               \
  do {                                                                                   \
    hipError_t e_ = hipGetLastError();                                                   \
    if (e_ != hipSuccess)                                                                \
      return set_err(JXG_ERR_DEVICE, "launch %s: %s", name, hipGetErrorString(e_));      \
  } while (0)

struct jxg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // A second stream (non-blocking, default priority) for the CfL maps, which
  // run beside the quant field; ev_fork / ev_join order it after XYB and
  // before the strategy search on `stream`.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // Quant tables: the caller's blob (jxg_ctx_set_tables; in use while
  // tables_set) in d_tables / d_selq, and the default tables
  // (jxg_default_quant_tables), one set per (distance tier, fmla flavour) in
  // slot 2 * tier + fused, each built and uploaded on first use.
  float *d_tables = nullptr;
  bool tables_set = false;
  float *d_tier_tables = nullptr;  // 2 * JXG_QUANT_TIERS blobs
  float *d_tier_selq = nullptr;    // their scan-ordered inverse weights
  bool tier_valid[2 * JXG_QUANT_TIERS] = {};
  jxg_derived *d_derived = nullptr;
  jxg_derived *h_derived = nullptr;  // pinned staging
  bool derived_valid = false;
  jxg_frame_params last_params;
  uint32_t *d_scans = nullptr;  // natural orders of DCT8/16/32: scan position -> storage index
  uint16_t *d_scanpos = nullptr;  // storage index -> scan position, every shape up to 64x64
  uint32_t h_order[64 + 256 + 1024];
  float *d_selq = nullptr;  // inverse quant weights of DCT8/16/32 per channel, in scan order
  float *d_mlut = nullptr;      // fwd matrix (9) + transfer LUT (256)
  float *h_mlut = nullptr;
  bool mlut_valid = false;
  float *d_exp2 = nullptr;      // fast_exp2 table
  float *d_pre = nullptr;
  size_t pre_cap = 0;
  float *d_noalloc_pre = nullptr;  // device address of the static no-allocation scratch
  float *d_xyb = nullptr;
  size_t xyb_cap = 0;
  uint32_t *d_misc = nullptr;  // [0] crossing count, [1] unsupported flag (zero between
                               // calls), [4..5] their last values, [6] finish ticket
  uint32_t *d_xt = nullptr;
  size_t xt_cap = 0;
  int8_t *d_zero_cmap = nullptr;
  size_t zero_cap = 0;
  uint8_t *d_rgb = nullptr;
  size_t rgb_cap = 0;
  float *d_tmp = nullptr;  // out-of-place staging of in-place filters
  size_t tmp_cap = 0;
  int8_t *d_cmap = nullptr;  // CfL maps when the caller does not want them
  size_t cmap_cap = 0;
  void *d_c64 = nullptr;  // split-bf16 DCT64 matrix (hi, lo), built on first use
  void *d_c64x6 = nullptr;  // the same in three terms (hi, mid, lo)
  float *d_rate_big = nullptr;  // log2f(1 + i), i < JXG_RATE_BIG_N, built at creation
  void *d_c128 = nullptr;  // split-bf16 DCT128 matrix (hi, lo), built on first use
  float *d_wc128 = nullptr;  // 128-point butterfly constants, built on first use
  uint32_t *d_astat = nullptr;  // alpha-stage counters
  uint8_t *d_alpha = nullptr;   // alpha plane of the last front end call
  size_t alpha_cap = 0;
  size_t alpha_used = 0;        // its size in bytes (0: that call had no alpha)
  uint8_t *d_alpha2 = nullptr;  // fringe sweep outputs
  size_t alpha2_cap = 0;
  uint32_t *d_vis = nullptr;  // visibility bits
  size_t vis_cap = 0;
  uint32_t *d_fxchg = nullptr;  // fill_invisible hand-over rows (3 x bands x W)
  size_t fxchg_cap = 0;
  float *d_afac = nullptr;  // alpha block factors
  size_t afac_cap = 0;
  float *d_pc = nullptr;  // re-tiling piece costs (4 planes of blocks)
  size_t pc_cap = 0;
  float *d_c32 = nullptr;  // raw DCT32X32 cost per 32x32 region's first block
  size_t c32_cap = 0;
  uint8_t *d_rf = nullptr;  // quant-field refinement: candidate costs / values, transform lists
  size_t rf_cap = 0;
  // jxg_ctx_profile: event pairs per stage, recorded while prof_on
  bool prof_on = false;
  hipEvent_t *prof_ev = nullptr;  // JXG_PROF_N * 2 * JXG_PROF_CAP
  uint32_t prof_n[JXG_PROF_N] = {};
};

// Event pair around a profiled kernel (no-ops unless profiling is on).
inline void prof_begin(jxg_ctx *ctx, int k, hipStream_t st) {
  if (ctx->prof_on && ctx->prof_n[k] < JXG_PROF_CAP)
    (void)hipEventRecord(ctx->prof_ev[(k * JXG_PROF_CAP + ctx->prof_n[k]) * 2], st);
}
inline void prof_end(jxg_ctx *ctx, int k, hipStream_t st) {
  if (ctx->prof_on && ctx->prof_n[k] < JXG_PROF_CAP) {
    (void)hipEventRecord(ctx->prof_ev[(k * JXG_PROF_CAP + ctx->prof_n[k]) * 2 + 1], st);
    ctx->prof_n[k]++;
  }
}

template <typename T>
inline int ensure(T **p, size_t *cap, size_t n) {
  if (*cap >= n && *p) return JXG_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  HIP_TRY(hipMalloc((void **)p, n * sizeof(T)));
  *cap = n;
  return JXG_OK;
}

// Shared argument checks and per-frame uploads (jxg_kernels.hip).
JXG_HIDDEN int check_image(uint32_t w, uint32_t h);
JXG_HIDDEN int check_planes(uint32_t w, uint32_t h, uint32_t stride, uint32_t padded_h);
// JXG_ERR_UNSUPPORTED for speeds whose layout decisions are not built.
JXG_HIDDEN int check_speed(const jxg_frame_params *p);
JXG_HIDDEN int upload_derived(jxg_ctx *ctx, const jxg_frame_params *p);
// The quant tables (blob and scan-ordered inverse weights) the stages use for
// a frame with parameters p: the caller's set if one was given, else the
// default set of p's distance tier, built on first use (a synchronous upload).
JXG_HIDDEN int tables_for(jxg_ctx *ctx, const jxg_frame_params *p, const float **tables,
                          const float **selq);
JXG_HIDDEN int fill_cmap(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                         uint32_t width, uint32_t height, const jxg_frame_params *p,
                         int8_t *d_ytox, int8_t *d_ytob, hipStream_t st);  // jxg_cfl.hip
JXG_HIDDEN int fill_u8(jxg_ctx *ctx, uint8_t *p, size_t n, uint8_t v);
// jxg_alpha.hip: the alpha stages of 2- and 4-channel input after XYB, and
// the alpha factors on the final quant field.
JXG_HIDDEN int alpha_front(jxg_ctx *ctx, const uint8_t *d_in, size_t row_bytes, int ch,
                           uint32_t W, uint32_t H, int flags, float *xyb, uint32_t stride,
                           uint32_t ph);
JXG_HIDDEN int alpha_finish(jxg_ctx *ctx, uint32_t W, uint32_t H, const uint8_t *cells,
                            uint8_t *qf);
// Batched stage launches (jxg_kernels.hip, jxg_cfl.hip); see jxg_batch.h.
struct BatchImg;
JXG_HIDDEN int batch_stages_pre(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                                const uint32_t *d_pfx, int n, const uint32_t total[],
                                int fused, const float *mlut, int aq_strip, int xyb_order);
JXG_HIDDEN int batch_cmap(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                          const uint32_t *d_pfx, int n, const uint32_t total[], int fused,
                          int cm_rows);
JXG_HIDDEN int batch_stages_post(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                                 const uint32_t *d_pfx, int n, const uint32_t total[],
                                 int fused, uint32_t *xt_list, uint32_t *xt_img, uint32_t xt_cap,
                                 uint32_t *counters, int xt_grid);
JXG_HIDDEN void batch_grids(uint32_t W, uint32_t H, uint32_t stride, uint32_t ph, int aq_strip,
                            int cm_rows, uint32_t gw[], uint32_t gh[]);
// jxg_chroma.hip: chroma-HF protection of the quant field.
JXG_HIDDEN int chroma_hf(jxg_ctx *ctx, const float *xyb, uint32_t stride, uint32_t ph, uint32_t W,
                         uint32_t H, float distance, int fused, uint8_t *qf);
// jxg_retile.hip: the coefficient re-tiling after k_select (mode = the
// derived `retile`), then adjust_quant_field; single image and batch.
JXG_HIDDEN int retile_stages(jxg_ctx *ctx, hipStream_t st, const float *xyb, uint32_t stride,
                             uint32_t ph, uint32_t W, uint32_t H, uint8_t *qf, const int8_t *ytox,
                             const int8_t *ytob, const float *tables, const jxg_frame_params *p,
                             int mode, uint8_t *cells, const float *c32, float *pc);
// jxg_refine.hip: refine_quant_field + qf_trellis after the adjustment.
JXG_HIDDEN int refine_stages(jxg_ctx *ctx, hipStream_t st, const float *xyb, uint32_t stride,
                             uint32_t ph, uint32_t W, uint32_t H, uint8_t *qf, const int8_t *ytox,
                             const int8_t *ytob, const float *tables, const jxg_frame_params *p,
                             const uint8_t *cells);
JXG_HIDDEN int batch_retile(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                            const uint32_t *d_pfx, int n, const uint32_t total[], int fused);
JXG_HIDDEN int rerank_stages(jxg_ctx *ctx, hipStream_t st, const float *xyb, uint32_t stride,
                             uint32_t ph, uint32_t W, uint32_t H, const uint8_t *qf,
                             const int8_t *ytox, const int8_t *ytob, const float *tables,
                             const jxg_frame_params *p, uint8_t *cells);
JXG_HIDDEN int batch_rerank(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                            const uint32_t *d_pfx, int n, const uint32_t total[], int fused);
JXG_HIDDEN void rerank_grids(uint32_t W, uint32_t H, uint32_t *g_list, uint32_t *g_small,
                             uint32_t *g_big, uint32_t *g_decide);
JXG_HIDDEN int batch_refine(jxg_ctx *ctx, hipStream_t st, const BatchImg *d_img,
                            const uint32_t *d_pfx, int n, const uint32_t total[], int fused);
JXG_HIDDEN void refine_grids(uint32_t W, uint32_t H, uint32_t *g_list, uint32_t *g0, uint32_t *g1,
                             uint32_t *g_trellis);
JXG_HIDDEN int zero_cmaps(jxg_ctx *ctx, uint32_t w, uint32_t h, const int8_t **x,
                          const int8_t **b);

#endif  // JXG_CTX_H_
