/*
 * jxo.h - CPU oracle of the VarDCT encoder front end.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * encoder's scalar CPU path (awxkee/jixel 0.5.14, Rust; no Rust toolchain
 * exists in this environment, so the reference itself cannot run).  Only the
 * tests, the smoke check and the benchmark's CPU-baseline leg use it, as the
 * checker; the product library never links it.
 *
 * Every function follows the operation order of the reference function it
 * cites.  The reference's `fmla` (src/dct.rs:33-53) is a fused multiply-add on
 * aarch64 and FMA-enabled x86 builds and a*b+c elsewhere; jxo_set_fused()
 * selects which.  Built with -ffp-contract=off so no other operation fuses.
 *
 * Images are planar f32 of W x H with stride W (the reference's Image3F,
 * src/image.rs:89-93); footprints past the image are gathered with edge
 * replication as the reference does (src/ac_strategy.rs:451-479).
 */
#ifndef JXO_H_
#define JXO_H_

#include <stddef.h>
#include <stdint.h>

#include "../include/jxg.h"

#ifdef __cplusplus
extern "C" {
#endif

void jxo_set_fused(int fused);
int jxo_get_fused(void);

/* ---- scalar helpers (pinned by the reference's known-answer tests) ---- */
float jxo_cbrtf(float x);                       /* src/xyb.rs:152 */
float jxo_fast_exp2(float d);                   /* src/adaptive_quant.rs:739 */
float jxo_dirty_log2f(float d);                 /* src/adaptive_quant.rs:768 */
float jxo_fast_exp(float x);                    /* src/util.rs:342 */
float jxo_distance_from_quality(float quality); /* src/encode_image.rs:1321 */
void jxo_srgb_lut(float out[256]);              /* src/color.rs:35 */

/* compute_distance_params (src/frame/quantization.rs:89-149). */
void jxo_distance_params(float distance, int32_t *global_scale, float *scale,
                         int32_t *quant_dc, float *scale_dc, uint32_t *x_qm_scale,
                         uint32_t *epf_iters, float dc_step[3]);

/* ---- quant tables ---- */
/* The stored library tables DEQUANT_MATRIX_8X8 (3 x 64) and
 * DEQUANT_MATRIX_16X8 (3 x 128) (src/quant_weights.rs:246,455); must be set
 * before jxo_fast_tables. */
void jxo_set_static_tables(const float *m8x8, const float *m16x8);
/* DequantMatrices::new_fast(distance, arbiter) (src/quant_weights.rs:2294)
 * as a blob in the layout of include/jxg.h, with the current fmla flavour.
 * Returns 0, or -1 when the static tables are not set. */
int jxo_fast_tables(float distance, int arbiter, float *blob);

/* ---- XYB ---- */
/* One pixel, scalar order (src/xyb.rs:172) or AVX2 production order
 * (src/avx/xyb.rs:234, SCALAR_ORDER=false, FUSED=true). */
void jxo_xyb_pixel(const float m[9], int order, float r, float g, float b, float out[3]);
/* u8 RGB(A) -> XYB planes (3 * W * H floats, stride W). */
void jxo_xyb_image(const uint8_t *rgb, size_t row_bytes, int channels, int width,
                   int height, const float m[9], int order, float *xyb);

/* ---- adaptive quantization ---- */
/* fill_quant_field_scalar for one DC group (src/adaptive_quant.rs:547-729). */
void jxo_fill_quant_field_group(const float *xyb, int width, int height, int x0, int y0,
                                int xblocks, int yblocks, float distance, float inv_scale,
                                uint8_t *qf /* xblocks*yblocks, stride xblocks */);
/* All DC groups (2048 px) into one xblocks_total-wide field. */
void jxo_fill_quant_field_image(const float *xyb, int width, int height, float distance,
                                float inv_scale, uint8_t *qf);
/* The same field before scaling and rounding: the per-block value
 * fast_exp2(out * log2 e) * mul + add of src/adaptive_quant.rs:724. */
void jxo_quant_field_float_image(const float *xyb, int width, int height, float distance,
                                 float *fq);

/* ---- transforms ---- */
/* Forward transform of `strategy` at pixel (px, py) of one plane
 * (forward_transform, src/ac_strategy.rs:500); returns size (cx*cy*64). */
int jxo_forward_transform(int strategy, const float *plane, int width, int height, int px,
                          int py, float *out);
/* DC values of one channel's coefficients (dc_from_coefficients,
 * src/group.rs:183). */
void jxo_dc_from_coefficients(int strategy, const float *coeffs, float *dc);
int jxo_covered_x(int strategy);
/* The DCT64 butterfly constants the oracle uses (32 floats). */
const float *jxo_wc64(void);
/* The AFV basis the oracle uses (16 x 16, row j = pixel j). */
const float *jxo_afv_basis(void);
int jxo_covered_y(int strategy);
/* The 128-point butterfly constants the oracle uses (64 floats). */
const float *jxo_wc128(void);
/* 128-family forward transforms: kind 0 128x128, 1 128 tall x 64 wide,
 * 2 64 tall x 128 wide; in = footprint with stride = its width. */
void jxo_dct128(int kind, const float *in, float *out);
/* decision pieces of the strategy search (src/ac_strategy/selection.rs) */
float jxo_partition_cost(const float children[2][2], const float merges[5], int mask); /* :91 */
int jxo_choose_merge_partition(const float children[2][2], const float baseline[2][2],
                               const float merges[5], const float accept[5]);        /* :115 */
int jxo_choose_square32(const float children[2][2], const float baseline[2][2], float merge,
                        float accept);
int jxo_aggregate_quant(int max_q, uint32_t sum, int covered, float distance);     /* :484 */
float jxo_aggregate_qac_2x2(const float qac[2][2], float scale, float distance);   /* :495 */
void jxo_cmap_factors(int ytox, int ytob, const jxg_frame_params *p, float out[3]); /* ac_strategy.rs:1841 */

/* ---- quantization ---- */
void jxo_ac_thresholds(int c, int xsize, int ysize, float distance, float qm_mult,
                       float thr[4]); /* src/group.rs:420 */
void jxo_quantize_block_ac(const float *in, int c, const float *qm, int32_t quant,
                           float scale, float qm_mult, float distance, int xsize,
                           int ysize, int32_t *out); /* src/group.rs:560 */
float jxo_dequant_level(float one, float numerator, int32_t q); /* src/group.rs:770 */

/* First-pass transform + quantization of every first block
 * (write_ac_group transform half, src/group.rs:1259-2240, no RDOQ). */
void jxo_forward_quantize(const float *xyb, int width, int height, const uint8_t *cells,
                          const uint8_t *qf, const int8_t *ytox, const int8_t *ytob,
                          const float *tables, const jxg_frame_params *p, int32_t *levels[3],
                          float *coeffs[3], int16_t *dc[3]);

/* ---- strategy search ---- */
/* Natural scan position of every raw index of a w x h coefficient block
 * (coeff_order::scan_pos_lut, src/coeff_order.rs:358). */
void jxo_scan_pos_lut(int w, int h, uint32_t *lut);
/* Natural order (scan position -> raw index) of a cx>=cy block
 * (natural_scan_generic, src/coeff_order.rs:206). */
void jxo_natural_scan(int cx, int cy, uint32_t *out);
/* strategy_cost with the coefficient model (src/ac_strategy.rs:724,1293). */
float jxo_strategy_cost(const float *xyb, int width, int height, int strategy, int px, int py,
                        float qac, float meta_r, const float cmap[3], const float *tables,
                        const jxg_frame_params *p);
/* The Fast / Fastest strategy decision for the whole image, per DC group:
 * select_band's square merges (src/ac_strategy/selection.rs:1000), the
 * coefficient-model re-tiling (layout_dp::retile_coefficient,
 * src/ac_strategy/selection/layout_dp.rs:2381) and adjust_quant_field
 * (selection.rs:503).  Updates qf in place and writes strategy cells. */
void jxo_select_strategy(const float *xyb, int width, int height, uint8_t *qf,
                         const int8_t *ytox, const int8_t *ytob, const float *tables,
                         const jxg_frame_params *p, uint8_t *cells);
/* plan_small (src/ac_strategy/selection/layout_dp.rs:1918): exact tiling of
 * a w x h (<= 8x8) cell grid by leaves, 2x1 / 1x2 pairs and 2x2 squares
 * around the `fixed` cells (bit y*w+x), after the live-mask pruning.
 * Returns the optimum; when finite and below replace_below, the plan's
 * pieces (cell index, choice 1 leaf / 2 wide / 3 tall / 4 square), last
 * cell first. */
float jxo_plan_small(const float *leaf, const float *wide, const float *tall, const float *square,
                     uint64_t fixed, int w, int h, float replace_below, int *n_out,
                     uint8_t *cells_out, uint8_t *choices_out);
void jxo_live_masks(const float *leaf, const float *wide, const float *tall, const float *square,
                    int w, int h, uint8_t *live);
/* One coefficient re-tiling pass (crossing 0: the aligned 64-px grid; 1:
 * the grid shifted by (4, 4) blocks) over the given layout, in place. */
void jxo_retile_pass(const float *xyb, int width, int height, const uint8_t *qf,
                     const int8_t *ytox, const int8_t *ytob, const float *tables,
                     const jxg_frame_params *p, uint8_t *cells, int crossing);
/* best_reconstruction_plan over nt tiles (x, y, w, h, cost as 5 floats
 * each) of a width x height footprint: the plan's cost, and per grid cell
 * (y * 4 + x) the index of the tile whose origin it is (0xFF: none). */
float jxo_rerank_plan(const float *tile_xywh_cost, int nt, int width, int height,
                      uint8_t *tiles_out);
/* Tests only: turn the rerank off (0) or back on (1) in jxo_select_strategy. */
void jxo_rerank_enable(int on);
/* The SSIM rerank's downgrade pass (Fast, d <= 1.709) over a given layout
 * and field (the field before adjust_quant_field), in place. */
void jxo_rerank_pass(const float *xyb, int width, int height, const uint8_t *qf,
                     const int8_t *ytox, const int8_t *ytob, const float *tables,
                     const jxg_frame_params *p, uint8_t *cells);
/* Square-candidate cost maps (DCT8 per block, DCT16X16 per 2x2 site,
 * DCT32X32 per 4x4 site; NaN where not evaluated). */
void jxo_strategy_costs(const float *xyb, int width, int height, const uint8_t *qf,
                        const int8_t *ytox, const int8_t *ytob, const float *tables,
                        const jxg_frame_params *p, float *cost8, float *cost16,
                        float *cost32);

/* ---- chroma from luma ---- */
/* One tile's (ytox, ytob) by the closed-form regression
 * (compute_cmap_tile, src/color_correlation.rs:441). */
void jxo_cmap_tile(const float *xyb, int width, int height, int tx, int ty, const float *tables,
                   const jxg_frame_params *p, int *ytox, int *ytob);
/* Both maps for the whole image, tiles_x = ceil(W/64) per row
 * (fill_cmap, src/color_correlation.rs:901, regression path). */
void jxo_fill_cmap(const float *xyb, int width, int height, const float *tables,
                   const jxg_frame_params *p, int8_t *ytox, int8_t *ytob);

/* ---- whole-plane filters ---- */
/* Normalised gaborish weights {centre, r, d, R, L, D} for `mul`
 * (src/gaborish.rs:44-70). */
void jxo_gaborish_weights(float mul, float w[6]);
/* gaborish_inverse on three W x H planes (stride W), in place
 * (src/gaborish.rs:44-77). */
void jxo_gaborish_inverse(float *xyb, int width, int height, float mul);

/* ---- EPF luma proxy (LumaProxy, src/epf_proxy.rs:151-245) ---- */
typedef struct jxo_epf jxo_epf;
/* Reference scales of a W x H luma plane (W, H >= 4). */
jxo_epf *jxo_epf_new(const float *y, int w, int h);
void jxo_epf_free(jxo_epf *p);
/* Six terms and the weighted score of a candidate plane; pooled sums in
 * pixel order in f32 (wide = 0, the reference) or in f64 (wide = 1). */
float jxo_epf_compare(const jxo_epf *p, const float *y, int wide, float terms[6]);
/* The 5x5 mirrored box mean (original_box5, src/epf_proxy/box_filter.rs:466). */
void jxo_epf_box5(const float *src, int w, int h, float *dst);

/* ---- chroma-HF protection (jxo_pixel.c) ---- */
/* chroma_hf_stats_scalar (src/adaptive_quant.rs:222): {chroma_sum,
 * chroma_grad, luma_grad} and the edge count of a w x h block at (px, py). */
void jxo_chroma_hf_stats(const float *xyb, int width, int height, int px, int py, int w, int h,
                         float out[3], int *edges);
/* The distance strength of apply_chroma_hf_protection (0 outside 0.35..2). */
float jxo_chroma_hf_strength(float distance);
/* apply_chroma_hf_protection (src/adaptive_quant.rs:303-408) on the whole
 * image's quant field (ceil(W/8) x ceil(H/8)). */
void jxo_chroma_hf_protection(const float *xyb, int width, int height, uint8_t *qf,
                              float distance);

/* ---- quant-field tuning objective (jxo_adjoint.c) ---- */
/* The tuning noise u in [-1/2, 1/2) of (seed, step, channel, first block,
 * storage index). */
float jxo_tune_noise(uint32_t seed, uint32_t step, int c, uint32_t first_block, int k);
int jxo_tune_supported(int strategy);
/* FP64 loss and per-block gradient dL/dtheta of the tuning objective. */
double jxo_tune_loss_grad(const float *xyb, int width, int height, const uint8_t *cells,
                          const double *theta, const float *tables, const jxg_frame_params *p,
                          double lambda, uint32_t seed, uint32_t step, double *grad);

/* ---- alpha (jxo_alpha.c) ---- */
/* alpha_fringe::clear for u8 samples (src/alpha_fringe.rs:77-225). */
void jxo_alpha_fringe_clear(uint8_t *alpha, int width, int height);
/* fill_invisible on one plane (src/invisible.rs:34-146); alpha is W x H. */
void jxo_fill_invisible_plane(float *plane, size_t stride, int width, int height,
                              const uint8_t *alpha);
int jxo_alpha_is_constant(const uint8_t *alpha, size_t n);
/* alpha_block_factors (src/adaptive_quant.rs:953-1030): ceil(W/8) x ceil(H/8). */
void jxo_alpha_block_factors(const uint8_t *alpha, int width, int height, float *factors);
/* apply_alpha_factors (src/adaptive_quant.rs:1036-1070) over a whole image. */
void jxo_apply_alpha_factors(uint8_t *qf, const uint8_t *cells, int xb, int yb,
                             const float *factors);

/* Whole front end on the CPU (XYB -> AQ -> CfL maps -> selection ->
 * quantize) with `threads` worker threads; used as the benchmark's CPU
 * baseline.  ytox / ytob receive the CfL maps (ceil(H/64) x ceil(W/64)).
 * With alpha (2 or 4 channels) and p->flags: JXG_FLAG_LOSSY_ALPHA clears
 * the alpha fringe, JXG_FLAG_FILL_INVISIBLE fills the colour under alpha 0
 * and scales the quant field by the alpha block factors; `alpha_out`
 * (W x H, may be NULL) receives the alpha plane the stages used. */
void jxo_front_end(const uint8_t *rgb, size_t row_bytes, int channels, int width, int height,
                   const jxg_frame_params *p, const float *tables, int8_t *ytox, int8_t *ytob,
                   int threads, uint8_t *qf, uint8_t *cells, int32_t *levels[3],
                   float *coeffs[3], int16_t *dc[3], uint8_t *alpha_out);

#ifdef __cplusplus
}
#endif

#endif /* JXO_H_ */
