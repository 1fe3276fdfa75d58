/*
 * jxg.h - C ABI of the MI355X (gfx950) VarDCT encoder front end.
 *
 * The front end is the data-parallel middle of jixel's lossy encoder
 * (awxkee/jixel 0.5.14): sRGB u8 -> linear -> XYB, the adaptive quant field,
 * the AC-strategy (transform layout) search and forward DCT + quantization
 * into per-block integer coefficient buffers.  Entropy coding, RDOQ, MA trees,
 * coefficient orders and the bit writer stay in the caller.
 *
 * Every entry point replaces one stage call of the reference's per-encode
 * dispatch table (src/encoding_context.rs:224-294) at stage granularity, one
 * call per image instead of one call per 8x8 block.  The replaced reference
 * function is cited on each declaration.
 *
 * Conventions
 *   - Plain C types only.  All functions return int: JXG_OK (0) or a negative
 *     JXG_ERR_* code; jxg_last_error() gives a message for the calling thread.
 *   - *_dev functions take device pointers and run on the context's stream;
 *     they are asynchronous unless stated.  Host-pointer functions copy in,
 *     run and copy out synchronously.
 *   - Images are planar f32 XYB planes of `stride` floats per row and
 *     `padded_h` rows (see jxg_plane_geometry).  Pixels past the image are
 *     filled by edge replication (columns repeat the row's last pixel, rows
 *     repeat the last row), which is exactly the reference's footprint gather
 *     (src/ac_strategy.rs:451-479), so transforms never need bounds checks.
 *   - Block maps (quant field, strategy cells, DC) are row-major with
 *     xblocks = ceil(W/8) columns and yblocks = ceil(H/8) rows.
 *   - No float atomics: results are independent of launch shape.
 */
#ifndef JXG_H_
#define JXG_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mapped by a caller onto jixel's EncodeError,
 *      src/util.rs:129-160) ---- */
#define JXG_OK 0
#define JXG_ERR_INVALID_ARG (-1)      /* bad pointer / stride / enum      */
#define JXG_ERR_EMPTY_IMAGE (-2)      /* EncodeError::EmptyImage          */
#define JXG_ERR_TOO_LARGE (-3)        /* EncodeError::DimensionTooLarge   */
#define JXG_ERR_INVALID_DISTANCE (-4) /* EncodeError::InvalidDistance     */
#define JXG_ERR_ALLOC (-5)            /* EncodeError::AllocationFailed    */
#define JXG_ERR_DEVICE (-6)           /* HIP runtime / launch failure     */
#define JXG_ERR_NO_DEVICE (-7)        /* no gfx950 device visible         */
#define JXG_ERR_UNSUPPORTED (-8)      /* a speed tier whose transform-layout
                                         decisions this library does not run
                                         (MEDIUM and slower)              */

/* Largest accepted image side (the format's 2^30 limit is far above what a
 * single HBM-resident image needs; this bound keeps block counts in int32). */
#define JXG_MAX_DIM (1u << 18)

/* ---- raw AC strategy codes (src/dc_group_data.rs:32-56) ---- */
enum {
  JXG_DCT = 0,
  JXG_DCT16X8 = 1, /* 8 px wide, 16 px tall */
  JXG_DCT8X16 = 2, /* 16 px wide, 8 px tall */
  JXG_DCT16X16 = 3,
  JXG_DCT32X32 = 4,
  JXG_DCT4X4 = 5,
  JXG_DCT4X8 = 6,
  JXG_DCT8X4 = 7,
  JXG_DCT32X16 = 8, /* 16 wide, 32 tall */
  JXG_DCT16X32 = 9, /* 32 wide, 16 tall */
  JXG_AFV0 = 10,
  JXG_AFV1 = 11,
  JXG_AFV2 = 12,
  JXG_AFV3 = 13,
  JXG_DCT64X64 = 14,
  JXG_DCT64X32 = 15, /* 32 wide, 64 tall */
  JXG_DCT32X64 = 16, /* 64 wide, 32 tall */
  JXG_IDENTITY = 17,
  JXG_DCT2X2 = 18,
  JXG_DCT128X128 = 19,
  JXG_DCT128X64 = 20,
  JXG_DCT64X128 = 21,
  JXG_DCT32X8 = 22, /* 8 wide, 32 tall */
  JXG_DCT8X32 = 23, /* 32 wide, 8 tall */
  JXG_NUM_STRATEGIES = 24
};

/* A strategy cell is (raw_strategy << 1) | first_block_bit, the reference's
 * AcStrategyImage encoding (src/dc_group_data.rs:85-170). */
#define JXG_CELL(strategy, first) ((uint8_t)(((strategy) << 1) | ((first) ? 1 : 0)))

/* ---- quant-matrix blob ----
 * One float array holding, for every table kind k, matrix[3][n_k] (step
 * sizes, DequantMatrices::matrix*) followed by inv[3][n_k] (reciprocals with
 * inv[c][0] = 0, DequantMatrices::inv_matrix*), kinds in the order below.
 * This is the data a caller hands over from its DequantMatrices
 * (src/quant_weights.rs:864-925) so the device uses exactly its tables. */
enum {
  JXG_QT_DCT8 = 0,     /* 64   */
  JXG_QT_DCT16X8 = 1,  /* 128, shared by 16X8 / 8X16, 8 rows x 16 cols */
  JXG_QT_DCT16X16 = 2, /* 256  */
  JXG_QT_DCT32X32 = 3, /* 1024 */
  JXG_QT_DCT32X16 = 4, /* 512, shared, 16 rows x 32 cols */
  JXG_QT_DCT8X32 = 5,  /* 256, shared, 8 rows x 32 cols */
  JXG_QT_DCT64X64 = 6, /* 4096 */
  JXG_QT_DCT64X32 = 7, /* 2048, shared, 32 rows x 64 cols */
  JXG_QT_IDENTITY = 8, /* 64 */
  JXG_QT_DCT2X2 = 9,   /* 64 */
  JXG_QT_DCT4X4 = 10,  /* 64 */
  JXG_QT_DCT4X8 = 11,  /* 64 */
  JXG_QT_AFV = 12,     /* 64 */
  JXG_QT_NUM_KINDS = 13
};
/* Total floats in a quant-matrix blob: 6 * sum(n_k). */
#define JXG_QT_BLOB_FLOATS 52224u

/* Element count and blob offset of table kind k; matrix of channel c starts
 * at offset + c * n, its inverse at offset + (3 + c) * n. */
uint32_t jxg_qt_size(int kind);
uint32_t jxg_qt_offset(int kind);
/* Table kind used by a raw strategy, or -1 (src/group.rs:1459-1495). */
int jxg_qt_kind_of(int strategy);

/* ---- frame parameters ----
 * The scalars the reference reads from DistanceParams
 * (src/frame/quantization.rs:36-50) and EncodingContext. */
typedef struct jxg_frame_params {
  float distance;
  float scale;    /* global_scale / 65536 */
  float scale_dc; /* quant_dc * scale */
  float dc_step[3];
  int32_t ytob_dc;
  float cfl_base_x;       /* CflFrame::base_x (0 by default)   */
  float cfl_base_b;       /* CflFrame::base_b (1 by default)   */
  float cfl_color_factor; /* CflFrame::color_factor (84)       */
  uint32_t x_qm_scale;    /* X AC multiplier 1.25^(x_qm_scale-2) */
  uint32_t b_qm_scale;    /* B AC multiplier table index, 2..7   */
  float dequant_one[3];   /* DequantBias::one per channel        */
  float dequant_numerator;
  int32_t fused;  /* 1: the reference's fmla is a fused multiply-add (aarch64,
                     FMA-enabled x86 builds); 0: a*b+c (generic x86-64)    */
  int32_t speed;  /* JXG_SPEED_* */
  int32_t global_scale;
  int32_t quant_dc;
  int32_t epf_iters;
  int32_t flags;  /* JXG_FLAG_* stage switches */
  float channel_weight[3]; /* EncodingContext::channel_weights (X, Y, B) */
  int32_t arbiter; /* JXG_ARBITER_*: selects the default quant tables */
  int32_t xyb_order;  /* JXG_XYB_SCALAR or JXG_XYB_AVX2: the front end's XYB flavour */
  float xyb_fwd[9];   /* XybMatrix::fwd of the frame (src/xyb.rs:81-96; the
                         yellow-opsin variants of src/yellow_opsin.rs:724
                         change it), row-major */
} jxg_frame_params;

/* jxg_frame_params::flags.  The first two are on after
 * jxg_frame_params_init, as in EncodeConfig's defaults
 * (src/encode_image.rs:682-740), and act only on 2- and 4-channel input
 * (gray + alpha, RGBA):
 *   JXG_FLAG_LOSSY_ALPHA     alpha_fringe::clear on the alpha plane first
 *                            (src/alpha_fringe.rs:77; the reference runs it
 *                            unless alpha_distance <= 0, src/encode_image.rs:3650)
 *   JXG_FLAG_FILL_INVISIBLE  fill_invisible on XYB (src/invisible.rs:34,
 *                            src/frame/encoder.rs:152) and, for non-constant
 *                            alpha, the alpha block factors on the final quant
 *                            field (src/frame/encoder.rs:60-69,
 *                            src/frame/dc_groups.rs:305). */
enum { JXG_FLAG_FILL_INVISIBLE = 1, JXG_FLAG_LOSSY_ALPHA = 2, JXG_FLAG_CHROMA_HF = 4 };
/* JXG_FLAG_CHROMA_HF (off after jxg_frame_params_init): apply the chroma-HF
 * protection to the AQ field, as setup_dc_group does when the effort has the
 * chroma policy and the frame classifier marked it x- or b-heavy
 * (src/frame/dc_groups.rs:159, flags set in src/frame/chroma.rs:442-443; the
 * classifier stays with the caller). */

/* Speed tiers (src/encode_image.rs Speed, src/effort.rs:305-441). */
enum {
  JXG_SPEED_ULTRAFAST = 0,
  JXG_SPEED_FASTEST = 1,
  JXG_SPEED_FAST = 2,
  JXG_SPEED_MEDIUM = 3,
  JXG_SPEED_MEDIUMSLOW = 4,
  JXG_SPEED_SLOW = 5,
  JXG_SPEED_EXTRASLOW = 6
};

/* XYB arithmetic flavour: the scalar kernel (src/xyb.rs:146-185) or the
 * x86-64 AVX2 production kernel (src/avx/xyb.rs:234-262: fast cube-root seed
 * and x*(num/den) Halley step, always fused). */
enum { JXG_XYB_SCALAR = 0, JXG_XYB_AVX2 = 1 };

/* ---- host-side parameter helpers (no device needed) ---- */

/* distance_from_quality (src/encode_image.rs:1321-1337). */
float jxg_distance_from_quality(float quality);

/* Fills every field of *out for `distance` and `speed` exactly as
 * compute_distance_params (src/frame/quantization.rs:89-149) and the
 * EncodingContext defaults (src/encoding_context.rs:609-700) do; `fused`
 * selects the fmla semantics.  Returns JXG_ERR_INVALID_DISTANCE for a
 * non-finite or non-positive distance. */
int jxg_frame_params_init(float distance, int speed, int fused, jxg_frame_params *out);

/* XybMatrix::SPEC's forward matrix (src/xyb.rs:35-47,81-96), row-major. */
void jxg_spec_xyb_matrix(float fwd[9]);

/* The 256-entry u8 sRGB -> linear table of TransferLut::build
 * (src/color.rs:35-45,295-340): computed in f64, stored as f32. */
void jxg_srgb_lut(float out[256]);

/* Arbiter: the metric the reference's lossy encoder tunes for
 * (Arbiter, src/encode_image.rs:292; per-arbiter tuning src/arbiter.rs:36-62).
 * The default is SSIMULACRA2 (src/encode_image.rs:717). */
enum { JXG_ARBITER_SSIMULACRA2 = 0, JXG_ARBITER_CVVDP = 1 };

/* The quant-matrix blob the reference codes with at its default effort:
 * DequantMatrices::new_fast(distance, arbiter) (src/quant_weights.rs:2294-2337),
 * the table set EncodingContext::new picks for speeds without rectangles
 * (src/encoding_context.rs:621-625).  Five distance tiers: the spec tables
 * below d = 0.3, the flat-B DCT8 rows FLAT_B8_BANDS_HQ / _MID
 * (src/quant_weights.rs:129-130) above it, the SS2 16 / 32 / 32x16 scales from
 * qm_ss2_distance (2.25, CVVDP 1.75) and the coarse DCT8 luma tilt from 3.5;
 * every tier carries the spec IDENTITY / DCT2X2 / DCT4X4 / AFV / DCT64
 * tables, the stored DCT16X8 table, the refitted DCT4X8 table and the
 * DCT8X32 table derived from the DCT16X8 bands (shared_tables, tables_4x8,
 * derive_8x32: src/quant_weights.rs:1729-1836,1986-2064).  `fused` is the
 * fmla flavour of the table formulas, as in jxg_frame_params.  The device
 * stages use these tables by default, chosen per frame from
 * jxg_frame_params::distance / arbiter / fused; jxg_ctx_set_tables replaces
 * them with a caller's DequantMatrices. */
int jxg_default_quant_tables(float distance, int arbiter, int fused,
                             float *blob /* JXG_QT_BLOB_FLOATS */);

/* Plane geometry used by the device stages for a W x H image: rows are
 * padded to a multiple of 64 pixels and the stride to a multiple of 64 floats
 * (256 B) so every 64x64 tile is whole and every row load is aligned. */
void jxg_plane_geometry(uint32_t width, uint32_t height, uint32_t *stride,
                        uint32_t *padded_h);

/* ---- context ---- */
typedef struct jxg_ctx jxg_ctx;

/* Creates a context on HIP device `device`, launching on `stream`
 * (a hipStream_t, or NULL for the null stream). */
int jxg_ctx_create(int device, void *stream, jxg_ctx **out);
void jxg_ctx_destroy(jxg_ctx *ctx);
int jxg_ctx_synchronize(jxg_ctx *ctx);
const char *jxg_last_error(void);

/* Per-kernel timing of the front end for measurement: with profiling on,
 * jxg_front_end_dev records a pair of HIP events around each of its kernels
 * (or short kernel chains) on the stream that kernel runs on, for up to
 * JXG_PROF_CAP calls.  jxg_ctx_profile(ctx, 1) resets and turns it on, 0
 * turns it off.  jxg_ctx_profile_read synchronizes and returns, per stage,
 * the summed milliseconds and the number of timed launches. */
enum {
  JXG_PROF_XYB = 0,        /* k_xyb                                    */
  JXG_PROF_ALPHA = 1,      /* alpha fringe + fill_invisible (2/4 ch)   */
  JXG_PROF_CMAP = 2,       /* k_cmap_blocks + k_cmap_solve (side stream) */
  JXG_PROF_AQ_PRE = 3,     /* k_aq_pre                                 */
  JXG_PROF_AQ_BLOCK = 4,   /* k_aq_block                               */
  JXG_PROF_SELECT = 5,     /* k_select                                 */
  JXG_PROF_ALPHA_QF = 6,   /* alpha block factors on the final field   */
  JXG_PROF_QUANTIZE = 7,   /* k_quantize                               */
  JXG_PROF_QUANTIZE_XT = 8,/* k_quantize_xt                            */
  JXG_PROF_FILL = 9,       /* k_fill_bands alone (inside JXG_PROF_ALPHA) */
  JXG_PROF_RETILE = 10,    /* piece costs, re-tiling passes, field adjustment */
  JXG_PROF_REFINE = 11,    /* quant-field refinement: candidate costs, QF trellis */
  JXG_PROF_RERANK = 12,    /* SSIM rerank (inside JXG_PROF_RETILE)     */
  JXG_PROF_N = 13
};
#define JXG_PROF_CAP 64
int jxg_ctx_profile(jxg_ctx *ctx, int enable);
int jxg_ctx_profile_read(jxg_ctx *ctx, double total_ms[JXG_PROF_N], uint32_t count[JXG_PROF_N]);

/* Uploads a quant-matrix blob (host pointer) into the context; the CfL,
 * strategy and quantize stages use it for every frame until it is replaced,
 * instead of the per-frame default tables (jxg_default_quant_tables).  A NULL
 * blob returns to the defaults. */
int jxg_ctx_set_tables(jxg_ctx *ctx, const float *blob);

/* ---- device stages ---- */

/* u8 sRGB (3 or 4 channels, alpha ignored; 1 or 2 channels = gray (+ alpha),
 * decoded to equal R = G = B as linearize_gray does, src/encode_image.rs:1571,
 * and snapped to X = 0, B = Y as snap_achromatic_xyb does,
 * src/frame/encoder.rs:351) -> linear -> XYB planes.
 * Replaces linearize_rgb (src/encode_image.rs:1498) followed by
 * to_xyb_in_place_with_fn (src/xyb.rs:439) with the selected ToXybBandFn
 * (src/xyb.rs:347).  `fwd` is XybMatrix::fwd (src/xyb.rs:81-96); `lut` the
 * 256-entry transfer table.  Writes planes X, Y, B at d_xyb + c*stride*padded_h. */
int jxg_xyb_dev(jxg_ctx *ctx, const uint8_t *d_rgb, size_t row_bytes, int channels,
                uint32_t width, uint32_t height, const float fwd[9], const float lut[256],
                int order, int fused, float *d_xyb, uint32_t stride, uint32_t padded_h);

/* Adaptive quant field of the whole image, one DC group (2048 px) at a time.
 * Replaces (ctx.fill_quant_field) = fill_quant_field_scalar
 * (src/adaptive_quant.rs:547-729) as called by setup_dc_group
 * (src/frame/dc_groups.rs:149).  d_qf is xblocks x yblocks bytes. */
int jxg_fill_quant_field_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride,
                             uint32_t padded_h, uint32_t width, uint32_t height,
                             float distance, float inv_scale, int fused, uint8_t *d_qf);

/* jxg_fill_quant_field_dev without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * AQ pre-map lives in a static device array of JXG_NOALLOC_AQ_PRE_FLOATS
 * floats (16 MiB, one per GPU, held until the process ends), so images up to
 * 2*ceil(W/8) * 2*ceil(H/8) <= JXG_NOALLOC_AQ_PRE_FLOATS (about 67
 * megapixels) are accepted and larger ones return JXG_ERR_TOO_LARGE.  Every
 * call on a GPU shares that array: calls that overlap on one GPU (two
 * streams, two contexts) race.  Output is identical to
 * jxg_fill_quant_field_dev's. */
#define JXG_NOALLOC_AQ_PRE_FLOATS (4u << 20)
int jxg_fill_quant_field_noalloc_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride,
                                     uint32_t padded_h, uint32_t width, uint32_t height,
                                     float distance, float inv_scale, int fused, uint8_t *d_qf);

/* Coefficient-model RD costs of every square candidate the Fast tiers
 * evaluate: per block its DCT8 cost, per 2x2-aligned site the DCT16X16 cost,
 * per 4x4-aligned site the DCT32X32 cost (NaN where a site cannot be placed).
 * Replaces strategy_cost (src/ac_strategy.rs:724 -> :1293) for those
 * candidates.  Outputs are xblocks x yblocks floats each (the 16/32 costs are
 * stored at the site's first block).  d_ytox / d_ytob are the per-64x64-tile
 * CfL maps, tiles_x = ceil(W/64) columns. */
int jxg_strategy_costs_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride,
                           uint32_t padded_h, uint32_t width, uint32_t height,
                           const uint8_t *d_qf, const int8_t *d_ytox, const int8_t *d_ytob,
                           const jxg_frame_params *params, float *d_cost8,
                           float *d_cost16, float *d_cost32);

/* Coefficient-model RD cost of one strategy at a list of sites: the cost
 * table the slower tiers' layout search (rectangles, sub-8x8 and DCT64
 * passes) compares.  Replaces strategy_cost (src/ac_strategy.rs:724 ->
 * strategy_cost_impl :1293) for every strategy up to 64x64, the DCT64
 * family's strategy_cost64 (:646, callers src/ac_strategy/selection.rs:
 * 2077-2189) and sub8_strategy_costs' per-candidate evaluation (:1746,
 * caller selection.rs:1266).  Site i is the first block
 * d_sites[i] = by << 16 | bx with quantizer d_qac[i] (the reference's qac,
 * scale * quant value); meta_r is the candidate's metadata charge (META_R = 4
 * at the chooser's call sites, 0 for refinement).  Costs use the model rate
 * (learned rate prices, src/ac_strategy/rate_prices.rs, stay on the CPU);
 * the selection quantizer applies to the 32-family as at the Fast tiers.
 * d_cost[i] receives the cost, NaN for a site whose footprint leaves the
 * padded planes.  DCT128 strategies are rejected. */
int jxg_strategy_cost_sites_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride,
                                uint32_t padded_h, uint32_t width, uint32_t height,
                                const int8_t *d_ytox, const int8_t *d_ytob,
                                const jxg_frame_params *params, int strategy, uint32_t n_sites,
                                const uint32_t *d_sites, const float *d_qac, float meta_r,
                                float *d_cost);

/* The transform-layout decision and the final quant field of speeds
 * ULTRAFAST (every block DCT8), FASTEST and FAST; MEDIUM and slower return
 * JXG_ERR_UNSUPPORTED (their rectangle, sub-8, DCT64 and layout-DP
 * decisions are not built; jxg_strategy_cost_sites_dev gives their
 * candidate costs).  Replaces the Fast / Fastest path of fill_ac_strategy:
 * the square-merge core (src/ac_strategy/selection.rs:1000-1180 select_band,
 * :238 select_super_block), layout_dp::retile_coefficient (:2194-2208,
 * selection/layout_dp.rs:2381; the (4, 4)-shifted grid and unaligned pieces
 * at FAST only), at FAST for d <= 1.709 the SSIM rerank (:2221-2330,
 * 2718-3232, 3568), adjust_quant_field (:503) and, at FAST for d >= 1.5,
 * refine_quant_field with qf_trellis (:534-896).  select_band_leaf_first
 * (d <= 1.5) gives the same layout as select_band at these tiers.  Writes
 * strategy cells and updates d_qf in place. */
int jxg_select_strategy_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride,
                            uint32_t padded_h, uint32_t width, uint32_t height,
                            uint8_t *d_qf, const int8_t *d_ytox, const int8_t *d_ytob,
                            const jxg_frame_params *params, uint8_t *d_cells);

/* Forward transform + quantization of every first block.  Replaces the
 * transform/quantize half of write_ac_group (src/group.rs:1259-2240) in its
 * first pass (no RDOQ): per channel the strategy's DCT, DC from the lowest
 * frequencies, Y DC quantization, the Y quantize/dequantize round trip, Y LLF
 * zeroing, CfL subtraction, X/B DC (with CfL) and X/B AC quantization.
 * Outputs per channel c (X=0, Y=1, B=2):
 *   d_levels[c]  int32 [yblocks*xblocks*64] quantized AC levels,
 *   d_coeffs[c]  f32   same layout: Y = source coefficients before the
 *                round trip, X/B = post-CfL coefficients (RDOQ inputs),
 *   d_dc[c]      int16 [yblocks*xblocks] quantized DC.
 * A transform covering cov_x x cov_y blocks at (bx, by) stores its
 * coefficients (normalised so the long side is the row) in chunks of
 * cov_x*64 values, chunk j at ((by+j)*xblocks + bx)*64. */
int jxg_forward_quantize_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride,
                             uint32_t padded_h, uint32_t width, uint32_t height,
                             const uint8_t *d_cells, const uint8_t *d_qf,
                             const int8_t *d_ytox, const int8_t *d_ytob,
                             const jxg_frame_params *params, int32_t *d_levels[3],
                             float *d_coeffs[3], int16_t *d_dc[3]);

/* Chroma-from-luma slope maps: per 64x64 tile the Y-to-X and Y-to-B
 * multipliers (i8, tiles_x = ceil(W/64) per row) by the closed-form
 * regression over the tile's DCT8 coefficients.  Replaces fill_cmap
 * (src/color_correlation.rs:901) on its regression path (compute_cmap_tile
 * :441), the path XYB frames take at Fast and faster, as called by
 * setup_dc_group (src/frame/dc_groups.rs:171). */
int jxg_fill_cmap_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                      uint32_t width, uint32_t height, const jxg_frame_params *params,
                      int8_t *d_ytox, int8_t *d_ytob);

/* The whole front end on device-resident input: XYB -> AQ -> CfL maps ->
 * strategy -> forward DCT + quantize.  d_ytox / d_ytob receive the CfL maps
 * (NULL: kept in context scratch).  XYB planes live in context scratch. */
int jxg_front_end_dev(jxg_ctx *ctx, const uint8_t *d_rgb, size_t row_bytes, int channels,
                      uint32_t width, uint32_t height, const jxg_frame_params *params,
                      int8_t *d_ytox, int8_t *d_ytob, uint8_t *d_qf, uint8_t *d_cells,
                      int32_t *d_levels[3], float *d_coeffs[3], int16_t *d_dc[3]);

/* ---- batched front end ----
 * Many images of uneven sizes and qualities through the front end with one
 * launch per stage for the whole batch (eight launches in all), the stage
 * chain captured once into a HIP graph and replayed.  The reference encodes
 * one image per encode_image call (src/encode_image.rs:1610); a batch is N
 * such calls' front ends.  Per image the outputs are bit-identical to
 * jxg_front_end_dev with the same parameters.  Batches take 1- and
 * 3-channel input (gray, RGB) without JXG_FLAG_CHROMA_HF; all images share
 * params.fused. */
typedef struct jxg_batch_image {
  const uint8_t *d_rgb; /* device input, row_bytes per row */
  size_t row_bytes;
  int32_t channels; /* 1 or 3 */
  uint32_t width, height;
  jxg_frame_params params;
  int8_t *d_ytox, *d_ytob; /* CfL maps, ceil(W/64) x ceil(H/64) */
  uint8_t *d_qf, *d_cells; /* xblocks x yblocks */
  int32_t *d_levels[3];
  float *d_coeffs[3];
  int16_t *d_dc[3];
} jxg_batch_image;
typedef struct jxg_batch jxg_batch;

/* Validates the images, allocates the batch's scratch (XYB planes, AQ and
 * CfL staging) and uploads the per-image descriptors.  The device buffers
 * named in `images` must stay valid for the batch's lifetime; their
 * contents are read / written at every run. */
int jxg_batch_create(jxg_ctx *ctx, uint32_t n, const jxg_batch_image *images, jxg_batch **out);
void jxg_batch_destroy(jxg_batch *batch);
/* Enqueues the batch's stage chain on the context's stream: kernel launches
 * only, so a caller may capture it into its own graph. */
int jxg_batch_run(jxg_batch *batch);
/* Runs the chain as a HIP graph: captured on the first call (and after a
 * retune) on the batch's own stream, then replayed; ordered after prior
 * work on the context's stream and before later work. */
int jxg_batch_launch(jxg_batch *batch);
/* Status of the last run (synchronizes): as jxg_quantize_status, summed over
 * the batch. */
int jxg_batch_status(jxg_batch *batch, uint32_t out[2]);
/* Autotuner: times every candidate tile shape of the tunable stages (AQ
 * strip of 16 / 32 / 64 blocks, CfL sub-tiles of 16 / 32 rows, 64 / 256 /
 * 1024 workgroups for the crossing-transform queue) on this batch, `reps`
 * graph replays each, keeps the fastest (results do not depend on it) and
 * writes a one-line-per-candidate report (may be NULL). */
int jxg_batch_autotune(jxg_batch *batch, int reps, char *report, size_t report_len);
/* The tile shapes in use: {aq_strip, cm_rows, xt_grid}. */
int jxg_batch_config(jxg_batch *batch, int out[3]);

/* ---- quant-field tuning by gradient descent ----
 * No counterpart in the reference (its quant field comes from the AQ
 * heuristics and the strategy search).  A differentiable proxy of the coding
 * loss of a fixed layout as a function of the per-block quant field theta:
 * for every DCT8..DCT32 transform T (theta_T = mean over its blocks) and
 * channel c, the AC coefficients get additive uniform noise of the
 * quantizer's step s_k = 1 / (inv_c[k] * scale * theta_T * qm_c), the exact
 * inverse transform reconstructs the block, and
 *   L = sum_c w_c |Xhat - X|^2 (in-image pixels) + lambda sum_k log2(1 + |Chat_k / s_k|).
 * The gradient comes from a hand-written adjoint (the adjoint of the
 * inverse transform is the unscaled forward product); Adam runs on
 * ln theta; the result is rounded back to an integer field that is
 * constant over every transform, so the quantize stage and the bitstream
 * take it as they take the heuristic field.  Transforms whose footprint
 * leaves their 32x32 region, sub-8x8 and DCT64 transforms keep their field
 * (counted by jxg_qf_tuner_excluded).  d_xyb and d_cells must stay valid
 * for the tuner's lifetime. */
typedef struct jxg_qf_tuner jxg_qf_tuner;
int jxg_qf_tuner_create(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                        uint32_t width, uint32_t height, const uint8_t *d_cells,
                        const jxg_frame_params *params, float lambda, uint32_t seed,
                        uint32_t max_steps, jxg_qf_tuner **out);
void jxg_qf_tuner_destroy(jxg_qf_tuner *tuner);
/* Transforms left out of the objective (synchronizes). */
int jxg_qf_tuner_excluded(jxg_qf_tuner *tuner, uint32_t *count);
/* L and dL/dtheta per block (xblocks x yblocks floats, 0 for left-out
 * blocks) at the field d_theta with the noise of step `step`; *loss (host,
 * may be NULL) synchronizes. */
int jxg_qf_tuner_loss_grad(jxg_qf_tuner *tuner, const float *d_theta, uint32_t step,
                           float *d_grad, float *loss);
/* `steps` Adam steps (b1 0.9, b2 0.999, eps 1e-8) at learning rate lr on
 * ln theta from the field d_qf_in, then the rounded field into d_qf_out
 * (may equal d_qf_in).  d_theta_out (may be NULL) receives the final real
 * field; loss_trace (host, steps + 1 floats, may be NULL; synchronizes)
 * the loss before each step and at the final field. */
int jxg_qf_tuner_run(jxg_qf_tuner *tuner, const uint8_t *d_qf_in, uint32_t steps, float lr,
                     uint8_t *d_qf_out, float *d_theta_out, float *loss_trace);

/* Host-buffer convenience form of jxg_front_end_dev for callers that own
 * host memory (the reference's encode_image input is a host slice,
 * src/encode_image.rs:1610): uploads, runs, downloads, synchronizes.
 * ytox / ytob may be NULL when the maps are not wanted. */
int jxg_front_end(jxg_ctx *ctx, const uint8_t *rgb, size_t row_bytes, int channels,
                  uint32_t width, uint32_t height, const jxg_frame_params *params,
                  int8_t *ytox, int8_t *ytob, uint8_t *qf, uint8_t *cells,
                  int32_t *levels[3], float *coeffs[3], int16_t *dc[3]);

/* Status of the last jxg_forward_quantize_dev on this context (synchronizes):
 * out[0] = transforms whose footprint crossed their 64x64 tile (handled by
 * the crossing pass), out[1] = 1 if a strategy cell was unsupported or did
 * not fit the image (those transforms were not written).  The reference
 * asserts on such cells (src/dc_group_data.rs:219-226); this ABI reports. */
int jxg_quantize_status(jxg_ctx *ctx, uint32_t out[2]);

/* DCT64X64 of every 64x64 block of the padded planes (ceil(W/64) x
 * ceil(H/64) blocks, edge replication past the image), the large-transform
 * workload: method JXG_DCT64_PLAIN runs the reference's FP32 butterflies in
 * its operation order (dct64x64_scalar_input, src/dct.rs:2549; bit-identical
 * to the quantize stage's DCT64X64); JXG_DCT64_MFMA_BF16X3 computes
 * C X C^T / 4096 on the bf16 matrix cores with split-precision operands
 * (hi + lo bf16, three products, f32 accumulation), an approximation whose
 * error against FP64 the caller measures; JXG_DCT64_MFMA_BF16X6 splits
 * every operand in three bf16 terms (exact for an f32) and keeps the six
 * products above 2^-24 of the leading one.  d_out[c] receives
 * 64 * (8 ceil(W/64)) * (8 ceil(H/64)) floats: the DCT64X64 coefficient
 * storage (u-major) in the chunk layout of jxg_forward_quantize_dev with
 * xblocks = 8 ceil(W/64).  No counterpart entry point in the reference (its
 * DCT64 runs per transform inside write_ac_group, src/group.rs:1534-1640). */
enum { JXG_DCT64_PLAIN = 0, JXG_DCT64_MFMA_BF16X3 = 1, JXG_DCT64_MFMA_BF16X6 = 2 };
int jxg_dct64_image_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                        uint32_t width, uint32_t height, int method, int fused,
                        float *d_out[3]);

/* DCT128X128 of every 128x128 block (ceil(W/128) x ceil(H/128) blocks; the
 * planes must reach whole 128-px blocks: stride >= 128 ceil(W/128) and
 * padded_h >= 128 ceil(H/128), edge-replicated), the 128 family's square
 * transform: method JXG_DCT64_PLAIN runs the reference's FP32 butterflies
 * in its order (dct128x128_scalar_input, src/dct128.rs:385; 128-point lines
 * dct1d_128 :330 over two 64-point halves), JXG_DCT64_MFMA_BF16X3 computes
 * C X C^T / 16384 on the bf16 matrix cores with split operands.  d_out[c]
 * receives 16384 coefficients per block (out[u * 128 + v]) in the chunk
 * layout with xblocks = 16 ceil(W/128).  The reference builds the 128
 * family only with its `extraslow` feature. */
int jxg_dct128_image_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride, uint32_t padded_h,
                         uint32_t width, uint32_t height, int method, int fused,
                         float *d_out[3]);

/* ---- EPF luma proxy ----
 * The six-term score the EPF fit ranks reconstructions with: structure pools
 * of the half- and quarter-resolution luma (5x5 mirrored box moments) and
 * error / excess / fourth-power pools of adjacent-pixel gradients against the
 * source.  Per-pixel values follow the reference's f32 order; pooled sums are
 * accumulated in f64 (the reference uses one f32 running sum). */
typedef struct jxg_epf_proxy jxg_epf_proxy;
/* LumaProxy::new (src/epf_proxy.rs:151; callers src/epf_filter.rs:832,1098):
 * copies the W x H source luma (row_stride floats per row; W, H >= 4, the
 * reference asserts) and builds the reference scales on the context's
 * stream. */
int jxg_epf_proxy_new(jxg_ctx *ctx, const float *d_y, uint32_t row_stride, uint32_t width,
                      uint32_t height, jxg_epf_proxy **out);
void jxg_epf_proxy_destroy(jxg_epf_proxy *proxy);
/* LumaProxy::compare (src/epf_proxy.rs:173): score and, if terms is not
 * NULL, the six terms {quarter structure, half structure, half structure
 * L4, gradient error, gradient excess, gradient loss L4}.  Synchronizes. */
int jxg_epf_proxy_compare(jxg_epf_proxy *proxy, const float *d_y, uint32_t row_stride,
                          float *score, float terms[6]);

/* Gaborish pre-inversion of the three XYB planes: a normalised symmetric
 * 5x5 sharpening with mirror-reflected image edges, so that the decoder's
 * gaborish smoothing reproduces the input.  Replaces gaborish_inverse
 * (src/gaborish.rs:44) as called by gaborize (src/frame/quantization.rs:241,
 * mul 0.9908511).  d_dst may equal d_src (in place, as the reference runs
 * it; staged through context scratch) or be a separate set of planes of the
 * same geometry.  Pixels past the image are written as the edge replication
 * of the filtered image.  mul must be finite and >= 0 (the reference
 * debug-asserts mul >= 0). */
int jxg_gaborish_inverse_dev(jxg_ctx *ctx, const float *d_src, float *d_dst, uint32_t stride,
                             uint32_t padded_h, uint32_t width, uint32_t height, float mul,
                             int fused);

/* apply_chroma_hf_protection (src/adaptive_quant.rs:303) on the quant
 * field of the whole image, in place: blocks whose chroma is strong, whose
 * chroma gradients are real and carry a large share of the edge energy get
 * their quant value raised by up to (0.35 + 1.4 share^2) * strength, the
 * strength fading in from distance 0.35 and out by 2.0.  Statistics and
 * weights follow the scalar path (chroma_hf_stats_scalar :222); `fused`
 * selects the fmla semantics. */
int jxg_chroma_hf_protection_dev(jxg_ctx *ctx, const float *d_xyb, uint32_t stride,
                                 uint32_t padded_h, uint32_t width, uint32_t height,
                                 float distance, int fused, uint8_t *d_qf);

/* ---- alpha stages (2- and 4-channel input) ----
 * jxg_front_end_dev runs these itself according to params->flags; they are
 * exported one by one for callers that stage the frame themselves.  Alpha
 * planes are W x H bytes, row stride W. */

/* alpha_fringe::clear (src/alpha_fringe.rs:77) in place: alpha samples 1..4
 * whose clipped 9x9 window of such samples has mean <= 1/2 become 0; then,
 * when near-opaque samples (251..254) are at least 1/12 of the plane, the
 * same sweep over 255 - alpha fills them with 255 if it changes at least
 * 1/12 of the plane.  Callers: the RGBA and gray + alpha paths of
 * encode_image_with_alpha (src/encode_image.rs:3650,3980). */
int jxg_alpha_fringe_clear_dev(jxg_ctx *ctx, uint8_t *d_alpha, uint32_t width, uint32_t height);

/* fill_invisible (src/invisible.rs:34) on the three planes, as called by
 * encode_frame (src/frame/encoder.rs:152): every alpha-0 sample, in
 * row-major order, becomes the weighted mean of its filled left and
 * upper-right neighbours and its visible right, lower-right, upper and lower
 * neighbours.  Bit-identical to the sequential scan.  The edge padding is
 * rewritten from the filled edge samples. */
int jxg_fill_invisible_dev(jxg_ctx *ctx, float *d_xyb, uint32_t stride, uint32_t padded_h,
                           uint32_t width, uint32_t height, const uint8_t *d_alpha);

/* alpha_block_factors (src/adaptive_quant.rs:953): per 8x8 block
 * (ceil(W/8) x ceil(H/8) floats) the largest alpha / 255, 1 for fully
 * transparent or opaque blocks, at least 0.1 otherwise. */
int jxg_alpha_block_factors_dev(jxg_ctx *ctx, const uint8_t *d_alpha, uint32_t width,
                                uint32_t height, float *d_factors);

/* apply_alpha_factors (src/adaptive_quant.rs:1036) on the final quant field:
 * every transform's quant value times the largest factor of its footprint,
 * rounded and clamped to 1..255.  The reference applies it only for
 * non-constant alpha (src/frame/encoder.rs:60-69; caller
 * src/frame/dc_groups.rs:305). */
int jxg_apply_alpha_factors_dev(jxg_ctx *ctx, const uint8_t *d_cells, const float *d_factors,
                                uint32_t width, uint32_t height, uint8_t *d_qf);

/* The alpha plane the last jxg_front_end_dev / jxg_front_end call used
 * (after fringe clearing), W x H bytes, copied to dst (host memory when
 * to_host, synchronously; device memory otherwise, stream-ordered).  The
 * caller's CPU back end encodes this plane as the alpha channel. */
int jxg_alpha_plane(jxg_ctx *ctx, uint8_t *dst, size_t bytes, int to_host);

#ifdef __cplusplus
}
#endif

#endif /* JXG_H_ */
