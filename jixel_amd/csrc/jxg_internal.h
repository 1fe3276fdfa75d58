// jxg_internal.h - constants shared by the host code and the gfx950 kernels
// of the VarDCT front end.  Everything here is plain data: the kernels read
// one `jxg_derived` block (built on the host from jxg_frame_params) from
// device memory.
#pragma once

#include <stdint.h>

#include "../../include/jxg.h"

#define JXG_RATE_LUT_N 1024
// log2f(1 + i) for the large levels, computed on the host by the C library's
// log2f (the function the reference's f32::log2 calls), so that no device
// log2f rounding enters a rate below 65536.
#define JXG_RATE_BIG_N 65536
#define JXG_AREA_CLASSES 7 /* cx*cy in {1,2,4,8,16,32,64} */

// Per-frame constants derived once on the host, in the reference's float
// operation order, so every kernel uses identical values.
typedef struct jxg_derived {
  float distance;
  float scale;
  float inv_scale;    // 1 / scale (the AQ stage's inv_scale)
  float inv_factor[3];  // INV_DC_QUANT[c] / dc_step[c] * scale_dc (src/group.rs:1295)
  float cfl_factor_x;   // dc_cfl_factor_x (src/color_correlation.rs:1608)
  float cfl_factor_b;   // dc_cfl_factor   (src/color_correlation.rs:1598)
  float cfl_base_x, cfl_base_b, cfl_inv_color_factor;
  float qm_mult[3];     // X: 1.25^(x_qm_scale-2), Y: 1, B: b_qm_mul
  float dequant_one[3];
  float dequant_numerator;
  float channel_weight[3];
  // quantize_ac_thresholds_scaled (src/group.rs:420) per channel and area class
  float thr[3][JXG_AREA_CLASSES][4];
  // MergeTuning at this distance/speed (src/ac_strategy.rs:288-350)
  float bias_16x16, bias_32x32, accept_16, accept_32, risk_k, site_k;
  int32_t selection_quantizer;  // selection_quantizer::enabled (Fast tiers, 1<=d<=4, SSIMULACRA2)
  int32_t run_selection;        // fill_ac_strategy reaches the chooser
  int32_t fused;
  int32_t retile;  // coefficient re-tiling: 0 off, 1 aligned grid (Fastest), 2 + unaligned
                   // pieces and the (4, 4)-shifted grid (Fast); src/effort.rs:311-349
  int32_t refine_steps;  // refine_quant_field: 0 off, 1 or 2 coarser steps (Fast, d >= 1.5)
  float recon_lambda;    // reconstruction_lambda(distance) (src/ac_strategy.rs:1667)
  int32_t rerank;        // the SSIM rerank's downgrade pass (Fast, d <= 1.7090991)
  float rerank_margin;   // MergeTuning::rerank_margin: 1, or 0 without the arbiter's downgrades
  uint32_t qt_off[JXG_QT_NUM_KINDS];  // quant-matrix blob layout
  uint32_t qt_n[JXG_QT_NUM_KINDS];
  float exp2_table[64];          // fast_exp2's 2^((i-32)/64)
  float rate_lut[JXG_RATE_LUT_N];  // log2(1+i) (src/inflated_cost.rs:57)
  const float *rate_big;  // device table of log2f(1+i), i < JXG_RATE_BIG_N (set by the context)
} jxg_derived;

#ifdef __cplusplus
extern "C" {
#endif
// Builds the derived constants; returns JXG_OK or an error code.
int jxg_derive(const jxg_frame_params *p, jxg_derived *out);
// Distance tier (0..4) of the default quant tables (jxg_tables.cpp).
int jxg_quant_tier(float distance, int arbiter);
#define JXG_QUANT_TIERS 5
// The format's f16 round trip (src/util.rs:427-518), jxg_params.cpp.
uint16_t jxg_f32_to_f16(float v);
float jxg_f16_to_f32(uint16_t b);
#ifdef __cplusplus
}
#endif

#ifdef __HIPCC__
#define JXG_HD __host__ __device__
#else
#define JXG_HD
#endif

// Scratch of the quant-field refinement for nb blocks: per block 4 candidate
// costs (f32), 4 values and a count (u8); two transform lists (u32) and their
// counters; per chain step its block (u32), candidate costs (4 f32), values
// (4 u8), count, chosen and fixed value and back pointers (u8 each).
static inline JXG_HD size_t jxg_refine_scratch_bytes(size_t nb) { return 61 * nb + 128; }

static inline JXG_HD int jxg_area_class(int cx, int cy) {
  int a = cx * cy, k = 0;
  while ((1 << k) < a) k++;
  return k;
}

// Quant-table kind of a raw strategy (src/group.rs:1459-1495).
static inline constexpr JXG_HD int jxg_qt_kind_dev(int s) {
  switch (s) {
    case JXG_DCT: return JXG_QT_DCT8;
    case JXG_DCT16X8:
    case JXG_DCT8X16: return JXG_QT_DCT16X8;
    case JXG_DCT16X16: return JXG_QT_DCT16X16;
    case JXG_DCT32X32: return JXG_QT_DCT32X32;
    case JXG_DCT32X16:
    case JXG_DCT16X32: return JXG_QT_DCT32X16;
    case JXG_DCT32X8:
    case JXG_DCT8X32: return JXG_QT_DCT8X32;
    case JXG_DCT64X64: return JXG_QT_DCT64X64;
    case JXG_DCT64X32:
    case JXG_DCT32X64: return JXG_QT_DCT64X32;
    case JXG_IDENTITY: return JXG_QT_IDENTITY;
    case JXG_DCT2X2: return JXG_QT_DCT2X2;
    case JXG_DCT4X4: return JXG_QT_DCT4X4;
    case JXG_DCT4X8:
    case JXG_DCT8X4: return JXG_QT_DCT4X8;
    default: return JXG_QT_AFV;
  }
}
