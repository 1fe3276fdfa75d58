/* jxo_internal.h - shared helpers of the oracle (TEST INFRASTRUCTURE ONLY). */
#ifndef JXO_INTERNAL_H_
#define JXO_INTERNAL_H_

#include <math.h>
#include <stdint.h>

extern int jxo_fused_flag;

/* The reference's fmla (src/dct.rs:33-53). */
static inline float jxo_fmla(float a, float b, float c) {
  return jxo_fused_flag ? fmaf(a, b, c) : a * b + c;
}

float jxo_round(float x);
int32_t jxo_sat_i32(float v);
int16_t jxo_sat_i16(float v);
int32_t jxo_round_i32(float v);
float jxo_aq_dampen(float distance);
uint16_t jxo_f32_to_f16(float v);
float jxo_f16_to_f32(uint16_t b);
const float *jxo_exp2_table(void);

/* Strategy pixel dimensions (inflated_cost::strategy_pixel_dims,
 * src/inflated_cost.rs:555-575). */
void jxo_strategy_pixel_dims(int strategy, int *w, int *h);
/* Quant-table kind of a strategy (same mapping as jxg_qt_kind_of). */
int jxo_qt_kind(int strategy);
uint32_t jxo_qt_size(int kind);
uint32_t jxo_qt_offset(int kind);

#endif
