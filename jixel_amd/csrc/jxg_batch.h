// jxg_batch.h - per-image descriptors of a batched front end (many images of
// uneven sizes and qualities in one launch per stage).
//
// Every stage kernel of the batch covers the workgroups of all images in one
// 1-D grid: stage s gives image i the workgroups [pfx[s][i], pfx[s][i+1]),
// laid out as that image's own 2-D grid of width gw[s] (row-major).  A
// workgroup finds its image by binary search over the prefix and then runs
// the single-image stage body with that image's pointers and parameters, so
// the batched and the single-image paths share every line of stage code.
#ifndef JXG_BATCH_H_
#define JXG_BATCH_H_

#include <stdint.h>

#include "jxg_internal.h"

enum {
  BS_XYB = 0,
  BS_AQPRE,
  BS_AQBLK,
  BS_CMAPB,
  BS_CMAPS,
  BS_SEL,
  BS_QUANT,
  BS_PIECE,
  BS_RETILE,
  BS_ADJUST,
  BS_RLIST,
  BS_RCOST_S,
  BS_RCOST_B,
  BS_TRELLIS,
  BS_RRLIST,
  BS_RRCOST_S,
  BS_RRCOST_B,
  BS_RRDECIDE,
  BS_N
};

struct BatchImg {
  const uint8_t *rgb;
  uint64_t row_bytes;
  int ch, W, H, stride, ph, xb, yb, tiles_x, tiles_y, pre_w, pre_h;
  int gw[BS_N];
  float *xyb;   // 3 planes of stride x ph
  float *pre;   // pre_w x pre_h AQ pre-pass
  float *part;  // CfL block sums, tiles * 64 * 4
  float *pc;    // re-tiling piece costs, 4 * xb * yb
  float *c32;   // raw DCT32X32 costs, xb * yb
  uint8_t *rf;  // quant-field refinement scratch (refine_scratch_bytes)
  uint8_t *qf, *cells;
  int8_t *ytox, *ytob;
  int32_t *lv[3];
  float *co[3];
  int16_t *dc[3];
  const jxg_derived *D;
  const float *tables, *selq;  // quant tables of the image's frame parameters
  float distance, inv_scale, cfl_base_x, cfl_base_b, cfl_color_factor;
};

// Image of flat workgroup `flat`: the i with pfx[i] <= flat < pfx[i + 1].
__device__ __forceinline__ int batch_image(const uint32_t *__restrict__ pfx, int n, uint32_t flat) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pfx[mid] <= flat)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

#endif  // JXG_BATCH_H_
