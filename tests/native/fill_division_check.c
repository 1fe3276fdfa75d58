/* The fill's division, checked against IEEE float division: for every
 * weight w in 1 .. 13 and every float sum s in the chosen exponent ranges,
 * (float) fma(fma(-q, w, s), r, q) with q = s r, r = 1 / w rounded (FP64)
 * must equal s / w (float), bit for bit; -0 is the one known exception (the
 * fill's sums are never -0).  Usage: fill_division_check [full]; without
 * "full" the biased exponents 0 .. 8, 118 .. 136 and 246 .. 254 of both
 * signs (subnormals, the smallest normals, the values near 1 and the
 * largest), with "full" every finite float.  Prints the mismatch count. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static int wanted(int ex, int full) {
  if (full) return ex < 255;
  return ex <= 8 || (ex >= 118 && ex <= 136) || (ex >= 246 && ex <= 254);
}

int main(int argc, char **argv) {
  const int full = argc > 1 && strcmp(argv[1], "full") == 0;
  unsigned long long bad = 0, checked = 0;
  for (int w = 1; w <= 13; w++) {
    const double wd = (double)w, rd = 1.0 / wd;
    for (int ex = 0; ex < 255; ex++) {
      if (!wanted(ex, full)) continue;
      unsigned long long b = 0;
#pragma omp parallel for reduction(+ : b) schedule(static)
      for (int64_t m = 0; m < (1LL << 24); m++) {
        const uint32_t u = ((uint32_t)(m >> 23) << 31) | ((uint32_t)ex << 23) | (uint32_t)(m & 0x7FFFFF);
        if (u == 0x80000000u) continue; /* -0 */
        float x;
        memcpy(&x, &u, 4);
        const float want = x / (float)w;
        const double sd = (double)x, q0 = sd * rd;
        const float got = (float)fma(fma(-q0, wd, sd), rd, q0);
        uint32_t a, c;
        memcpy(&a, &want, 4);
        memcpy(&c, &got, 4);
        b += a != c;
      }
      bad += b;
      checked += 1ULL << 24;
    }
  }
  printf("checked %llu mismatches %llu\n", checked, bad);
  return bad != 0;
}
