/* st_div_small(a, n, 1/n) against the f32 division a / n, n = 1..6.
 * usage: div_small_check STRIDE
 *   STRIDE 1: every one of the 2^32 bit patterns of a (the exhaustive run);
 *   STRIDE > 1: every STRIDE-th bit pattern, plus every a whose biased exponent
 *   is below 4 or above 251 (gradual underflow, where n = 6 needs the division
 *   fallback, and the overflow end), both signs; +-0, +-inf and NaN are in there.
 * Prints mismatches per n; exit status 1 on any. */
#include <stdio.h>
#include <stdlib.h>
#include "st_math.h"

static int same(float a, float b) { return st_f2u(a) == st_f2u(b) || (a != a && b != b); }

int main(int argc, char** argv) {
  long long stride = argc > 1 ? atoll(argv[1]) : 1;
  if (stride < 1) stride = 1;
  int fail = 0;
  for (int n = 1; n <= 6; n++) {
    volatile float fn = (float)n;
    float y = 1.0f / fn;
    unsigned long long bad = 0, seen = 0;
    long long steps = ((1LL << 32) + stride - 1) / stride;
#pragma omp parallel for reduction(+ : bad, seen) schedule(static)
    for (long long k = 0; k < steps; k++) {
      float a = st_u2f((uint32_t)(k * stride));
      seen++;
      if (!same(a / fn, st_div_small(a, fn, y))) bad++;
    }
    if (stride > 1) {
      /* exponent fields 0..3 and 252..255, both signs: 16 blocks of 2^23 */
#pragma omp parallel for reduction(+ : bad, seen) schedule(static)
      for (long long k = 0; k < (16LL << 23); k++) {
        uint32_t blk = (uint32_t)(k >> 23), m = (uint32_t)(k & 0x7fffff);
        uint32_t e = (blk & 7u) < 4u ? (blk & 7u) : 248u + (blk & 7u);
        uint32_t u = ((blk >> 3) << 31) | (e << 23) | m;
        float a = st_u2f(u);
        seen++;
        if (!same(a / fn, st_div_small(a, fn, y))) bad++;
      }
    }
    printf("n=%d checked=%llu mismatches=%llu\n", n, seen, bad);
    if (bad) fail = 1;
  }
  return fail;
}
