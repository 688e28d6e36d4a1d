/* Exhaustive-by-sampling check of the reciprocal division used by the fused
 * stage kernels (div_rn in gmcore_amd/csrc/gmc_internal.hpp):
 *     q0 = a * y,  e = fma(-q0, d, a),  q = e == 0 ? q0 : fma(e, y, q0)
 * with y = 1 / d correctly rounded must equal the IEEE quotient a / d.
 *
 * usage: div_rn_check DENOMS_FILE SAMPLES_PER_DENOM SEED
 * DENOMS_FILE holds raw doubles (the divisor metric terms of real meshes);
 * adversarial divisors (powers of two, all-ones and one-ulp-above-one
 * significands) are added here.  For each divisor the numerators are random
 * doubles over 200 binades and, for the hard cases of rounding, a = RN(d m)
 * and its four nearest neighbours for random rounding midpoints m (d m is
 * exact in binary128), so that a / d lies within a few ulps of 2^-53 of a
 * midpoint.  Prints "checked N mismatches M" and exits 1 on a mismatch. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t st;
static uint64_t rnd(void) {
  st ^= st << 13;
  st ^= st >> 7;
  st ^= st << 17;
  return st;
}
static double from_bits(uint64_t b) {
  double x;
  memcpy(&x, &b, 8);
  return x;
}
static uint64_t bits(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  return b;
}
static double rdbl(int e0, int span) {
  return from_bits((rnd() & 0x000fffffffffffffull) | ((uint64_t)(1023 + e0 + (int)(rnd() % (uint64_t)span)) << 52));
}
static double div_rn(double a, double d, double y) {
  const double q0 = a * y;
  const double e = fma(-q0, d, a);
  return e == 0.0 ? q0 : fma(e, y, q0);
}

static long checked, bad;
static void check(double a, double d, double y) {
  const double want = a / d, got = div_rn(a, d, y);
  checked++;
  if (bits(want) != bits(got)) {
    if (bad < 5) printf("mismatch a=%a d=%a want=%a got=%a\n", a, d, want, got);
    bad++;
  }
}

static void one_divisor(double d, long n) {
  const double y = 1.0 / d;
  check(0.0, d, y);
  check(-0.0, d, y);
  for (long k = 0; k < n; k++) {
    double a = rdbl(-100, 200);
    check((rnd() & 1) ? -a : a, d, y);
    /* near-midpoint quotients: a / d ~ q + ulp(q) / 2 */
    const double q = rdbl(-40, 80);
    const __float128 m = (__float128)q + ((__float128)nextafter(q, INFINITY) - (__float128)q) / 2;
    const double a0 = (double)((__float128)d * m);
    for (int off = -2; off <= 2; off++) check(from_bits(bits(a0) + (int64_t)off), d, y);
  }
}

int main(int argc, char **argv) {
  if (argc < 4) return 2;
  const long n = atol(argv[2]);
  st = 0x9E3779B97F4A7C15ull ^ (uint64_t)atol(argv[3]);
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  double d;
  long nd = 0;
  while (fread(&d, sizeof d, 1, f) == 1) {
    if (d != 0.0 && isfinite(d)) { one_divisor(d, n); nd++; }
  }
  fclose(f);
  for (int e = -30; e <= 30; e += 3) {
    one_divisor(ldexp(1.0, e), n);
    one_divisor(ldexp(2.0 - ldexp(1.0, -52), e), n);
    one_divisor(ldexp(1.0 + ldexp(1.0, -52), e), n);
    one_divisor(-ldexp(1.5, e), n);
    nd += 4;
  }
  printf("divisors %ld checked %ld mismatches %ld\n", nd, checked, bad);
  return bad != 0;
}
