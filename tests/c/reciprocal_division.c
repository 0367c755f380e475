/* The face flux divides its nine reconstructions by one interface volume
 * b through div_by_reciprocal (gamera-cube_amd/csrc/gcs_math.hpp): q = a y
 * with y = 1 / b, then q + (a - b q) y with fused multiply-adds, and the
 * plain division where that correction is not exact (zeros, infinities,
 * NaNs, subnormal dividends, quotients near underflow or overflow). This
 * checks the same function, bit for bit against the IEEE quotient a / b,
 * on random normal operands with significands crowded at both ends of
 * [1, 2), and on special and extreme operands. Prints the number of
 * mismatches. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 88172645463325252ull;
static uint64_t rnd(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double mk(uint64_t mant, int e, int neg) {
  uint64_t bits = ((uint64_t)(neg ? 1 : 0) << 63) | ((uint64_t)(e + 1023) << 52) | (mant & ((1ull << 52) - 1));
  double d; memcpy(&d, &bits, 8); return d;
}
/* the device function, statement for statement */
static double div_by_reciprocal(double a, double b, double y) {
  const double q = a * y;
  const double r = fma(-q, b, a);
  const double x = fma(r, y, q);
  const double ax = fabs(a), xx = fabs(x);
  if (ax >= 0x1p-960 && xx >= 0x1p-960 && xx <= 0x1p1000) return x;
  return a / b;
}
static long bad = 0, n = 0;
static void check(double a, double b) {
  const double q = div_by_reciprocal(a, b, 1.0 / b);
  const double ref = a / b;
  ++n;
  /* NaN results: any NaN matches (their payloads are not specified) */
  if (isnan(q) && isnan(ref)) return;
  if (memcmp(&q, &ref, 8) != 0) {
    if (bad < 5) printf("mismatch a=%a b=%a q=%a ref=%a\n", a, b, q, ref);
    ++bad;
  }
}
int main(void) {
  for (long it = 0; it < 20000000L; ++it) {
    uint64_t r1 = rnd(), r2 = rnd(), r3 = rnd();
    int mode = r3 & 7;
    uint64_t mb = r2;
    if (mode == 1) mb = (1ull << 52) - 1 - (r2 & 1023);      /* b near 2 */
    if (mode == 2) mb = r2 & 1023;                            /* b near 1 */
    uint64_t ma = r1;
    if (mode == 3) ma = (1ull << 52) - 1 - (r1 & 1023);
    if (mode == 4) ma = r1 & 1023;
    const double b = mk(mb, (int)((r3 >> 8) % 200) - 100, 0);
    const double a = mk(ma, (int)((r3 >> 16) % 200) - 100, (int)((r3 >> 24) & 1));
    check(a, b);
  }
  /* special and extreme operands: signed zeros, infinities, NaN, subnormal
   * dividends, quotients at the edges of the normal range */
  const double specials[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, 0x1p-1074, -0x1p-1060,
                             0x1p-1022, 0x1.8p-1000, 0x1p-960, 0x1p1000, 0x1.fffffffffffffp1023,
                             -0x1.fffffffffffffp1023, 1.0, -3.0};
  const int ns = (int)(sizeof(specials) / sizeof(specials[0]));
  for (int i = 0; i < ns; ++i)
    for (long it = 0; it < 2000; ++it) {
      const uint64_t r = rnd();
      const double b = mk(rnd(), (int)(r % 2000) - 1000, 0);
      check(specials[i], b);
      /* a random dividend against an extreme divisor */
      check(mk(rnd(), (int)((r >> 16) % 2000) - 1000, (int)((r >> 40) & 1)), b);
    }
  printf("n=%ld mismatches=%ld\n", n, bad);
  return 0;
}
