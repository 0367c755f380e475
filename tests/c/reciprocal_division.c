/* The face flux divides its nine reconstructions by one interface volume
 * b through div_by_reciprocal (gamera-cube_amd/csrc/gcs_math.hpp): q = a y
 * with y = 1 / b, then q + (a - b q) y with fused multiply-adds. This checks
 * on random normal operands, with significands crowded at both ends of
 * [1, 2), that the result is bit for bit the IEEE quotient a / b. Prints the
 * number of mismatches. */
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
int main(void) {
  long bad = 0, n = 0;
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
    const double y = 1.0 / b;
    double q = a * y;
    const double r = fma(-q, b, a);
    q = fma(r, y, q);
    const double ref = a / b;
    ++n;
    if (memcmp(&q, &ref, 8) != 0) { if (bad < 5) printf("mismatch a=%a b=%a q=%a ref=%a\n", a, b, q, ref); ++bad; }
  }
  printf("n=%ld mismatches=%ld\n", n, bad);
  return 0;
}
