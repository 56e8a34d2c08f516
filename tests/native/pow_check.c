/* Host check of the product's pow(x,4), BPR and integral against glibc
 * (TEST INFRASTRUCTURE).  Prints the number of mismatching samples.
 * usage: pow_check N SEED */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include "../../traffic_assignment_amd/csrc/tapmath.h"

static uint64_t s;
static double urand(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return (s >> 11) * 0x1.0p-53; }

int main(int argc, char **argv)
{
  long n = argc > 1 ? atol(argv[1]) : 1000000;
  s = argc > 2 ? strtoull(argv[2], 0, 10) : 88172645463325252ull;
  long bad_pow = 0, bad_bpr = 0, bad_int = 0;
  for (long i = 0; i < n; i++) {
    double x;
    int kind = i % 4;
    if (kind == 0) x = urand() * 5.0;                       /* v/c ratios  */
    else if (kind == 1) x = ldexp(urand() + 0.5, (int)(urand() * 120) - 60);
    else if (kind == 2) x = floor(urand() * 1e6) / (4000.0 + floor(urand() * 22000));
    else x = 1e3 + urand() * 3e4;                           /* capacities  */
    double a = pow(x, 4.0), b = tap_pow4(x);
    if (a != b) { if (bad_pow < 5) printf("pow mismatch x=%a glibc=%a mine=%a\n", x, a, b); bad_pow++; }
    double t0 = 0.5 + urand() * 10, cap = 500 + urand() * 20000, Q = urand() * 30000;
    double r1 = t0 * (1.0 + 0.15 * pow(Q / cap, 4.0)), r2 = tap_bpr(t0, cap, 0.15, 4.0, Q);
    if (r1 != r2) bad_bpr++;
    double i1 = Q * t0 * (1 + ((0.15 / pow(cap, 4.0)) / (4.0 + 1)) * pow(Q, 4.0));
    double i2 = tap_integral(t0, cap, 0.15, 4.0, Q);
    if (i1 != i2) bad_int++;
  }
  printf("samples %ld pow4_mismatch %ld bpr_mismatch %ld integral_mismatch %ld\n", n, bad_pow, bad_bpr, bad_int);
  return (bad_pow || bad_bpr || bad_int) ? 1 : 0;
}
