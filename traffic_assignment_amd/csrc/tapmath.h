/*
 * tapmath.h -- link-cost arithmetic shared by the gfx950 kernels and the
 * host-side checks.  Compiled for the device by hipcc and for the host by
 * g++ (the CPU test of the power function against glibc), always with
 * -ffp-contract=off so every + and * is one IEEE double rounding.
 *
 * pow(x, 4) is evaluated correctly rounded through an exact double-double
 * square of an exact double-double square (fma based), which is what glibc's
 * pow returns for these arguments (checked on the host against libm over
 * random samples); other integer powers 0..2 are exact special cases, and
 * any other power falls back to the platform pow (no parity pin).
 */
#ifndef TAPMATH_H
#define TAPMATH_H

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define TAP_HD __host__ __device__ __forceinline__
#define TAP_FMA(a, b, c) __builtin_fma((a), (b), (c))
#else
#include <math.h>
#define TAP_HD static inline
#define TAP_FMA(a, b, c) fma((a), (b), (c))
#endif

/* x^4 rounded once: x^2 = h1 + l1 and h1^2 = h2 + l2 exactly, then
   x^4 = h2 + (l2 + 2 h1 l1 + l1^2) with the tail far below h2's ulp. */
TAP_HD double tap_pow4(double x)
{
  const double h1 = x * x;
  const double h2 = h1 * h1;
  if (!(h2 < 1e300) || h2 < 1e-280)      /* overflow, zero or near-subnormal */
    return h2;
  const double l1 = TAP_FMA(x, x, -h1);
  const double l2 = TAP_FMA(h1, h1, -h2);
  double t = TAP_FMA(2.0 * h1, l1, l2);
  t = TAP_FMA(l1, l1, t);
  return h2 + t;
}

/* any other power: the platform pow, kept out of line so that the loops
   over step sizes that call tap_pow stay small enough to unroll (their
   sums then live in registers) */
#if defined(__HIPCC__) || defined(__HIP__)
__host__ __device__ __attribute__((noinline)) static double tap_pow_other(double x, double p)
#else
static double tap_pow_other(double x, double p)
#endif
{
  return pow(x, p);
}

TAP_HD double tap_pow(double x, double p)
{
  if (p == 4.0)
    return tap_pow4(x);
  if (p == 1.0)
    return x;
  if (p == 2.0)
    return x * x;
  if (p == 0.0)
    return 1.0;
  return tap_pow_other(x, p);
}

/* BPR cost t0*(1 + B*(Q/c)^p), trunk/tap/tap_functions.c:55-65 */
TAP_HD double tap_bpr(double t0, double cap, double B, double p, double Q)
{
  return t0 * (1.0 + B * tap_pow(Q / cap, p));
}

/* integral of BPR over [0, Q] in the reference's form,
   trunk/tap/tap_functions.c:112-126 */
TAP_HD double tap_integral(double t0, double cap, double B, double p, double Q)
{
  if (!(p >= 0))
    return 0.0;
  return Q * t0 * (1.0 + ((B / tap_pow(cap, p)) / (p + 1.0)) * tap_pow(Q, p));
}

/* compensated accumulator (TwoSum based; order-fixed callers give
   deterministic results) */
typedef struct {
  double s, c;
} tap_acc;

TAP_HD void tap_acc_add(tap_acc *a, double x)
{
  const double t = a->s + x;
  const double bp = t - a->s;
  const double e = (a->s - (t - bp)) + (x - bp);
  a->s = t;
  a->c += e;
}

TAP_HD void tap_acc_merge(tap_acc *a, tap_acc b)
{
  tap_acc_add(a, b.s);
  a->c += b.c;
}

TAP_HD double tap_acc_value(tap_acc a)
{
  return a.s + a.c;
}

#endif /* TAPMATH_H */
