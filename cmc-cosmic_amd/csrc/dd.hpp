// Double-double (unevaluated sum hi + lo) arithmetic for the potential scan.
// Error-free transformations: TwoSum (Knuth), FastTwoSum (Dekker) and the FMA
// remainder of a division.  Requires IEEE double with no reassociation
// (-ffp-contract=off, no fast-math), which the library build enforces.
#pragma once
#include <hip/hip_runtime.h>

struct dd { double hi, lo; };

__device__ __forceinline__ dd dd_two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
__device__ __forceinline__ dd dd_fast_two_sum(double a, double b)
{
    double s = a + b;
    double e = b - (s - a);
    return {s, e};
}
__device__ __forceinline__ dd dd_add(dd a, dd b)
{
    dd s = dd_two_sum(a.hi, b.hi);
    dd t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return dd_fast_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
__device__ __forceinline__ dd dd_from(double a) { return {a, 0.0}; }
// a / r as a double-double (quotient plus FMA-exact remainder over r)
__device__ __forceinline__ dd dd_div_d(double a, double r)
{
    double q = a / r;
    double e = fma(-q, r, a);
    return dd_fast_two_sum(q, e / r);
}
// (hi + lo) / r
__device__ __forceinline__ dd dd_div_dd_d(dd a, double r)
{
    double q1 = a.hi / r;
    double p = fma(-q1, r, a.hi);
    p += a.lo;
    double q2 = p / r;
    return dd_fast_two_sum(q1, q2);
}
__device__ __forceinline__ double dd_round(dd a) { return a.hi + a.lo; }
