// sin and cos for the relaxation's scattering angles, rounded from a double-double
// evaluation (about 100 correct bits), so the result is the correctly rounded value
// except for arguments whose exact sine lies within ~2^-100 relative of a rounding
// midpoint.  The result therefore does not depend on the device math library.
// sincos() first tries a shorter evaluation (sincos_fast, error below 2^-71 of the
// result) and keeps its rounding when that bound decides it, about 99.99% of calls;
// otherwise it takes the double-double one (sincos_full).  Both give the same bits.
//
// Method: Cody-Waite reduction x = k pi/2 + r with pi/2 split into five 33-bit
// pieces (k * piece is exact for |k| < 2^20, so the reduction is exact up to the
// 2^-160 remainder of the split); then r = j pi/128 + t, |j| <= 32, |t| <= pi/256,
// with pi/128 split into three pieces (j times the two 47-bit ones is exact), and
//   sin r = sin(j pi/128) cos t + cos(j pi/128) sin t,
//   cos r = cos(j pi/128) cos t - sin(j pi/128) sin t,
// the table values in double-double (gen_trig_table.py: mpmath at 200 bits) and the
// Taylor series of sin t / t and cos t in z = t^2 to z^6, the terms up to z^3 in
// double-double and the smaller ones (below 2^-68 of the sum) in double.  pi/2 comes
// from Machin's formula 16 atan(1/5) - 4 atan(1/239) in exact rational arithmetic;
// the series coefficients are 1/n! split into a double and its remainder.
//
// Arguments with |x| >= 1e5 (never produced by the relaxation, whose angles lie in
// [0, 2 pi]) fall back to the library functions.
#pragma once
#include <hip/hip_runtime.h>
#include "dd_trig_fine.hpp"

namespace ddtrig {

struct dd { double hi, lo; };

__device__ __forceinline__ dd two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

__device__ __forceinline__ dd fast_two_sum(double a, double b)
{
    double s = a + b;
    double e = b - (s - a);
    return {s, e};
}

__device__ __forceinline__ dd add(dd a, dd b)
{
    dd s = two_sum(a.hi, b.hi);
    dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}

// big + small for |big.hi| >= |small.hi| with no cancellation (the Horner steps of
// the series, where the coefficient dominates the product by a factor >= 2e4)
__device__ __forceinline__ dd add_ord(dd big, dd small)
{
    dd s = fast_two_sum(big.hi, small.hi);
    s.lo += big.lo + small.lo;
    return fast_two_sum(s.hi, s.lo);
}

__device__ __forceinline__ dd add_d(dd a, double b)
{
    dd s = two_sum(a.hi, b);
    s.lo += a.lo;
    return fast_two_sum(s.hi, s.lo);
}

__device__ __forceinline__ dd mul(dd a, dd b)
{
    double p = a.hi * b.hi;
    double e = __builtin_fma(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p, e);
}

// 1/n! as (hi, lo), n = 0..29, with the sign of the series term folded in by the caller
__device__ __constant__ const double kInvFact[30][2] = {
    {0x1p+0, 0x0p+0}, {0x1p+0, 0x0p+0}, {0x1p-1, 0x0p+0},
    {0x1.5555555555555p-3, 0x1.5555555555555p-57}, {0x1.5555555555555p-5, 0x1.5555555555555p-59},
    {0x1.1111111111111p-7, 0x1.1111111111111p-63}, {0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65},
    {0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73}, {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76},
    {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73}, {0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76},
    {0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80}, {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83},
    {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87}, {0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92},
    {0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97}, {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101},
    {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103}, {0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107},
    {0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112}, {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120},
    {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120}, {0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124},
    {0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130}, {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135},
    {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139}, {0x1.88e85fc6a4e5ap-89, -0x1.71c37ebd16540p-143},
    {0x1.d1ab1c2dccea3p-94, 0x1.054d0c78aea14p-149}, {0x1.0a18a2635085dp-98, 0x1.b9e2e28e1aa54p-153},
    {0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157}};

__device__ __forceinline__ dd coef(int n, bool neg)
{
    dd c{kInvFact[n][0], kInvFact[n][1]};
    if (neg) { c.hi = -c.hi; c.lo = -c.lo; }
    return c;
}

// sin and cos of j pi/128, j = 0..32: {sin hi, sin lo, cos hi, cos lo}
__device__ __constant__ const double kTab[33][4] = {
    {0x0.0p+0, 0x0.0p+0, 0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.92155f7a3667ep-6, -0x1.b1d63091a0130p-64, 0x1.ffd886084cd0dp-1, -0x1.1354d4556e4cbp-55},
    {0x1.91f65f10dd814p-5, -0x1.912bd0d569a90p-61, 0x1.ff621e3796d7ep-1, -0x1.c57bc2e24aa15p-57},
    {0x1.2d52092ce19f6p-4, -0x1.9a088a8bf6b2cp-59, 0x1.fe9cdad01883ap-1, 0x1.521ecd0c67e35p-57},
    {0x1.917a6bc29b42cp-4, -0x1.e2718d26ed688p-60, 0x1.fd88da3d12526p-1, -0x1.87df6378811c7p-55},
    {0x1.f564e56a9730ep-4, 0x1.a2704729ae56dp-59, 0x1.fc26470e19fd3p-1, 0x1.1ec8668ecaceep-55},
    {0x1.2c8106e8e613ap-3, 0x1.13000a89a11e0p-58, 0x1.fa7557f08a517p-1, -0x1.7a0a8ca13571fp-55},
    {0x1.5e214448b3fc6p-3, 0x1.531ff779ddac6p-57, 0x1.f8764fa714ba9p-1, 0x1.ab256778ffcb6p-56},
    {0x1.8f8b83c69a60bp-3, -0x1.26d19b9ff8d82p-57, 0x1.f6297cff75cb0p-1, 0x1.562172a361fd3p-56},
    {0x1.c0b826a7e4f63p-3, -0x1.af1439e521935p-62, 0x1.f38f3ac64e589p-1, -0x1.d7bafb51f72e6p-56},
    {0x1.f19f97b215f1bp-3, -0x1.42deef11da2c4p-57, 0x1.f0a7efb9230d7p-1, 0x1.52c7adc6b4989p-56},
    {0x1.111d262b1f677p-2, 0x1.824c20ab7aa9ap-56, 0x1.ed740e7684963p-1, 0x1.e82c791f59cc2p-56},
    {0x1.294062ed59f06p-2, -0x1.5d28da2c4612dp-56, 0x1.e9f4156c62ddap-1, 0x1.760b1e2e3f81ep-55},
    {0x1.4135c94176601p-2, 0x1.0c97c4afa2518p-56, 0x1.e6288ec48e112p-1, -0x1.16b56f2847754p-57},
    {0x1.58f9a75ab1fddp-2, -0x1.efdc0d58cf620p-62, 0x1.e212104f686e5p-1, -0x1.014c76c126527p-55},
    {0x1.7088530fa459fp-2, -0x1.44b19e0864c5dp-56, 0x1.ddb13b6ccc23cp-1, 0x1.83c37c6107db3p-55},
    {0x1.87de2a6aea963p-2, -0x1.72cedd3d5a610p-57, 0x1.d906bcf328d46p-1, 0x1.457e610231ac2p-56},
    {0x1.9ef7943a8ed8ap-2, 0x1.6da81290bdbabp-57, 0x1.d4134d14dc93ap-1, -0x1.4ef5295d25af2p-55},
    {0x1.b5d1009e15cc0p-2, 0x1.5b362cb974183p-57, 0x1.ced7af43cc773p-1, -0x1.e7b6bb5ab58aep-58},
    {0x1.cc66e9931c45ep-2, 0x1.6850e59c37f8fp-58, 0x1.c954b213411f5p-1, -0x1.2fb761e946603p-58},
    {0x1.e2b5d3806f63bp-2, 0x1.e0d891d3c6841p-58, 0x1.c38b2f180bdb1p-1, -0x1.6e0b1757c8d07p-56},
    {0x1.f8ba4dbf89abap-2, -0x1.2ec1fc1b776b8p-60, 0x1.bd7c0ac6f952ap-1, -0x1.825a732ac700ap-55},
    {0x1.073879922ffeep-1, -0x1.a5a014347406cp-55, 0x1.b728345196e3ep-1, -0x1.bc69f324e6d61p-55},
    {0x1.11eb3541b4b23p-1, -0x1.ef23b69abe4f1p-55, 0x1.b090a58150200p-1, -0x1.926da300ffccep-55},
    {0x1.1c73b39ae68c8p-1, 0x1.b25dd267f6600p-55, 0x1.a9b66290ea1a3p-1, 0x1.9f630e8b6dac8p-60},
    {0x1.26d054cdd12dfp-1, -0x1.5da743ef3770cp-55, 0x1.a29a7a0462782p-1, -0x1.128bb015df175p-56},
    {0x1.30ff7fce17035p-1, -0x1.efcc626f74a6fp-57, 0x1.9b3e047f38741p-1, -0x1.30ee286712474p-55},
    {0x1.3affa292050b9p-1, 0x1.e3e25e3954964p-56, 0x1.93a22499263fbp-1, 0x1.3d419a920df0bp-55},
    {0x1.44cf325091dd6p-1, 0x1.8076a2cfdc6b3p-57, 0x1.8bc806b151741p-1, -0x1.2c5e12ed1336dp-55},
    {0x1.4e6cabbe3e5e9p-1, 0x1.3c293edceb327p-57, 0x1.83b0e0bff976ep-1, -0x1.6f420f8ea3475p-56},
    {0x1.57d69348ceca0p-1, -0x1.75720992bfbb2p-55, 0x1.7b5df226aafafp-1, -0x1.0f537acdf0ad7p-56},
    {0x1.610b7551d2cdfp-1, -0x1.251b352ff2a37p-56, 0x1.72d0837efff96p-1, 0x1.0d4ef0f1d915cp-55},
    {0x1.6a09e667f3bcdp-1, -0x1.bdd3413b26456p-55, 0x1.6a09e667f3bcdp-1, -0x1.bdd3413b26456p-55},
};

// The double-double evaluation above, rounded once at the end.
__device__ __forceinline__ void sincos_full(double x, double *s_out, double *c_out)
{
    if (!(__builtin_fabs(x) < 1.0e5)) {
        *s_out = ::sin(x);
        *c_out = ::cos(x);
        return;
    }
    const double two_over_pi = 0x1.45f306dc9c883p-1;
    const double P1 = 0x1.921fb544p+0, P2 = 0x1.0b4611a6p-34, P3 = 0x1.3198a2ep-69, P4 = 0x1.b839a252p-104,
                 P5 = 0x1.27044533p-142;
    double k = __builtin_rint(x * two_over_pi);
    dd r = two_sum(x - k * P1, -(k * P2));
    r = add_d(r, -(k * P3));
    r = add_d(r, -(k * P4));
    r = add_d(r, -(k * P5));
    // r = j pi/128 + t
    const double H1 = 0x1.921fb54442d00p-6, H2 = 0x1.8469898cc5180p-54, H3 = -0x1.fc8f8cbb5bf6cp-103;
    const double j = __builtin_rint(r.hi * 0x1.45f306dc9c883p+5);      // 128/pi
    dd t = two_sum(r.hi - j * H1, -(j * H2));
    t = add_d(t, r.lo);
    t = add_d(t, -(j * H3));
    dd z = mul(t, t);
    // sin t / t = sum_n (-1)^n z^n / (2n+1)!,   cos t = sum_n (-1)^n z^n / (2n)!
    double ts = 0.0, tc = 0.0;
    for (int n = 6; n >= 4; n--) ts = ts * z.hi + ((n & 1) ? -kInvFact[2 * n + 1][0] : kInvFact[2 * n + 1][0]);
    for (int n = 6; n >= 4; n--) tc = tc * z.hi + ((n & 1) ? -kInvFact[2 * n][0] : kInvFact[2 * n][0]);
    dd S{ts, 0.0}, Cc{tc, 0.0};
    for (int n = 3; n >= 0; n--) S = add_ord(coef(2 * n + 1, n & 1), mul(S, z));
    for (int n = 3; n >= 0; n--) Cc = add_ord(coef(2 * n, n & 1), mul(Cc, z));
    S = mul(S, t);
    const int aj = (int)__builtin_fabs(j);
    dd Sj{kTab[aj][0], kTab[aj][1]}, Cj{kTab[aj][2], kTab[aj][3]};
    if (j < 0) { Sj.hi = -Sj.hi; Sj.lo = -Sj.lo; }
    dd sr = add(mul(Sj, Cc), mul(Cj, S));
    dd p2 = mul(Sj, S);
    p2.hi = -p2.hi; p2.lo = -p2.lo;
    dd cr2 = add(mul(Cj, Cc), p2);
    double srd = sr.hi + sr.lo, crd = cr2.hi + cr2.lo;
    int q = (int)((long long)k & 3);
    double s, c;
    switch (q) {
    case 0: s = srd; c = crd; break;
    case 1: s = crd; c = -srd; break;
    case 2: s = -srd; c = -crd; break;
    default: s = -crd; c = srd; break;
    }
    *s_out = s;
    *c_out = c;
}

// True when hi is the correctly rounded value of any number within `rel` |hi| of
// hi + lo (hi + lo normalised, |lo| <= ulp(hi) / 2): hi + lo then lies farther than
// that from the rounding midpoints next to hi.  Below a power of two (lo pointing
// toward zero) the midpoint is half as far.
__device__ __forceinline__ bool rounds_surely(double hi, double lo, double rel)
{
    if (hi == 0.0) return lo == 0.0;
    int e;
    (void)__builtin_frexp(hi, &e);                 // |hi| in [2^(e-1), 2^e)
    const bool pow2 = (__double_as_longlong(hi) & 0x000fffffffffffffll) == 0;
    const bool down = lo != 0.0 && ((lo < 0.0) != (hi < 0.0));
    const double half_ulp = __builtin_ldexp(1.0, (pow2 && down) ? e - 55 : e - 54);
    return __builtin_fabs(lo) + rel * __builtin_fabs(hi) < half_ulp;
}

// The same values by a shorter evaluation, accepted only where its error bound decides
// the rounding.  The reduction to r is the exact one above; then r = j pi/1024 + t,
// |j| <= 256, |t| <= pi/2048, with pi/1024 = G1 + G2 + G3 (j times the two 44-bit
// pieces is exact) and sin, cos of j pi/1024 in double-double (dd_trig_fine.hpp).
// With z = t_hi^2 rounded and t^2 = z + zl (zl from the fma residual and 2 t_hi t_lo):
//   sin t = t + sl,          sl = t_lo + t_hi z (-1/3! + z (1/5! - z/7!)) - t_hi zl/3!
//   cos t = 1 - z/2 + cq,    cq = z^2 (1/4! - z/6!) - zl/2
// (the omitted terms are below 2^-89).  In
//   sin r = Sj + Cj t_hi - Sj z/2 + [Sj cq + Cj sl + Cj_lo t_hi + Sj_lo (1 - z/2)]
//   cos r = Cj - Sj t_hi - Cj z/2 + [Cj cq - Sj sl - Sj_lo t_hi + Cj_lo (1 - z/2)]
// the three leading terms are summed exactly (two_sum, fma residuals), so only the
// bracket, below 6.1e-10 + 2^-53 |Sj| in magnitude, is rounded: its few roundings stay
// below 2^-79 absolute, at most 2^-71 of the result (at worst for sin with j = 1), and
// sl's own error below 2^-73 |t|.  Returns (hi, lo) of sin r and cos r and k.
__device__ __forceinline__ void sincos_short_dd(double x, dd *sr_out, dd *cr_out, double *k_out)
{
    const double two_over_pi = 0x1.45f306dc9c883p-1;
    const double P1 = 0x1.921fb544p+0, P2 = 0x1.0b4611a6p-34, P3 = 0x1.3198a2ep-69, P4 = 0x1.b839a252p-104,
                 P5 = 0x1.27044533p-142;
    const double k = __builtin_rint(x * two_over_pi);
    dd r = two_sum(x - k * P1, -(k * P2));
    r = add_d(r, -(k * P3));
    r = add_d(r, -(k * P4));
    r = add_d(r, -(k * P5));
    const double j = __builtin_rint(r.hi * 0x1.45f306dc9c883p+8);      // 1024/pi
    dd t = two_sum(r.hi - j * kG1, -(j * kG2));
    t = add_d(t, r.lo);
    t = add_d(t, -(j * kG3));
    const double z = t.hi * t.hi;
    const double zl = __builtin_fma(t.hi, t.hi, -z) + 2.0 * t.hi * t.lo;
    const double ps = z * (-kInvFact[3][0] + z * (kInvFact[5][0] - z * kInvFact[7][0]));
    const double sl = t.lo + t.hi * (ps - kInvFact[3][1] * z) - t.hi * zl * kInvFact[3][0];
    const double cq = z * z * (kInvFact[4][0] - z * kInvFact[6][0]) - 0.5 * zl;
    const double ch = -0.5 * z;                                      // exact
    const int aj = (int)__builtin_fabs(j);
    double sjh = kTabFine[aj][0], sjl = kTabFine[aj][1];
    const double cjh = kTabFine[aj][2], cjl = kTabFine[aj][3];
    if (j < 0) { sjh = -sjh; sjl = -sjl; }
    // sin r
    const double a1 = cjh * t.hi, e1 = __builtin_fma(cjh, t.hi, -a1);
    const double p1 = sjh * ch, f1 = __builtin_fma(sjh, ch, -p1);
    dd u = two_sum(sjh, a1);
    dd v = two_sum(u.hi, p1);
    double lo = u.lo + v.lo + (e1 + f1 + (sjh * cq + cjh * sl + cjl * t.hi + sjl * (1.0 + ch)));
    *sr_out = fast_two_sum(v.hi, lo);
    // cos r
    const double a2 = sjh * t.hi, e2 = __builtin_fma(sjh, t.hi, -a2);
    const double p2 = cjh * ch, f2 = __builtin_fma(cjh, ch, -p2);
    u = two_sum(cjh, -a2);
    v = two_sum(u.hi, p2);
    lo = u.lo + v.lo + (f2 - e2 + (cjh * cq - sjh * sl - sjl * t.hi + cjl * (1.0 + ch)));
    *cr_out = fast_two_sum(v.hi, lo);
    *k_out = k;
}

// sincos_short_dd rounded, when 2^-68 of the result (8 times its error bound) cannot
// move the rounding; returns false otherwise.
__device__ __forceinline__ bool sincos_fast(double x, double *s_out, double *c_out)
{
    if (!(__builtin_fabs(x) < 1.0e5)) return false;
    dd sr, cr;
    double k;
    sincos_short_dd(x, &sr, &cr, &k);
    const double rel = 0x1.0p-68;
    if (!rounds_surely(sr.hi, sr.lo, rel) || !rounds_surely(cr.hi, cr.lo, rel)) return false;
    const double srd = sr.hi, crd = cr.hi;
    int q = (int)((long long)k & 3);
    double s, c;
    switch (q) {
    case 0: s = srd; c = crd; break;
    case 1: s = crd; c = -srd; break;
    case 2: s = -srd; c = -crd; break;
    default: s = -crd; c = srd; break;
    }
    *s_out = s;
    *c_out = c;
    return true;
}

__device__ __forceinline__ void sincos(double x, double *s_out, double *c_out)
{
    if (!sincos_fast(x, s_out, c_out)) sincos_full(x, s_out, c_out);
}

} // namespace ddtrig
