/*
 * sin/cos from a double-double evaluation — TEST INFRASTRUCTURE ONLY.
 *
 * The same arithmetic, operation for operation, as the device's relaxation
 * angles (Cody-Waite reduction by pi/2 in five 33-bit pieces, Taylor series of
 * sin r / r and cos r in z = r^2 on |r| <= pi/4 with the leading terms in
 * double-double, final rounding of the double-double).  ORC_LIBM_DD selects it in
 * orc_dynamics_apply so that a whole Henon step can be compared with the device
 * bit for bit; ORC_LIBM_SYSTEM keeps the C library's sin/cos as the reference
 * calls them (cmc_dynamics_helper.c:860-866, cmc_dynamics.c:457-459).
 *
 * Pinned by tests/test_oracle_trig.py: correctly rounded results (exact rational
 * Taylor sums) on random and near-multiple-of-pi/2 arguments.
 */
#include "cmc_oracle.h"
#include <math.h>

typedef struct { double hi, lo; } dd_t;

static dd_t two_sum(double a, double b)
{
    double s = a + b, bb = s - a;
    dd_t r = {s, (a - (s - bb)) + (b - bb)};
    return r;
}
static dd_t fast_two_sum(double a, double b)
{
    double s = a + b;
    dd_t r = {s, b - (s - a)};
    return r;
}
static dd_t dd_add(dd_t a, dd_t b)
{
    dd_t s = two_sum(a.hi, b.hi), t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
static dd_t dd_add_d(dd_t a, double b)
{
    dd_t s = two_sum(a.hi, b);
    s.lo += a.lo;
    return fast_two_sum(s.hi, s.lo);
}
static dd_t dd_mul(dd_t a, dd_t b)
{
    double p = a.hi * b.hi, e = fma(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p, e);
}

/* 1/n!, n = 0..29, as (hi, lo) */
static const double inv_fact[30][2] = {
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

static dd_t coef(int n, int neg)
{
    dd_t c = {inv_fact[n][0], inv_fact[n][1]};
    if (neg) { c.hi = -c.hi; c.lo = -c.lo; }
    return c;
}

void orc_dd_sincos(double x, double *s_out, double *c_out)
{
    if (!(fabs(x) < 1.0e5)) { *s_out = sin(x); *c_out = cos(x); return; }
    const double two_over_pi = 0x1.45f306dc9c883p-1;
    const double P1 = 0x1.921fb544p+0, P2 = 0x1.0b4611a6p-34, P3 = 0x1.3198a2ep-69, P4 = 0x1.b839a252p-104,
                 P5 = 0x1.27044533p-142;
    double k = rint(x * two_over_pi);
    dd_t r = two_sum(x - k * P1, -(k * P2));
    r = dd_add_d(r, -(k * P3));
    r = dd_add_d(r, -(k * P4));
    r = dd_add_d(r, -(k * P5));
    dd_t z = dd_mul(r, r);
    double ts = 0.0, tc = 0.0;
    for (int n = 14; n >= 8; n--) ts = ts * z.hi + ((n & 1) ? -inv_fact[2 * n + 1][0] : inv_fact[2 * n + 1][0]);
    for (int n = 14; n >= 9; n--) tc = tc * z.hi + ((n & 1) ? -inv_fact[2 * n][0] : inv_fact[2 * n][0]);
    dd_t S = {ts, 0.0}, Cc = {tc, 0.0};
    for (int n = 7; n >= 0; n--) S = dd_add(dd_mul(S, z), coef(2 * n + 1, n & 1));
    for (int n = 8; n >= 0; n--) Cc = dd_add(dd_mul(Cc, z), coef(2 * n, n & 1));
    S = dd_mul(S, r);
    double sr = S.hi + S.lo, cr = Cc.hi + Cc.lo;
    switch ((int)((long long)k & 3)) {
    case 0: *s_out = sr; *c_out = cr; break;
    case 1: *s_out = cr; *c_out = -sr; break;
    case 2: *s_out = -sr; *c_out = -cr; break;
    default: *s_out = -cr; *c_out = sr; break;
    }
}
