/*
 * sin/cos from a double-double evaluation — TEST INFRASTRUCTURE ONLY.
 *
 * The same arithmetic, operation for operation, as the device's relaxation
 * angles (cmc-cosmic_amd/csrc/dd_trig.hpp): Cody-Waite reduction by pi/2 in five
 * 33-bit pieces; r = j pi/128 + t (pi/128 in three pieces); sin and cos of j pi/128
 * from a double-double table; Taylor series of sin t / t and cos t in z = t^2 with
 * the leading terms in double-double; the angle-addition formulas; final rounding
 * of the double-double.  ORC_LIBM_DD selects it in
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
static dd_t dd_add_ord(dd_t big, dd_t small)
{
    dd_t s = fast_two_sum(big.hi, small.hi);
    s.lo += big.lo + small.lo;
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

/* sin and cos of j pi/128, j = 0..32: {sin hi, sin lo, cos hi, cos lo}
 * (cmc-cosmic_amd/csrc/gen_trig_table.py) */
static const double trig_tab[33][4] = {
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

void orc_dd_trig_table(int j, double *out4)
{
    for (int i = 0; i < 4; i++) out4[i] = (j >= 0 && j <= 32) ? trig_tab[j][i] : 0.0;
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
    const double H1 = 0x1.921fb54442d00p-6, H2 = 0x1.8469898cc5180p-54, H3 = -0x1.fc8f8cbb5bf6cp-103;
    const double j = rint(r.hi * 0x1.45f306dc9c883p+5);
    dd_t t = two_sum(r.hi - j * H1, -(j * H2));
    t = dd_add_d(t, r.lo);
    t = dd_add_d(t, -(j * H3));
    dd_t z = dd_mul(t, t);
    double ts = 0.0, tc = 0.0;
    for (int n = 6; n >= 4; n--) ts = ts * z.hi + ((n & 1) ? -inv_fact[2 * n + 1][0] : inv_fact[2 * n + 1][0]);
    for (int n = 6; n >= 4; n--) tc = tc * z.hi + ((n & 1) ? -inv_fact[2 * n][0] : inv_fact[2 * n][0]);
    dd_t S = {ts, 0.0}, Cc = {tc, 0.0};
    for (int n = 3; n >= 0; n--) S = dd_add_ord(coef(2 * n + 1, n & 1), dd_mul(S, z));
    for (int n = 3; n >= 0; n--) Cc = dd_add_ord(coef(2 * n, n & 1), dd_mul(Cc, z));
    S = dd_mul(S, t);
    const int aj = (int)fabs(j);
    dd_t Sj = {trig_tab[aj][0], trig_tab[aj][1]}, Cj = {trig_tab[aj][2], trig_tab[aj][3]};
    if (j < 0) { Sj.hi = -Sj.hi; Sj.lo = -Sj.lo; }
    dd_t sr = dd_add(dd_mul(Sj, Cc), dd_mul(Cj, S));
    dd_t p2 = dd_mul(Sj, S);
    p2.hi = -p2.hi; p2.lo = -p2.lo;
    dd_t cr2 = dd_add(dd_mul(Cj, Cc), p2);
    double srd = sr.hi + sr.lo, crd = cr2.hi + cr2.lo;
    switch ((int)((long long)k & 3)) {
    case 0: *s_out = srd; *c_out = crd; break;
    case 1: *s_out = crd; *c_out = -srd; break;
    case 2: *s_out = -srd; *c_out = -crd; break;
    default: *s_out = -crd; *c_out = srd; break;
    }
}
