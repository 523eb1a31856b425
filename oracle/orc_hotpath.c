#define _GNU_SOURCE
/*
 * Oracle restatement of the reference hot-path functions — TEST INFRASTRUCTURE ONLY.
 *
 * Plain C over the SoA arrays of orc_state.  Expressions are written in the same
 * operand order as the reference so that, with -ffp-contract=off, every double
 * result is the one the reference computes on the same inputs.  function_q keeps
 * the reference's x87 long double evaluation in ORC_Q_LONG_DOUBLE mode; the
 * ORC_Q_FP64 mode evaluates the same expression in double (the device semantics).
 *
 * Pinning (tests named in brackets):
 *  - potential_calculate: exact rational telescoped closed form on small N
 *    [test_oracle_potential_closed_form] and the Plummer analytic potential.
 *  - FindZero_r / potential(r): brute-force scan for the bracketing index
 *    [test_oracle_findzero_bruteforce]; interpolant exactness at nodes.
 *  - find_zero_Q / calc_orbit_new: brute-force Q over all k on tiny N gives
 *    kmin/kmax; Kepler special case (point mass) gives a(1-e), a(1+e)
 *    [test_oracle_apsides_*].
 *  - find_root_vr (GSL Brent in the reference; GSL is not in the reference tree
 *    nor in this image, so the Brent-Dekker iteration below is written from the
 *    published method): the root equals the exact root of the interval's quadratic
 *    in 1/r (50-digit) to the reference's tolerance, also on near-tangent orbits
 *    that reach it through calc_orbit_new [test_find_root_vr_*].
 *  - dynamics_apply relaxation: pair momentum and kinetic energy conservation;
 *    the angle between w and w_new recovered from the outputs equals beta, beta^2
 *    scales with dt, the clamp at pi/2 [test_relaxation_kick_angle*].
 *  - get_positions sampling: fixed-draw closed forms + KS test of accepted r
 *    against the 1/|v_r| density [test_oracle_sampling_ks].
 *  - sort: numpy.lexsort on (id, r) [test_oracle_sort].
 *  - simul_relax_new, calc_n_local, calc_sigma_r: direct window sums recomputed in
 *    exact rational arithmetic on small N [test_oracle_timestep_sigma].
 */
#include "cmc_oracle.h"
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>

/* cmc_utils.c:27-30, 39-42 */
double orc_sqr(double x) { return x * x; }
double orc_cub(double x) { return x * x * x; }

/* include/cmc/cmc.h:2027-2028: SQR(double) */
static inline double SQR(double a) { return a * a; }
/* include/cmc/cmc.h:2241-2242 */
#define OMAX(a, b) ((a) >= (b) ? (a) : (b))
#define OMIN(a, b) ((a) <= (b) ? (a) : (b))

static void set_error(orc_state *s, int code, int64_t idx)
{
    if (!s->error) { s->error = code; s->error_index = idx; }
}

/* include/cmc/cmc.h:2175 MPI_PHI_S(rad, j) in double context */
static inline double phi_s(const orc_state *s, double rad, int64_t j)
{
    return (rad >= s->r[j] ? s->m[j] / rad : s->m[j] / s->r[j]) * (1.0 / s->N_STAR);
}

/* cmc_nr.c:62-82 (the out-of-range dprintf is diagnostic only) */
int64_t orc_FindZero_r(const orc_state *s, int64_t kmin, int64_t kmax, double r)
{
    int64_t ktry;
    do {
        ktry = (kmin + kmax + 1) / 2;
        if (s->r[ktry] < r) kmin = ktry;
        else kmax = ktry - 1;
    } while (kmax != kmin);
    return kmin;
}

/* cmc_utils.c:112-165 with SEARCH_GRID = 0 */
double orc_potential(orc_state *s, double r)
{
    int64_t i, kmin, kmax;
    if (r == 0.0) return s->phi[0];
    if (r < s->r[1]) return s->phi[0] - (s->phi[0] - s->phi[1]) * s->r[1] / r;
    kmax = s->N_MAX + 1;
    kmin = 1;
    if (kmin == kmax - 1) i = kmin;
    else i = orc_FindZero_r(s, kmin, kmax, r);
    if (s->r[i] > r || s->r[i + 1] < r) {
        set_error(s, ORC_E_BRACKET, i);
        return NAN;
    }
    if (i == 0) return s->phi[1];
    return s->phi[i] + (s->phi[i + 1] - s->phi[i]) * (1.0 / s->r[i] - 1.0 / r) /
                           (1.0 / s->r[i] - 1.0 / s->r[i + 1]);
}

/* cmc_utils.c:675-736: N_MAX, Mtotal, Rtidal, sentinels, reverse recurrence */
int64_t orc_potential_calculate(orc_state *s)
{
    int64_t k = 1;
    double mprev = 0.0;
    while (s->r[k] < ORC_SF_INFINITY && k <= s->N_MAX) {
        mprev += s->m[k];
        if (isnan(mprev)) { set_error(s, ORC_E_NAN, k); return -1; }
        k++;
    }
    s->N_MAX = k - 1;
    s->Mtotal = mprev * s->madhoc + s->cenma_m * s->madhoc;
    s->Rtidal = s->orbit_r * pow(s->Mtotal, 1.0 / 3.0);
    s->r[s->N_MAX + 1] = ORC_SF_INFINITY;
    s->phi[s->N_MAX + 1] = 0.0;
    mprev = s->Mtotal;
    for (k = s->N_MAX; k >= 1; k--) {
        s->phi[k] = s->phi[k + 1] - mprev * (1.0 / s->r[k] - 1.0 / s->r[k + 1]);
        mprev -= s->m[k] / s->N_STAR;
        if (isnan(s->phi[k])) { set_error(s, ORC_E_NAN, k); return -1; }
    }
    s->phi[0] = s->phi[1] + s->cenma_m * s->madhoc / s->r[1];
    return s->N_MAX;
}

/* cmc_utils.c:1654-1662 function_q: in the reference E, J, r, pot are long double,
 * (J/r) is squared by the double SQR inline, and the long double result is
 * returned as double. */
static double fq_ld(const orc_state *s, int64_t j, long double r, long double pot,
                    long double E, long double J)
{
    double res;
    long double Jr, phis;
    Jr = SQR((J) / (r));
    phis = ((r) >= s->r[j] ? s->m[j] / (r) : s->m[j] / s->r[j]) * (1.0 / s->N_STAR);
    res = (2.0 * ((E) - (pot + phis)) - Jr);
    return res;
}

static double fq_d(const orc_state *s, int64_t j, double r, double pot, double E, double J)
{
    double Jr = SQR(J / r);
    double phis = (r >= s->r[j] ? s->m[j] / r : s->m[j] / s->r[j]) * (1.0 / s->N_STAR);
    return 2.0 * (E - (pot + phis)) - Jr;
}

double orc_function_q(const orc_state *s, int64_t j, double r, double pot, double E, double J)
{
    if (s->q_mode == ORC_Q_LONG_DOUBLE) return fq_ld(s, j, r, pot, E, J);
    return fq_d(s, j, r, pot, E, J);
}

/* cmc_nr.c:316-383.  The reference loops until kmax == kmin; a bracket that can
 * never close (kmin == kmax on entry with the wrong sign) would loop forever
 * there, here it stops after 200 halvings and reports ORC_E_BRACKET. */
int64_t orc_find_zero_Q(const orc_state *s, int64_t j, int64_t kmin, int64_t kmax, double E, double J)
{
    int64_t ktry;
    double qmin, qmax, qtry;
    int guard = 0;
    qmin = orc_function_q(s, j, s->r[kmin], s->phi[kmin], E, J);
    qmax = orc_function_q(s, j, s->r[kmax], s->phi[kmax], E, J);
    if (qmin < qmax) {
        do {
            ktry = (kmin + kmax + 1) / 2;
            qtry = orc_function_q(s, j, s->r[ktry], s->phi[ktry], E, J);
            if (qtry < 0.e0) kmin = ktry;
            else kmax = ktry - 1;
        } while (kmax != kmin && ++guard < 200);
    } else {
        do {
            ktry = (kmin + kmax + 1) / 2;
            qtry = orc_function_q(s, j, s->r[ktry], s->phi[ktry], E, J);
            if (qtry > 0.e0) kmin = ktry;
            else kmax = ktry - 1;
        } while (kmax != kmin && ++guard < 200);
    }
    if (guard >= 200) return -1;
    return kmin;
}

/* cmc_orbit.c:30-59 (no newly created stars at one rank) */
int64_t orc_get_positive_Q_index(orc_state *s, int64_t index, double E, double J)
{
    int64_t ktemp = index;
    double Qtemp = orc_function_q(s, index, s->r[ktemp], s->phi[ktemp], E, J);
    if (Qtemp <= 0.0) {
        s->n_linear_scan++;
        ktemp = -1;
        do {
            ktemp++;
            Qtemp = orc_function_q(s, index, s->r[ktemp], s->phi[ktemp], E, J);
        } while (Qtemp <= 0.0 && ktemp <= s->N_MAX);
        if (ktemp >= s->N_MAX) return -1;
    }
    return ktemp;
}

/* cmc_orbit.c:722-740 */
void orc_set_a_b(const orc_state *s, int64_t index, int64_t k, double *a, double *b)
{
    if (k == 0) {
        *a = -s->cenma_m * s->madhoc;
        *b = s->phi[1] + s->cenma_m * s->madhoc / s->r[1] + s->m[index] * s->madhoc / s->r[index];
    } else {
        double rk = s->r[k], rk1 = s->r[k + 1];
        double Uk = s->phi[k] + phi_s(s, rk, index);
        double Uk1 = s->phi[k + 1] + phi_s(s, rk1, index);
        *a = (Uk1 - Uk) / (1. / rk1 - 1. / rk);
        *b = (Uk / rk1 - Uk1 / rk) / (1. / rk1 - 1. / rk);
    }
}

/* cmc_nr.c:239-269 calc_pot_within_interval and :393-403 calc_vr_within_interval */
static double vr2_in_interval(const orc_state *s, int64_t index, int64_t k, double E, double J, double r)
{
    double pot;
    if (fabs(r - s->r[k]) < DBL_EPSILON) pot = s->phi[k];
    else if (fabs(r - s->r[k + 1]) < DBL_EPSILON) pot = s->phi[k + 1];
    else if (r < s->r[1]) pot = s->phi[0] - s->cenma_m * s->madhoc / r;
    else pot = (s->phi[k] + (s->phi[k + 1] - s->phi[k]) * (1.0 / s->r[k] - 1.0 / r) /
                                (1.0 / s->r[k] - 1.0 / s->r[k + 1]));
    return orc_function_q(s, index, r, pot, E, J);
}

/* Interval convergence test with the semantics documented for GSL's
 * gsl_root_test_interval: |hi - lo| < epsabs + epsrel * min(|lo|, |hi|) when the
 * interval does not contain 0 (else min term = 0). */
static int interval_converged(double lo, double hi, double epsabs, double epsrel)
{
    double alo = fabs(lo), ahi = fabs(hi), min_abs;
    if ((lo > 0.0 && hi > 0.0) || (lo < 0.0 && hi < 0.0)) min_abs = alo < ahi ? alo : ahi;
    else min_abs = 0.0;
    return fabs(hi - lo) < epsabs + epsrel * min_abs;
}

/* cmc_nr.c:500-603 find_root_vr: root of v_r^2(r) in [r_k, r_{k+1}] by the
 * Brent-Dekker method (R. P. Brent, Algorithms for Minimization without
 * Derivatives, 1973, ch. 4), APSIDES_PRECISION = 1e-11, APSIDES_MAX_ITER = 100
 * (cmc_io.c:1752-1754), then the residual fix-up of cmc_nr.c:587-600.  GSL's own
 * iterate path is not reproduced (GSL is absent); the root is pinned to the exact
 * root within the reference's tolerance. */
double orc_find_root_vr(orc_state *s, int64_t index, int64_t k, double E, double J)
{
    const double prec = 1e-11;
    int iter = 100;
    double a = s->r[k], b = s->r[k + 1];
    double fa = vr2_in_interval(s, index, k, E, J, a);
    double fb = vr2_in_interval(s, index, k, E, J, b);
    double c, fc, d, e;
    double lo, hi;
    s->n_brent++;
    if ((fa > 0) == (fb > 0) && (fa < 0) == (fb < 0)) { set_error(s, ORC_E_BRENT, index); return NAN; }
    c = a; fc = fa; d = b - a; e = d;
    lo = a < b ? a : b; hi = a < b ? b : a;
    int converged = 0;
    double prev_apsis = -1.0;
    while (!converged && iter-- > 0) {
        /* keep b the best estimate and [b, c] a bracket */
        if ((fb > 0 && fc > 0) || (fb < 0 && fc < 0)) { c = a; fc = fa; d = b - a; e = d; }
        if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
        double tol = 2.0 * DBL_EPSILON * fabs(b);
        double m = 0.5 * (c - b);
        if (fb == 0.0) { lo = hi = b; converged = 1; break; }
        if (fabs(e) < tol || fabs(fa) <= fabs(fb)) {
            d = m; e = m;                         /* bisection */
        } else {
            double p, q, r2, s2 = fb / fa;
            if (a == c) {                         /* secant */
                p = 2.0 * m * s2; q = 1.0 - s2;
            } else {                              /* inverse quadratic interpolation */
                q = fa / fc; r2 = fb / fc;
                p = s2 * (2.0 * m * q * (q - r2) - (b - a) * (r2 - 1.0));
                q = (q - 1.0) * (r2 - 1.0) * (s2 - 1.0);
            }
            if (p > 0) q = -q; else p = -p;
            if (2.0 * p < fmin(3.0 * m * q - fabs(tol * q), fabs(e * q))) { e = d; d = p / q; }
            else { d = m; e = m; }
        }
        a = b; fa = fb;
        if (fabs(d) > tol) b += d;
        else b += (m > 0 ? tol : -tol);
        fb = vr2_in_interval(s, index, k, E, J, b);
        lo = b < c ? b : c; hi = b < c ? c : b;
        if ((fb > 0 && fc > 0) || (fb < 0 && fc < 0)) { lo = a < b ? a : b; hi = a < b ? b : a; }
        converged = interval_converged(lo, hi, prec, prec);
        /* cmc_nr.c:552-566: in the last ten iterations also accept two successive
         * estimates within APSIDES_CONVERGENCE = 5e-13 (documented test_delta). */
        if (!converged && iter < 10) {
            if (prev_apsis >= 0.0 && fabs(b - prev_apsis) < 5e-13 + 5e-13 * fabs(b)) converged = 1;
            prev_apsis = b;
        }
    }
    if (!converged) { set_error(s, ORC_E_BRENT, index); }
    double apsis = b;
    if (vr2_in_interval(s, index, k, E, J, apsis) < 0.) {
        if (vr2_in_interval(s, index, k, E, J, lo) < 0.) apsis = hi;
        if (vr2_in_interval(s, index, k, E, J, hi) < 0.) apsis = lo;
    }
    if (vr2_in_interval(s, index, k, E, J, apsis) < 0.) { set_error(s, ORC_E_BRENT, index); }
    return apsis;
}

/* cmc_orbit.c:413-540 calc_orbit_new (also known as calc_pericenter_and_apocenter) */
orc_orbit_rs orc_calc_orbit_new(orc_state *s, int64_t index, double E, double J)
{
    orc_orbit_rs o;
    int64_t ktemp, kmin, kmax;
    double rmin, rmax, a, b, dQdr_min, dQdr_max, inside_sqrt;
    int circular = 0;

    ktemp = orc_get_positive_Q_index(s, index, E, J);
    if (ktemp <= -1) {
        o.rp = o.ra = s->r[index];
        o.kmin = o.kmax = index;
        o.dQdrp = o.dQdra = 0.0;
        o.circular_flag = 1;
        return o;
    }
    kmin = orc_find_zero_Q(s, index, 0, ktemp, E, J);
    kmax = orc_find_zero_Q(s, index, ktemp, s->N_MAX + 1, E, J);
    if (kmin < 0 || kmax < 0) {
        set_error(s, ORC_E_BRACKET, index);
        o.rp = o.ra = s->r[index]; o.kmin = o.kmax = index; o.dQdrp = o.dQdra = 0.0; o.circular_flag = 1;
        return o;
    }

    orc_set_a_b(s, index, kmin, &a, &b);
    inside_sqrt = a * a - 2.0 * J * J * (b - E);
    if (inside_sqrt < 0.0) {
        rmin = -1.0 * J * J / a;
        if ((rmin < s->r[kmin]) || (rmin > s->r[kmin + 1])) {
            rmin = orc_find_root_vr(s, index, kmin, E, J);
            if (s->flags) s->flags[index] |= ORC_F_BRENT;
        }
    } else {
        rmin = J * J / (-a + sqrt(inside_sqrt));
    }
    dQdr_min = 2.0 * J * J / (rmin * rmin * rmin) + 2.0 * a / (rmin * rmin);

    orc_set_a_b(s, index, kmax, &a, &b);
    inside_sqrt = a * a - 2.0 * J * J * (b - E);
    if (inside_sqrt < 0.0) {
        rmax = -a / (2.0 * (b - E));
        if ((rmax < s->r[kmax]) || (rmax > s->r[kmax + 1])) {
            rmax = orc_find_root_vr(s, index, kmax, E, J);
            if (s->flags) s->flags[index] |= ORC_F_BRENT;
        }
    } else {
        rmax = (-a + sqrt(inside_sqrt)) / (2.0 * (b - E));
    }
    dQdr_max = 2.0 * J * J / (rmax * rmax * rmax) + 2.0 * a / (rmax * rmax);

    if ((rmin > rmax) || (dQdr_min <= 0.) || (dQdr_max >= 0.)) circular = 1;

    if (circular) {
        o.rp = o.ra = s->r[index];
        o.kmin = o.kmax = index;
        o.dQdrp = o.dQdra = 0.0;
        o.circular_flag = 1;
    } else {
        o.rp = rmin; o.ra = rmax; o.kmin = kmin; o.kmax = kmax;
        o.dQdrp = dQdr_min; o.dQdra = dQdr_max; o.circular_flag = 0;
    }
    return o;
}

/* cmc_evolution_thr.c:655-716 remove_star (logging omitted) */
static void remove_star(orc_state *s, int64_t j)
{
    double E = s->E[j], J = s->J[j], m = s->m[j];
    s->rnew[j] = ORC_SF_INFINITY;
    s->vrnew[j] = 0.0;
    s->vtnew[j] = 0.0;
    s->Eescaped += E * m / s->N_STAR;
    s->Jescaped += J * m / s->N_STAR;
    s->TidalMassLoss += m / s->N_STAR;
    s->Etidal += E * m / s->N_STAR;
}

/* cmc_dynamics_helper.c:131-153 destroy_obj: zero_star keeps r and phi, then
 * remove_star_center (cmc_evolution_thr.c:766-771). */
static void destroy_obj(orc_state *s, int64_t j)
{
    s->vr[j] = s->vt[j] = s->E[j] = s->J[j] = 0.0;
    s->id[j] = 0;
    s->X[j] = s->Y[j] = s->r_peri[j] = s->r_apo[j] = 0.0;
    s->rOld[j] = s->Uoldrold[j] = s->Uoldrnew[j] = s->vrold[j] = s->vtold[j] = s->EI[j] = 0.0;
    s->rnew[j] = ORC_SF_INFINITY;
    s->m[j] = DBL_MIN;
    s->vrnew[j] = 0.0;
    s->vtnew[j] = 0.0;
}

/* cmc_evolution_thr.c:784-937 get_positions_loop and :944-1023 get_positions */
double orc_get_positions(orc_state *s)
{
    double max_rad = 0.0;
    orc_t113 cur = {{0, 0, 0, 0}};     /* substream of slot j: slot 1's, jumped 2^20 per slot */
    const int64_t j_lo = s->range_hi > 0 ? s->range_lo : 1;
    const int64_t j_hi = s->range_hi > 0 ? s->range_hi : s->N_MAX_NEW;
    for (int64_t j = j_lo; j <= j_hi; j++) {
        double E, J, rmin, rmax, dQdr_min, dQdr_max, g1, g2, F, s0, g0, drds, pot, Q;
        double r = 0.0, vr = 0.0, X = 0.0;
        orc_orbit_rs o;
        orc_t113 local, *rng;
        int64_t k;

        if (s->ntry) s->ntry[j] = 0;
        if (s->flags) s->flags[j] = 0;
        if (s->pot_rnew) s->pot_rnew[j] = 0.0;
        if (s->rng_mode != ORC_RNG_REF_SEQ)
            cur = j == j_lo ? orc_stream_orbit(&s->seeded, s->step, (uint64_t)j) : orc_t113_jump_pow2(cur, 20);
        E = s->E[j] + phi_s(s, s->r[j], j);
        J = s->J[j];

        if (s->m[j] < ORC_ZERO) {
            destroy_obj(s, j);
            s->n_removed_mass++;
            if (s->flags) s->flags[j] |= ORC_F_DESTROYED;
            continue;
        }
        if (E >= 0.0) {
            remove_star(s, j);
            s->n_removed_E++;
            if (s->flags) s->flags[j] |= ORC_F_REMOVED_E;
            continue;
        }
        o = orc_calc_orbit_new(s, j, E, J);
        if (o.circular_flag == 1) {
            s->rnew[j] = s->r[j];
            s->vrnew[j] = s->vr[j];
            s->vtnew[j] = s->vt[j];
            s->n_circular++;
            if (s->flags) s->flags[j] |= ORC_F_CIRCULAR;
            continue;
        }
        rmin = o.rp; rmax = o.ra; dQdr_min = o.dQdrp; dQdr_max = o.dQdra;
        if (rmax >= s->Rtidal) {
            s->r_apo[j] = rmax;
            s->r_peri[j] = rmin;
            remove_star(s, j);
            s->n_removed_tidal++;
            if (s->flags) s->flags[j] |= ORC_F_REMOVED_RT;
            continue;
        }
        g1 = sqrt(3.0 * (rmax - rmin) / dQdr_min);
        g2 = sqrt(-3.0 * (rmax - rmin) / dQdr_max);
        F = 1.2 * OMAX(g1, g2);

        if (s->rng_mode == ORC_RNG_REF_SEQ) rng = &s->seq;
        else { local = cur; rng = &local; }

        for (k = 1; k <= ORC_N_TRY; k++) {
            X = orc_t113_dbl(rng);
            s0 = 2.0 * X - 1.0;
            g0 = F * orc_t113_dbl(rng);
            r = 0.5 * (rmin + rmax) + 0.25 * (rmax - rmin) * (3.0 * s0 - s0 * s0 * s0);
            pot = orc_potential(s, r) + phi_s(s, r, j);
            drds = 0.25 * (rmax - rmin) * (3.0 - 3.0 * s0 * s0);
            Q = 2.0 * E - 2.0 * pot - J * J / r / r;
            if (Q >= 0.0) {
                vr = sqrt(Q);
            } else {
                if (isnan(Q)) { set_error(s, ORC_E_NAN, j); return -1.0; }
                vr = 0;
            }
            if (g0 < 1.0 / vr * drds) break;
        }
        if (k == ORC_N_TRY + 1) { set_error(s, ORC_E_NTRY, j); return -1.0; }
        if (s->ntry) s->ntry[j] = (int32_t)k;
        if (s->pot_rnew) s->pot_rnew[j] = pot;

        s->X[j] = X;
        s->r_peri[j] = rmin;
        s->r_apo[j] = rmax;
        if (orc_t113_dbl(rng) < 0.5) vr = -vr;
        s->rnew[j] = r;
        s->vrnew[j] = vr;
        s->vtnew[j] = J / r;
        if (r > max_rad) max_rad = r;
    }
    s->max_r = max_rad;
    return max_rad;
}

/* cmc_dynamics_helper.c:768-786 */
double orc_calc_n_local(const orc_state *s, int64_t k, int64_t p, int64_t N_LIMIT)
{
    int64_t kmin = k - p, kmax = k + p;
    if (kmin < 0) { kmin = 0; kmax = 2 * p + 1; }
    else if (kmax > N_LIMIT) { kmax = N_LIMIT; kmin = N_LIMIT - 2 * p - 1; }
    /* below 2p+2 stars the reference indexes outside [0, N+1] (undefined); both this
     * restatement and the device clamp to the array there */
    if (kmax > N_LIMIT + 1) kmax = N_LIMIT + 1;
    if (kmin < 0) kmin = 0;
    return (2.0 * ((double)p)) * 3.0 / (4.0 * ORC_PI * (orc_cub(s->r[kmax]) - orc_cub(s->r[kmin])));
}

/* cmc_dynamics_helper.c:2821-2986 at one rank (no ghost exchange; myid == 0). */
void orc_calc_sigma_r(orc_state *s)
{
    int64_t p = s->avekernel, N = s->N_MAX, si, k, simin, simax, siminlast = 1, simaxlast = 0;
    double Mv2ave = 0.0, Mave = 0.0;
    for (si = 1; si <= N; si++) {
        simin = si - p;
        simax = simin + (2 * p - 1);
        if (simin < 1) { simin = 1; simax = simin + (2 * p - 1); }
        else if (simax > s->N_MAX) { simax = N; simin = simax - (2 * p - 1); }
        if (simin < 1) simin = 1;           /* N < 2p: window clamped to the live stars, as on the device */
        if (simax > N) simax = N;
        for (k = siminlast; k < simin; k++) {
            Mv2ave -= s->m[k] * s->madhoc * (orc_sqr(s->vr[k]) + orc_sqr(s->vt[k]));
            Mave -= s->m[k] * s->madhoc;
        }
        for (k = simaxlast + 1; k <= simax; k++) {
            Mv2ave += s->m[k] * s->madhoc * (orc_sqr(s->vr[k]) + orc_sqr(s->vt[k]));
            Mave += s->m[k] * s->madhoc;
        }
        if (s->sigma) s->sigma[si] = sqrt(Mv2ave / Mave);
        siminlast = simin;
        simaxlast = simax;
    }
}

/* cmc_dynamics_helper.c:2526-2579 */
double orc_simul_relax_new(const orc_state *s)
{
    int64_t p = s->avekernel, si, k, simin, simax;
    double dt, dtmin = INFINITY, W, n_local, Mv2ave, Mave, M2ave, sigma;
    for (si = 1 + p; si < s->N_MAX - p; si += 2 * p) {
        simin = si - p;
        simax = simin + (2 * p - 1);
        Mv2ave = 0.0; Mave = 0.0; M2ave = 0.0;
        for (k = simin; k <= simax; k++) {
            double tmp = s->m[k] * s->madhoc;
            Mv2ave += tmp * (orc_sqr(s->vr[k]) + orc_sqr(s->vt[k]));
            Mave += tmp;
            M2ave += orc_sqr(tmp);
        }
        Mv2ave /= (double)(2 * p);
        Mave /= (double)(2 * p);
        M2ave /= (double)(2 * p);
        sigma = sqrt(Mv2ave / Mave);
        W = 4.0 * sigma / sqrt(3.0 * ORC_PI);
        n_local = orc_calc_n_local(s, si, p, s->N_MAX);
        dt = orc_sqr(2.0 * s->theta_se_max / ORC_PI) * (ORC_PI / 32.0) * orc_cub(W) /
             (((double)s->N_STAR) * n_local * (4.0 * M2ave));
        dtmin = OMIN(dtmin, dt);
    }
    return dtmin;
}

/* sin/cos of the relaxation.  ORC_LIBM_SYSTEM calls the C library's sin() and cos()
 * separately, as the reference's documented (unoptimised) build does;
 * ORC_LIBM_SINCOS calls sincos() for each same-argument pair, as an optimising gcc
 * compiles the reference's source; ORC_LIBM_DD uses the double-double functions
 * that the device uses (orc_trig.c); ORC_LIBM_LONG_DOUBLE rounds the long double
 * functions to double.  The last two and SINCOS are last-bit variants used to
 * measure how sensitive the evolution is to the math library.  The library
 * functions are called through volatile pointers so that the compiler cannot merge
 * them. */
static double (*volatile libm_sin)(double) = sin;
static double (*volatile libm_cos)(double) = cos;
static void (*volatile libm_sincos)(double, double *, double *) = sincos;
static double o_sin(const orc_state *s, double x)
{
    double sn, cs;
    if (s->libm_mode == ORC_LIBM_DD) { orc_dd_sincos(x, &sn, &cs); return sn; }
    if (s->libm_mode == ORC_LIBM_SINCOS) { libm_sincos(x, &sn, &cs); return sn; }
    return s->libm_mode == ORC_LIBM_LONG_DOUBLE ? (double)sinl((long double)x) : libm_sin(x);
}
static double o_cos(const orc_state *s, double x)
{
    double sn, cs;
    if (s->libm_mode == ORC_LIBM_DD) { orc_dd_sincos(x, &sn, &cs); return cs; }
    if (s->libm_mode == ORC_LIBM_SINCOS) { libm_sincos(x, &sn, &cs); return cs; }
    return s->libm_mode == ORC_LIBM_LONG_DOUBLE ? (double)cosl((long double)x) : libm_cos(x);
}

/* cmc_dynamics.c:17-542, relaxation branch (:427-498) with the configs' switches:
 * no binaries, no collisions (S = 0 so P_enc = 0), no three-body binaries, no
 * central mass.  Three draws per pair: Y (cmc_dynamics_helper.c:839), the P_enc test
 * (cmc_dynamics.c:406) and psi (:465). */
void orc_dynamics_apply(orc_state *s, double dt)
{
    int64_t p = s->avekernel, N = s->N_MAX, si = 1, pair = 0;
    /* pair p's substream starts 3p draws into the step's stream: walk it */
    int64_t si_end = N - N % 2 - 1;
    orc_t113 cur;
    if (s->range_hi > 0) {                   /* one rank's pairs: its chunk holds whole pairs */
        si = s->range_lo;
        if (s->range_hi < si_end + 1) si_end = s->range_hi - 1;
        pair = (si - 1) / 2;
    }
    cur = orc_stream_relax(&s->seeded, s->step, (uint64_t)pair);
    if (!s->skip_sigma) orc_calc_sigma_r(s);
    while (si <= si_end) {
        int64_t k = si, kp = si + 1;
        double v[4], vp[4], w[4], w1[4], w2[4], w_new[4], v_new[4], vp_new[4];
        double W, phi, n_local, mass_k, mass_kp, P_enc, Trel12, beta, wp, psi;
        orc_t113 local, *rng;
        int j;
        si += 2;
        if (s->rng_mode == ORC_RNG_REF_SEQ) rng = &s->seq;
        else {
            local = cur; rng = &local;
            orc_t113_int(&cur); orc_t113_int(&cur); orc_t113_int(&cur);
        }
        pair++;

        /* calc_encounter_dyns, cmc_dynamics_helper.c:831-871 */
        s->Y[k] = orc_t113_dbl(rng);
        s->Y[kp] = s->Y[k];
        phi = s->Y[k] * 2.0 * ORC_PI;
        v[1] = s->vt[k]; v[2] = 0.0; v[3] = s->vr[k];
        vp[1] = s->vt[kp] * o_cos(s, phi);
        vp[2] = s->vt[kp] * o_sin(s, phi);
        vp[3] = s->vr[kp];
        for (j = 1; j <= 3; j++) w[j] = vp[j] - v[j];
        W = sqrt(orc_sqr(w[1]) + orc_sqr(w[2]) + orc_sqr(w[3]));
        if (W == 0.0) { set_error(s, ORC_E_WZERO, k); return; }

        n_local = orc_calc_n_local(s, k, p, N);
        mass_k = s->m[k];
        mass_kp = s->m[kp];
        P_enc = 0.0; /* S = 0 with binaries and collisions off (cmc_dynamics.c:391-398) */
        if (orc_t113_dbl(rng) < P_enc) continue; /* never taken */

        Trel12 = (ORC_PI / 32.0) * orc_cub(W) /
                 (((double)s->N_STAR) * n_local * orc_sqr((mass_k + mass_kp) * s->madhoc));
        beta = (ORC_PI / 2.0) * sqrt(dt / Trel12);
        if (beta > ORC_PI / 2.0) beta = ORC_PI / 2.0;
        wp = sqrt(orc_sqr(w[1]) + orc_sqr(w[2]));
        if (wp == 0.0) { set_error(s, ORC_E_WPZERO, k); return; }
        w1[1] = -w[2] * W / wp;
        w1[2] = w[1] * W / wp;
        w1[3] = 0.0;
        w2[1] = -w[1] * w[3] / wp;
        w2[2] = -w[2] * w[3] / wp;
        w2[3] = wp;
        psi = orc_t113_dbl(rng) * 2 * ORC_PI;
        for (j = 1; j <= 3; j++)
            w_new[j] = w[j] * o_cos(s, beta) + w1[j] * o_sin(s, beta) * o_cos(s, psi) + w2[j] * o_sin(s, beta) * o_sin(s, psi);
        for (j = 1; j <= 3; j++) {
            v_new[j] = v[j] - mass_kp / (mass_k + mass_kp) * (w_new[j] - w[j]);
            vp_new[j] = vp[j] + mass_k / (mass_k + mass_kp) * (w_new[j] - w[j]);
        }
        s->vr[k] = v_new[3];
        s->vt[k] = sqrt(orc_sqr(v_new[1]) + orc_sqr(v_new[2]));
        s->vr[kp] = vp_new[3];
        s->vt[kp] = sqrt(orc_sqr(vp_new[1]) + orc_sqr(vp_new[2]));
        /* set_star_EJ, cmc_dynamics_helper.c:878-882 (old phi, old r) */
        s->E[k] = s->phi[k] + 0.5 * (orc_sqr(s->vr[k]) + orc_sqr(s->vt[k]));
        s->J[k] = s->r[k] * s->vt[k];
        s->E[kp] = s->phi[kp] + 0.5 * (orc_sqr(s->vr[kp]) + orc_sqr(s->vt[kp]));
        s->J[kp] = s->r[kp] * s->vt[kp];
    }
}

/* cmc_sort.c:245-250 compare_type extended with an id tie-break, which the device sort
 * also applies (libc qsort is unstable, so the reference leaves equal-r order open). */
static const orc_state *g_sort_state;
static int cmp_r_id(const void *pa, const void *pb)
{
    int64_t a = *(const int64_t *)pa, b = *(const int64_t *)pb;
    double ra = g_sort_state->r[a], rb = g_sort_state->r[b];
    if (ra < rb) return -1;
    if (ra > rb) return 1;
    int64_t ia = g_sort_state->id[a], ib = g_sort_state->id[b];
    if (ia < ib) return -1;
    if (ia > ib) return 1;
    return 0;
}

static void permute_d(double *a, const int64_t *perm, int64_t n, double *tmp)
{
    if (!a) return;
    for (int64_t i = 0; i < n; i++) tmp[i] = a[perm[i]];
    memcpy(a + 1, tmp, (size_t)n * sizeof(double));
}

/* qsorts_new (cmc_utils.c:2232-2254) -> sample_sort at P=1 (cmc_sort.c:375-412:
 * local qsort of whole records, then remove_stripped_stars :343-354), followed by
 * post_sort_comm (cmc_utils.c:2259-2288).  Returns the new N_MAX. */
int64_t orc_sort_by_r(orc_state *s)
{
    int64_t n = s->N_MAX_NEW;
    int64_t *perm = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    double *tmp = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    int64_t *tmpi = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) perm[i] = i + 1;
    g_sort_state = s;
    qsort(perm, (size_t)n, sizeof(int64_t), cmp_r_id);
    double *fields[] = { s->r, s->m, s->vr, s->vt, s->E, s->J, s->rnew, s->vrnew, s->vtnew,
                         s->X, s->Y, s->r_peri, s->r_apo, s->rOld, s->Uoldrold, s->Uoldrnew,
                         s->vrold, s->vtold, s->EI };
    for (size_t f = 0; f < sizeof(fields) / sizeof(fields[0]); f++) permute_d(fields[f], perm, n, tmp);
    for (int64_t i = 0; i < n; i++) tmpi[i] = s->id[perm[i]];
    memcpy(s->id + 1, tmpi, (size_t)n * sizeof(int64_t));
    /* remove_stripped_stars: drop the r == SF_INFINITY tail */
    while (n > 0 && s->r[n] == ORC_SF_INFINITY) n--;
    s->N_MAX_NEW = n;
    s->N_MAX = n;
    free(perm); free(tmp); free(tmpi);
    return n;
}
