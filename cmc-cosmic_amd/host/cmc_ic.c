/*
 * Seeded initial-condition generators for the benchmark clusters (host side).
 *
 * Algorithms follow the reference's legacy generators, restated:
 *   Plummer: radii from the cumulative mass, r = a/sqrt(u^(-2/3) - 1); speeds by
 *     von Neumann rejection on q^2 (1-q^2)^(7/2); isotropic directions
 *     (src/utils/mkplummer/mkplummer.c:151-170).
 *   King (1966) lowered isothermal model: the dimensionless Poisson equation
 *     W'' + 2W'/x = -9 rho(W)/rho(W0), rho(W) = e^W erf(sqrt W) - sqrt(4W/pi)(1+2W/3)
 *     (src/utils/mkking/mkking.c:61-86) integrated in FP64 with classical RK4 from
 *     the centre to W = 0; positions from the tabulated M(x); speeds by rejection
 *     from f ~ (e^{W - v^2/2} - 1) v^2 (mkking.c:446-458).
 *   Scaling to Henon units with the shell-potential recurrence of
 *     mkplummer.c:234-274 (PE = -1/2, KE = 1/4 after scaling).
 *   Kroupa IMF by inverse CDF per power-law segment (src/utils/setimf/setimf.c:343-355).
 * Random numbers: the taus113 generator (src/libs/taus113-v2.c) seeded with `seed`.
 */
#include "../../include/cmc_hotpath.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint32_t z[4]; } t113;

static uint32_t t113_int(t113 *s)
{
    uint32_t b;
    b = ((s->z[0] << 6) ^ s->z[0]) >> 13;  s->z[0] = ((s->z[0] & 4294967294u) << 18) ^ b;
    b = ((s->z[1] << 2) ^ s->z[1]) >> 27;  s->z[1] = ((s->z[1] & 4294967288u) << 2) ^ b;
    b = ((s->z[2] << 13) ^ s->z[2]) >> 21; s->z[2] = ((s->z[2] & 4294967280u) << 7) ^ b;
    b = ((s->z[3] << 3) ^ s->z[3]) >> 12;  s->z[3] = ((s->z[3] & 4294967168u) << 13) ^ b;
    return s->z[0] ^ s->z[1] ^ s->z[2] ^ s->z[3];
}
static double t113_dbl(t113 *s) { return t113_int(s) / 4294967296.0; }
static void t113_seed(uint64_t seed, t113 *s)
{
    uint64_t z;
    if (seed == 0) seed = 1;
    z = (69069ull * seed) & 0xffffffffull;  if (z < 2) z += 2;     s->z[0] = (uint32_t)z;
    z = (69069ull * s->z[0]) & 0xffffffffull; if (z < 8) z += 8;   s->z[1] = (uint32_t)z;
    z = (69069ull * s->z[1]) & 0xffffffffull; if (z < 16) z += 16; s->z[2] = (uint32_t)z;
    z = (69069ull * s->z[2]) & 0xffffffffull; if (z < 128) z += 128; s->z[3] = (uint32_t)z;
    for (int i = 0; i < 10; i++) t113_int(s);
}

typedef struct { double r, m, vr, vt, w; } rec_t;   /* w: King potential W at r (sigma^2 units) */

static int cmp_rec(const void *a, const void *b)
{
    double ra = ((const rec_t *)a)->r, rb = ((const rec_t *)b)->r;
    return ra < rb ? -1 : (ra > rb ? 1 : 0);
}

int cmc_ic_kroupa(int64_t n, uint64_t seed, double *m)
{
    if (n < 0 || (n > 0 && !m)) return CMCG_E_ARG;
    t113 rng;
    t113_seed(seed ^ 0x9e3779b97f4a7c15ull, &rng);
    const double a1 = 1.3, a2 = 2.3, lo = 0.08, mid = 0.5, hi = 150.0;
    /* continuity at 0.5: c2 = c1 * 0.5^(a2 - a1) */
    double w1 = (pow(lo, 1.0 - a1) - pow(mid, 1.0 - a1)) / (a1 - 1.0);
    double w2 = pow(mid, a2 - a1) * (pow(mid, 1.0 - a2) - pow(hi, 1.0 - a2)) / (a2 - 1.0);
    double p1 = w1 / (w1 + w2);
    for (int64_t i = 0; i < n; i++) {
        double u = t113_dbl(&rng), x = t113_dbl(&rng);
        double a, b, al;
        if (u < p1) { a = lo; b = mid; al = a1; } else { a = mid; b = hi; al = a2; }
        double ea = pow(a, 1.0 - al), eb = pow(b, 1.0 - al);
        m[i] = pow(ea + x * (eb - ea), 1.0 / (1.0 - al));
    }
    return CMCG_OK;
}

/* PE and KE of the sorted records with the shell recurrence (mkplummer.c:240-250);
 * masses sum to 1. */
static void energies(const rec_t *s, int64_t n, double *PE, double *KE)
{
    double U = 0.0, MM = 1.0, pe = 0.0, ke = 0.0;
    for (int64_t i = n - 1; i >= 0; i--) {
        double rn = i + 1 < n ? s[i + 1].r : 1.0e10;
        U -= MM * (1.0 / s[i].r - 1.0 / rn);
        MM -= s[i].m;
        pe += 0.5 * U * s[i].m;
        ke += 0.5 * (s[i].vr * s[i].vr + s[i].vt * s[i].vt) * s[i].m;
    }
    *PE = pe;
    *KE = ke;
}

/* relative potential psi_i = Phi(r_t) - Phi(r_i) of the sorted records with the
 * shell recurrence (masses normalised to their sum), r_t = rt */
static void rel_potential(const rec_t *s, int64_t n, double rt, double *psi)
{
    double msum = 0.0;
    for (int64_t i = 0; i < n; i++) msum += s[i].m;
    double U = 0.0, MM = 1.0;
    for (int64_t i = n - 1; i >= 0; i--) {
        double rn = i + 1 < n ? s[i + 1].r : rt;
        U -= MM * (1.0 / s[i].r - 1.0 / rn);
        MM -= s[i].m / msum;
        psi[i] = -U;            /* Phi(r_t) = -1/r_t cancels in the difference */
    }
}

static void finish(rec_t *s, int64_t n, cmcg_host_soa *out, double rcut_model, double *rtid, int potential_units)
{
    /* masses sum to one */
    double msum = 0.0;
    for (int64_t i = 0; i < n; i++) msum += s[i].m;
    for (int64_t i = 0; i < n; i++) s[i].m /= msum;
    qsort(s, (size_t)n, sizeof(rec_t), cmp_rec);
    double PE, KE;
    energies(s, n, &PE, &KE);
    /* virial scaling (PE = -1/2, KE = 1/4); for a model not in virial equilibrium
     * (mass-segregated) velocities scale with the potential unit instead, which keeps
     * every star's binding energy sign */
    double rfac = -PE * 2.0, vfac = potential_units ? 1.0 / sqrt(rfac) : 1.0 / sqrt(4.0 * KE);
    for (int64_t i = 0; i < n; i++) { s[i].r *= rfac; s[i].vr *= vfac; s[i].vt *= vfac; }
    out->n = n;
    out->r[0] = 1.0e-20; out->m[0] = 0.0; out->vr[0] = 0.0; out->vt[0] = 0.0; out->id[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        out->r[i + 1] = s[i].r;
        out->m[i + 1] = s[i].m * (double)n;
        out->vr[i + 1] = s[i].vr;
        out->vt[i + 1] = s[i].vt;
        out->id[i + 1] = i + 1;
    }
    out->r[n + 1] = 1.0e10; out->m[n + 1] = 0.0; out->vr[n + 1] = 0.0; out->vt[n + 1] = 0.0; out->id[n + 1] = 0;
    if (rtid) *rtid = rcut_model * rfac;
}

static int assign_masses(rec_t *s, int64_t n, uint64_t seed, int kroupa)
{
    if (!kroupa) {
        for (int64_t i = 0; i < n; i++) s[i].m = 1.0;
        return CMCG_OK;
    }
    double *m = (double *)malloc((size_t)n * sizeof(double));
    if (!m) return CMCG_E_NOMEM;
    cmc_ic_kroupa(n, seed, m);
    for (int64_t i = 0; i < n; i++) s[i].m = m[i];
    free(m);
    return CMCG_OK;
}

int cmc_ic_plummer(int64_t n, uint64_t seed, int kroupa, cmcg_host_soa *out, double *rtid)
{
    if (n < 2 || !out || !out->r || !out->m || !out->vr || !out->vt || !out->id) return CMCG_E_ARG;
    rec_t *s = (rec_t *)malloc((size_t)n * sizeof(rec_t));
    if (!s) return CMCG_E_NOMEM;
    t113 rng;
    t113_seed(seed, &rng);
    const double a = 3.0 * 3.14159265358979323 / 16.0;
    for (int64_t i = 0; i < n; i++) {
        double u;
        do { u = t113_dbl(&rng); } while (u <= 0.0 || u > 0.999999);
        double r = a / sqrt(pow(u, -2.0 / 3.0) - 1.0);
        double vesc = sqrt(2.0) * pow(r * r + a * a, -0.25);
        double q, g, X2;
        do {
            q = t113_dbl(&rng);
            X2 = t113_dbl(&rng);
            g = q * q * pow(1.0 - q * q, 3.5);
        } while (0.1 * X2 > g);
        double v = q * vesc;
        double X3 = t113_dbl(&rng);
        s[i].r = r;
        s[i].vr = (1.0 - 2.0 * X3) * v;
        s[i].vt = sqrt(v * v - s[i].vr * s[i].vr);
    }
    int e = assign_masses(s, n, seed, kroupa);
    if (e) { free(s); return e; }
    /* isolated model: a far tidal boundary (300 model scale lengths) */
    finish(s, n, out, 300.0 * a, rtid, 0);
    free(s);
    return CMCG_OK;
}

/* King ODE right-hand side, mkking.c:65-78 */
static double king_rho(double w)
{
    if (w <= 0.0) return 0.0;
    return exp(w) * erf(sqrt(w)) - sqrt(4.0 * w / 3.14159265358979323) * (1.0 + 2.0 * w / 3.0);
}

int cmc_ic_king(int64_t n, double w0, uint64_t seed, int kroupa, int segregate, cmcg_host_soa *out, double *rtid)
{
    if (n < 2 || !out || !out->r || !out->m || !out->vr || !out->vt || !out->id || !(w0 > 0.0) || w0 > 16.0)
        return CMCG_E_ARG;
    /* integrate y = (W, W') in x from 0 until W <= 0 */
    const double h = 1e-4;
    const int64_t maxsteps = 20000000;
    double rho0 = king_rho(w0);
    int64_t cap = 1 << 16, nt = 0;
    double *tx = (double *)malloc(cap * sizeof(double)), *tw = (double *)malloc(cap * sizeof(double)),
           *tm = (double *)malloc(cap * sizeof(double));
    if (!tx || !tw || !tm) { free(tx); free(tw); free(tm); return CMCG_E_NOMEM; }
    double x = 0.0, W = w0, Wp = 0.0;
#define KING_F(xx, ww, wp) ((xx) > 0.0 ? -9.0 * king_rho(ww) / rho0 - 2.0 * (wp) / (xx) : -3.0 * king_rho(ww) / rho0)
    tx[0] = 0.0; tw[0] = w0; tm[0] = 0.0; nt = 1;
    int64_t step = 0;
    const int64_t every = 10;
    while (W > 0.0 && step < maxsteps) {
        double k1w = Wp, k1p = KING_F(x, W, Wp);
        double k2w = Wp + 0.5 * h * k1p, k2p = KING_F(x + 0.5 * h, W + 0.5 * h * k1w, Wp + 0.5 * h * k1p);
        double k3w = Wp + 0.5 * h * k2p, k3p = KING_F(x + 0.5 * h, W + 0.5 * h * k2w, Wp + 0.5 * h * k2p);
        double k4w = Wp + h * k3p, k4p = KING_F(x + h, W + h * k3w, Wp + h * k3p);
        double Wn = W + h / 6.0 * (k1w + 2 * k2w + 2 * k3w + k4w);
        double Wpn = Wp + h / 6.0 * (k1p + 2 * k2p + 2 * k3p + k4p);
        double xn = x + h;
        step++;
        if (Wn <= 0.0) {
            /* linear interpolation to the zero crossing */
            double t = W / (W - Wn);
            x = x + t * h;
            Wp = Wp + t * (Wpn - Wp);
            W = 0.0;
        } else {
            x = xn; W = Wn; Wp = Wpn;
        }
        if (step % every == 0 || W <= 0.0) {
            if (nt == cap) {
                cap *= 2;
                tx = (double *)realloc(tx, cap * sizeof(double));
                tw = (double *)realloc(tw, cap * sizeof(double));
                tm = (double *)realloc(tm, cap * sizeof(double));
                if (!tx || !tw || !tm) { free(tx); free(tw); free(tm); return CMCG_E_NOMEM; }
            }
            tx[nt] = x; tw[nt] = W < 0.0 ? 0.0 : W; tm[nt] = -x * x * Wp;   /* M(x) ~ -x^2 W' */
            nt++;
        }
    }
#undef KING_F
    double xt = x, Mt = tm[nt - 1];
    for (int64_t i = 0; i < nt; i++) tm[i] /= Mt;
    /* enforce monotone cumulative mass for inversion */
    for (int64_t i = 1; i < nt; i++) if (tm[i] < tm[i - 1]) tm[i] = tm[i - 1];
    tm[nt - 1] = 1.0;

    rec_t *s = (rec_t *)malloc((size_t)n * sizeof(rec_t));
    if (!s) { free(tx); free(tw); free(tm); return CMCG_E_NOMEM; }
    t113 rng;
    t113_seed(seed, &rng);
    for (int64_t i = 0; i < n; i++) {
        double X;
        do { X = t113_dbl(&rng); } while (X <= 0.0);
        /* bisection in the cumulative mass table */
        int64_t lo = 0, hi = nt - 1;
        while (hi - lo > 1) {
            int64_t mid = (lo + hi) / 2;
            if (tm[mid] < X) lo = mid; else hi = mid;
        }
        double t = (tm[hi] > tm[lo]) ? (X - tm[lo]) / (tm[hi] - tm[lo]) : 0.0;
        double r = tx[lo] + t * (tx[hi] - tx[lo]);
        double psi = tw[lo] + t * (tw[hi] - tw[lo]);
        if (!(r > 0.0)) r = 0.5 * tx[1];
        if (!(psi > 1e-12)) psi = 1e-12;
        /* speed: f(v) ~ (e^{psi - v^2/2} - 1) v^2 on [0, sqrt(2 psi)], bound 2 e^{psi - 1} */
        double vmax = sqrt(2.0 * psi), F = 2.0 * exp(psi - 1.0), v0, f0, f;
        do {
            v0 = t113_dbl(&rng) * vmax;
            f0 = t113_dbl(&rng) * F;
            f = (exp(psi - v0 * v0 / 2.0) - 1.0) * v0 * v0;
        } while (f0 > f);
        double X3 = t113_dbl(&rng);
        s[i].r = r;
        s[i].w = psi;
        s[i].vr = (1.0 - 2.0 * X3) * v0;
        s[i].vt = sqrt(v0 * v0 - s[i].vr * s[i].vr);
        s[i].m = 1.0;
    }
    free(tx); free(tw); free(tm);
    qsort(s, (size_t)n, sizeof(rec_t), cmp_rec);
    int e = assign_masses(s, n, seed, kroupa);
    if (e) { free(s); return e; }
    if (segregate && n >= 100) {
        /* the heaviest 1% of masses go to radius ranks drawn uniformly (without
         * replacement) from the innermost 10%; the rest keep their random order */
        int64_t K = n / 100, inner = n / 10;
        int64_t *idx = (int64_t *)malloc((size_t)n * sizeof(int64_t));
        double *ms = (double *)malloc((size_t)n * sizeof(double));
        char *heavy_slot = (char *)calloc((size_t)n, 1);
        if (!idx || !ms || !heavy_slot) { free(idx); free(ms); free(heavy_slot); free(s); return CMCG_E_NOMEM; }
        for (int64_t i = 0; i < n; i++) ms[i] = s[i].m;
        /* threshold = K-th largest mass: sort a copy descending */
        double *sorted = (double *)malloc((size_t)n * sizeof(double));
        memcpy(sorted, ms, (size_t)n * sizeof(double));
        int cmpd(const void *a, const void *b);
        qsort(sorted, (size_t)n, sizeof(double), cmpd);
        /* inner ranks chosen by partial Fisher-Yates */
        for (int64_t i = 0; i < inner; i++) idx[i] = i;
        for (int64_t i = 0; i < K; i++) {
            int64_t j = i + (int64_t)(t113_dbl(&rng) * (double)(inner - i));
            if (j >= inner) j = inner - 1;
            int64_t tmp = idx[i]; idx[i] = idx[j]; idx[j] = tmp;
            heavy_slot[idx[i]] = 1;
        }
        /* heavy masses (descending) to the chosen slots; others in original order */
        int64_t h_i = 0, l_i = 0;
        double thr = sorted[K - 1];
        int64_t n_thr_heavy = 0;
        for (int64_t i = 0; i < K; i++) if (sorted[i] == thr) n_thr_heavy++;
        double *light = (double *)malloc((size_t)n * sizeof(double));
        int64_t nl = 0, taken_thr = 0;
        for (int64_t i = 0; i < n; i++) {
            if (ms[i] > thr) continue;
            if (ms[i] == thr && taken_thr < n_thr_heavy) { taken_thr++; continue; }
            light[nl++] = ms[i];
        }
        for (int64_t i = 0; i < n; i++) {
            if (heavy_slot[i]) s[i].m = sorted[h_i++];
            else s[i].m = light[l_i++];
        }
        free(light); free(sorted); free(idx); free(ms); free(heavy_slot);
        /* speeds follow the deeper segregated potential: keep each star's v^2 / (2 psi),
         * converting from sigma units (psi = W) to G = M = 1 units of the model radius */
        double *psi_seg = (double *)malloc((size_t)n * sizeof(double));
        if (!psi_seg) { free(s); return CMCG_E_NOMEM; }
        rel_potential(s, n, xt, psi_seg);
        for (int64_t i = 0; i < n; i++) {
            double f = s[i].w > 0.0 && psi_seg[i] > 0.0 ? sqrt(psi_seg[i] / s[i].w) : 0.0;
            s[i].vr *= f;
            s[i].vt *= f;
        }
        free(psi_seg);
    }
    finish(s, n, out, xt, rtid, segregate && n >= 100);
    free(s);
    return CMCG_OK;
}

int cmpd(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return x > y ? -1 : (x < y ? 1 : 0);   /* descending */
}
