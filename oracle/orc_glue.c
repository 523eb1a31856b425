/*
 * Oracle restatement of the Henon-step glue around the hot path — TEST
 * INFRASTRUCTURE ONLY.  One rank (P = 1), the configs' switch set: PERTURB =
 * RELAXATION = 1, TIDALLY_STRIP_STARS = 1 with TIDAL_TREATMENT = 0, stellar
 * evolution, binaries, collisions and the central black hole off.
 *
 * Pinning: energy bookkeeping is checked through conservation over a step
 * (total energy change against the reference's own Etot + Eoops definition,
 * cmc_io.c:2347) and by the per-field parity tests of the device path; the
 * excess-energy chain of set_velocities3 is pinned by hand-built cases whose
 * result is known in closed form [test_oracle_set_velocities3_chain].
 */
#include "cmc_oracle.h"
#include <math.h>
#include <float.h>
#include <string.h>
#include <stdlib.h>

static inline double phi_s(const orc_state *s, double rad, int64_t j)
{
    return (rad >= s->r[j] ? s->m[j] / rad : s->m[j] / s->r[j]) * (1.0 / s->N_STAR);
}
#define OMIN(a, b) ((a) <= (b) ? (a) : (b))

/* cmc_utils.c:2153-2172 */
void orc_energy_conservation1(orc_state *s)
{
    for (int64_t i = 1; i <= s->N_MAX_NEW; i++) {
        s->vtold[i] = s->vt[i];
        s->vrold[i] = s->vr[i];
        s->Uoldrold[i] = s->phi[i] + phi_s(s, s->r[i], i);
    }
}

/* cmc_utils.c:2217-2227 */
void orc_energy_conservation2(orc_state *s)
{
    for (int64_t i = 1; i <= s->N_MAX_NEW; i++)
        s->Uoldrnew[i] = orc_potential(s, s->rnew[i]) + phi_s(s, s->rnew[i], i);
}

/* cmc_evolution_thr.c:443-646 with TIDAL_TREATMENT = 0 (logging omitted);
 * destroy_obj (cmc_dynamics_helper.c:131-153) zeroes the star record. */
void orc_tidally_strip_stars(orc_state *s)
{
    double DTidalMassLoss;
    s->Etidal = 0.0;
    do {
        s->Rtidal = s->orbit_r * pow(s->Mtotal - (s->TidalMassLoss - s->OldTidalMassLoss), 1.0 / 3.0);
        DTidalMassLoss = 0.0;
        for (int64_t i = 1; i <= s->N_MAX_NEW; i++) {
            double m = s->m[i];
            if (s->r_apo[i] > s->Rtidal && s->rnew[i] < 1000000) {
                s->rnew[i] = ORC_SF_INFINITY;
                s->vrnew[i] = 0.0;
                s->vtnew[i] = 0.0;
                s->Eescaped += s->E[i] * m / s->N_STAR;
                s->Jescaped += s->J[i] * m / s->N_STAR;
                DTidalMassLoss += m / s->N_STAR;
                s->Etidal += s->E[i] * m / s->N_STAR;
                /* destroy_obj */
                s->vr[i] = s->vt[i] = s->E[i] = s->J[i] = 0.0;
                s->id[i] = 0;
                s->X[i] = s->Y[i] = s->r_peri[i] = s->r_apo[i] = 0.0;
                s->rOld[i] = s->Uoldrold[i] = s->Uoldrnew[i] = s->vrold[i] = s->vtold[i] = s->EI[i] = 0.0;
                s->m[i] = DBL_MIN;
                s->n_removed_tidal++;
                if (s->flags) s->flags[i] |= ORC_F_STRIPPED;
            }
        }
        s->TidalMassLoss += DTidalMassLoss;
    } while (DTidalMassLoss > 0);
}

/* cmc_utils.c:368-391 */
void orc_compute_intermediate_energy(orc_state *s)
{
    for (int64_t j = 1; j <= s->N_MAX_NEW; j++)
        if (s->rnew[j] < 1.0e6)
            s->EI[j] = orc_sqr(s->vr[j]) + orc_sqr(s->vt[j]) + s->phi[j] - orc_potential(s, s->rnew[j]);
    for (int64_t j = 1; j <= s->N_MAX_NEW; j++) {
        s->rOld[j] = s->r[j];
        s->r[j] = s->rnew[j];
        s->vr[j] = s->vrnew[j];
        s->vt[j] = s->vtnew[j];
    }
}

/* cmc_utils.c:254-354 set_velocities3 at one rank.  The hand-off loop
 * (:313-350) restarts from Eexcess_prev = 0 on rank 0, so at P = 1 it changes
 * nothing and the excess left after the first loop is dropped (Eoops += 0). */
void orc_set_velocities3(orc_state *s)
{
    const double q = 0.5, E_dump_factor = 0.8;
    double Eexcess = 0.0;
    s->n_excess_events = 0;
    s->n_excess_dumps = 0;
    for (int64_t i = 1; i <= s->N_MAX_NEW; i++) {
        double m = s->m[i];
        double Unewrold = orc_potential(s, s->rOld[i]) + phi_s(s, s->rOld[i], i);
        double Unewrnew = s->phi[i] + phi_s(s, s->r[i], i);
        double vold2 = s->vtold[i] * s->vtold[i] + s->vrold[i] * s->vrold[i];
        double vnew2 = vold2 + 2.0 * (1.0 - q) * (s->Uoldrold[i] - s->Uoldrnew[i])
                       + 2.0 * q * (Unewrold - Unewrnew);
        if (vnew2 <= 0.0) {
            s->n_excess_events++;
            Eexcess += 0.5 * (orc_sqr(s->vr[i]) + orc_sqr(s->vt[i]) - vnew2) * m;
        } else {
            double alpha = sqrt(vnew2 / (orc_sqr(s->vr[i]) + orc_sqr(s->vt[i])));
            s->vr[i] *= alpha;
            s->vt[i] *= alpha;
            double v2 = orc_sqr(s->vr[i]) + orc_sqr(s->vt[i]);
            s->vrnew[i] = s->vr[i];
            s->vtnew[i] = s->vt[i];
            double E_dump_capacity = E_dump_factor * 0.5 * m * v2;
            double E_dump = OMIN(E_dump_capacity, Eexcess);
            if (Eexcess > 0) {
                s->n_excess_dumps++;
                double exc_ratio = sqrt((v2 - 2 * E_dump / m) / v2);
                s->vr[i] *= exc_ratio;
                s->vt[i] *= exc_ratio;
                Eexcess -= E_dump;
            }
        }
    }
}

/* cmc_utils.c:604-653 ComputeEnergy (binaries and internal energies are zero) */
void orc_compute_energy(orc_state *s)
{
    double K = 0.0, P = 0.0, phi0 = s->phi[0];
    for (int64_t i = 1; i <= s->N_MAX; i++) {
        s->E[i] = s->phi[i] + 0.5 * (orc_sqr(s->vr[i]) + orc_sqr(s->vt[i]));
        s->J[i] = s->r[i] * s->vt[i];
    }
    for (int64_t i = 1; i <= s->N_MAX; i++) {
        K += 0.5 * (orc_sqr(s->vr[i]) + orc_sqr(s->vt[i])) * s->m[i] * s->madhoc;
        P += s->phi[i] * s->m[i] * s->madhoc;
        P += phi0 * s->cenma_m * s->madhoc / s->N_MAX;
    }
    P *= 0.5;
    s->Ekin = K;
    s->Epot = P;
    s->Etot = K + P + 0.0 + 0.0;
    s->Etot += 0.0 + s->Eescaped + 0.0 + 0.0;
}

/* cmc_utils.c:2132-2148 calc_timestep with GetTimeStep (cmc_evolution_thr.c:317-438):
 * only the relaxation time step is finite with the configs' switches. */
double orc_calc_timestep(orc_state *s)
{
    double Dt = orc_simul_relax_new(s);
    if (Dt == INFINITY) Dt = 0.001;
    if ((s->TidalMassLoss_old - s->OldTidalMassLoss) > 0.01) {
        Dt = s->Prev_Dt * 0.8;
    } else if (Dt > 1.1 * s->Prev_Dt && s->Prev_Dt > 0 && (s->TidalMassLoss - s->OldTidalMassLoss) > 0.02) {
        Dt = s->Prev_Dt * 1.1;
    }
    s->Dt = Dt;
    s->TotalTime += Dt;
    return Dt;
}

/* cmc.c:151-165 subset: potential, velocity dispersion, energies */
void orc_init(orc_state *s)
{
    s->madhoc = 1.0 / ((double)s->N_STAR);
    orc_potential_calculate(s);
    orc_calc_sigma_r(s);
    orc_compute_energy(s);
    s->N_MAX_NEW = s->N_MAX;
}

/* One Henon step in the order of cmc.c:240-407 for the configs' switches. */
void orc_step(orc_state *s)
{
    double Eescaped_old = s->Eescaped, Jescaped_old = s->Jescaped, Etidal_old = s->Etidal;
    s->TidalMassLoss_old = s->TidalMassLoss;
    s->Eescaped = 0.0; s->Jescaped = 0.0; s->TidalMassLoss = 0.0; s->Etidal = 0.0;

    orc_calc_timestep(s);
    s->N_MAX_NEW = s->N_MAX;
    orc_dynamics_apply(s, s->Dt);
    if (s->error) return;
    s->Prev_Dt = s->Dt;
    orc_energy_conservation1(s);
    /* new_orbits_calculate, cmc_utils.c:2194-2211 */
    s->OldTidalMassLoss = s->TidalMassLoss_old;
    orc_get_positions(s);
    if (s->error) return;
    s->TidalMassLoss += s->TidalMassLoss_old;
    orc_energy_conservation2(s);
    orc_tidally_strip_stars(s);
    orc_compute_intermediate_energy(s);
    orc_sort_by_r(s);
    orc_potential_calculate(s);
    orc_set_velocities3(s);
    /* compute_energy_new, cmc_utils.c:2017-2028 */
    s->Eescaped += Eescaped_old;
    s->Jescaped += Jescaped_old;
    s->Etidal += Etidal_old;
    orc_compute_energy(s);
    s->step++;
}

/* cmc_utils.c:872-982 comp_mass_percent with one rank (get_global_idx(k) = k, the
 * MPI_Exscan prefixes are zero): running sums in the reference's order. */
void orc_comp_mass_percent(const orc_state *s, const double *pc, int count, int bh_include,
                           double initial_total_mass, double minimum_r, orc_mass_pc_out *o)
{
    const double PI = 3.14159265358979323;             /* include/cmc/cmc.h:62 */
    double mprev;
    int mcount = 0;
    if (bh_include) {
        mprev = s->cenma_m * s->madhoc;
        for (mcount = 0; mcount < count; mcount++) {
            if (mprev / s->Mtotal > pc[mcount]) {          /* cmc_utils.c:879-890 */
                o->mass_r[mcount] = minimum_r;
                o->ave_mass_r[mcount] = 0.0;
                o->no_star_r[mcount] = 0;
                o->densities_r[mcount] = 0.0;
                o->ke_rad_r[mcount] = 0.0;
                o->ke_tan_r[mcount] = 0.0;
                o->v2_rad_r[mcount] = 0.0;
                o->v2_tan_r[mcount] = 0.0;
            } else {
                break;
            }
        }
    } else {
        mprev = 0.0;
        mcount = 0;
    }
    const int64_t nn = s->N_MAX_NEW;
    double *kr = calloc((size_t)nn + 1, sizeof(double)), *kt = calloc((size_t)nn + 1, sizeof(double));
    double *vr2 = calloc((size_t)nn + 1, sizeof(double)), *vt2 = calloc((size_t)nn + 1, sizeof(double));
    for (int64_t k = 1; k <= nn; k++) {                     /* cmc_utils.c:906-911 */
        kr[k] = kr[k - 1] + 0.5 * s->m[k] * s->madhoc * s->vr[k] * s->vr[k];
        kt[k] = kt[k - 1] + 0.5 * s->m[k] * s->madhoc * s->vt[k] * s->vt[k];
        vr2[k] = vr2[k - 1] + s->vr[k] * s->vr[k];
        vt2[k] = vt2[k - 1] + s->vt[k] * s->vt[k];
    }
    if (mcount < count) {
        for (int64_t k = 1; k <= s->N_MAX; k++) {          /* cmc_utils.c:931-971 */
            mprev += s->m[k] / s->N_STAR;
            if (mprev > pc[mcount] * s->Mtotal) {
                o->mass_r[mcount] = s->r[k];
                o->ave_mass_r[mcount] = mprev / s->Mtotal / k * initial_total_mass;
                o->no_star_r[mcount] = k;
                o->densities_r[mcount] = mprev / (4.0 / 3.0 * PI * pow(s->r[k], 3));
                o->ke_rad_r[mcount] = kr[k];
                o->ke_tan_r[mcount] = kt[k];
                o->v2_rad_r[mcount] = vr2[k];
                o->v2_tan_r[mcount] = vt2[k];
                mcount++;
                if (mcount == count) break;
            }
        }
    }
    free(kr); free(kt); free(vr2); free(vt2);
}

/* central_calculate, cmc_utils.c:1239-1345 (first part; J = 6), and
 * clusdyn_calculate, cmc_utils.c:1614-1624.  sqr/cub are the reference's
 * x*x and x*x*x (cmc_utils.c:27-42), fb_cub the same (fewbody_utils.c:72-75). */
int orc_central_calculate(const orc_state *s, orc_central_out *o)
{
    const double PI = 3.14159265358979323;             /* include/cmc/cmc.h:62 */
    const long J = 6;
    double m = 0.0;
    int64_t nave = 1;
    while (m < 0.5 * s->Mtotal) {                        /* cmc_utils.c:1247-1251 */
        if (nave > s->N_MAX) return -1;
        m += s->m[nave] / s->N_STAR;
        nave++;
    }
    if (s->N_STAR <= 2 * J || nave >= s->N_STAR - 6) return -1;   /* :1254-1257 */
    if ((nave - J / 2 > 1 ? nave - J / 2 : 1) + J > s->N_MAX + 1) return -1;   /* last jmax past the live stars */
    double *rhoj = malloc((size_t)(nave + 1) * sizeof(double));
    for (int64_t i = 1; i <= nave; i++) {                /* :1263-1274 */
        int64_t jmin = i - J / 2 > 1 ? i - J / 2 : 1;
        int64_t jmax = jmin + J;
        double mrho = 0.0;
        for (int64_t j = jmin + 1; j <= jmax - 1; j++) mrho += s->m[j] * s->madhoc;
        double a = s->r[jmax], b = s->r[jmin];
        double Vrj = 4.0 / 3.0 * PI * (a * a * a - b * b * b);
        rhoj[i] = mrho / Vrj;
    }
    double rhojsum = 0.0, rhoj2sum = 0.0, rho = 0.0, v_rms = 0.0, rc = 0.0, rc_nb = 0.0, m_ave = 0.0;
    for (int64_t i = 1; i <= nave; i++) {                /* :1283-1294 */
        rhojsum += rhoj[i];
        rhoj2sum += rhoj[i] * rhoj[i];
        rho += rhoj[i] * rhoj[i];
        v_rms += rhoj[i] * (s->vr[i] * s->vr[i] + s->vt[i] * s->vt[i]);
        rc += rhoj[i] * s->r[i];
        double q = rhoj[i] * s->r[i];
        rc_nb += q * q;
        m_ave += rhoj[i] * s->m[i] * s->madhoc;
    }
    free(rhoj);
    rho /= rhojsum;                                      /* :1323-1345 */
    rho *= 4.0 / 5.0;
    rho *= 2.0;
    v_rms /= rhojsum;
    v_rms = sqrt(v_rms);
    rc /= rhojsum;
    rc_nb = sqrt(rc_nb / rhoj2sum);
    m_ave /= rhojsum;
    double n = rho / m_ave;
    o->rc_spitzer = sqrt(3.0 * (v_rms * v_rms) / (4.0 * PI * rho));
    o->N_core = 4.0 / 3.0 * PI * (rc * rc * rc) * (n / 2.0);
    o->N_core_nb = 4.0 / 3.0 * PI * (rc_nb * rc_nb * rc_nb) * (n / 2.0);
    o->Trc = 0.065 * (v_rms * v_rms * v_rms) / (rho * m_ave);
    o->nave = nave;
    o->rh = s->r[nave];                                  /* clusdyn_calculate: same loop, rh = star_r[k] */
    o->rhojsum = rhojsum;
    o->rhoj2sum = rhoj2sum;
    o->rho = rho;
    o->v_rms = v_rms;
    o->rc = rc;
    o->rc_nb = rc_nb;
    o->m_ave = m_ave;
    o->n = n;
    return 0;
}
