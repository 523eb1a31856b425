/*
 * The reference's hot-path call surface on top of the cmcg C-ABI (one device,
 * one process-wide context).  Each function keeps the reference name, arguments
 * and effect on the cluster state; the state itself lives on the device and the
 * scalar globals are mirrored in cmc_hp after every call, and written to the
 * reference's own globals when those are linked in (below).
 */
#include "../../include/cmc_hotpath.h"
#include <stdio.h>
#include <stdlib.h>
#include <float.h>

cmc_hotpath_globals cmc_hp;
cmc_hotpath_lagrange cmc_lag = {0, 0, 0.0, 1.0, {0}, {0}, {0}, {0}, {0}, {0}, {0}, {0}, {0}};
cmc_hotpath_central cmc_cen;

/* The reference's own globals that the path updates (include/cmc/cmc_vars.h:11-12,
 * 32-33,38,188,293; clus_struct_t and Etotal_struct_t, include/cmc/cmc.h:1690-1745).
 * Weak references: linked into the reference they bind to its definitions, and every
 * call leaves them as the reference's own functions would, so its main loop reads
 * them unchanged; linked alone they are null and only cmc_hp is kept. */
typedef struct { long N_MAX, N_MAX_NEW, N_STAR, N_STAR_NEW, N_BINARY; } cmc_ref_clus;
typedef struct { double tot, New, ini, K, P, Eint, Eb; } cmc_ref_etotal;
extern cmc_ref_clus clus __attribute__((weak));
extern cmc_ref_etotal Etotal __attribute__((weak));
extern double Mtotal __attribute__((weak)), Rtidal __attribute__((weak)), max_r __attribute__((weak)),
    Dt __attribute__((weak)), TotalTime __attribute__((weak)), TidalMassLoss __attribute__((weak)),
    Eescaped __attribute__((weak)), Jescaped __attribute__((weak)), Etidal __attribute__((weak)),
    Eoops __attribute__((weak));

static void set_ref(double *g, double v)
{
    if (g) *g = v;
}

static cmcg_ctx *g_ctx = NULL;
static uint64_t g_step = 0;   /* RNG step counter (tcount - 1 in the reference's main loop) */

static void default_handler(int code, int64_t index, const char *where)
{
    fprintf(stderr, "cmc hot path: %s in %s (star %lld)\n", cmcg_error_string(code), where, (long long)index);
    exit(code < 0 ? -code : code);   /* exit_cleanly, src/cmc/cmc_utils.c:208-220 */
}
static cmc_hotpath_error_handler g_handler = default_handler;

void cmc_hotpath_set_error_handler(cmc_hotpath_error_handler h) { g_handler = h ? h : default_handler; }

static void refresh(void)
{
    cmcg_scalars s;
    cmcg_get_scalars(g_ctx, &s);
    cmc_hp.N_MAX = (long)s.n_max;
    cmc_hp.N_MAX_NEW = (long)s.n_max_new;
    cmc_hp.Mtotal = s.Mtotal;
    cmc_hp.Rtidal = s.Rtidal;
    cmc_hp.max_r = s.max_r;
    cmc_hp.Dt = s.Dt;
    cmc_hp.TotalTime = s.TotalTime;
    cmc_hp.TidalMassLoss = s.TidalMassLoss;
    cmc_hp.Eescaped = s.Eescaped;
    cmc_hp.Jescaped = s.Jescaped;
    cmc_hp.Etidal = s.Etidal;
    cmc_hp.Etot = s.Etot;
    cmc_hp.Ekin = s.Ekin;
    cmc_hp.Epot = s.Epot;
    cmc_hp.Eoops = s.Eoops;
    if (&clus) {
        clus.N_MAX = cmc_hp.N_MAX;
        clus.N_MAX_NEW = cmc_hp.N_MAX_NEW;
    }
    if (&Etotal) {
        Etotal.tot = s.Etot;
        Etotal.K = s.Ekin;
        Etotal.P = s.Epot;
    }
    set_ref(&Mtotal, s.Mtotal);
    set_ref(&Rtidal, s.Rtidal);
    set_ref(&max_r, s.max_r);
    set_ref(&Dt, s.Dt);
    set_ref(&TotalTime, s.TotalTime);
    set_ref(&TidalMassLoss, s.TidalMassLoss);
    set_ref(&Eescaped, s.Eescaped);
    set_ref(&Jescaped, s.Jescaped);
    set_ref(&Etidal, s.Etidal);
    set_ref(&Eoops, s.Eoops);
}

static int check(int rc, const char *where)
{
    if (rc < 0) {
        cmcg_scalars s;
        (void)s;
        g_handler(rc, -1, where);
    }
    return rc;
}

int cmc_hotpath_init(const cmcg_params *p, int64_t n_cap, const cmcg_host_soa *stars, int device)
{
    if (g_ctx) cmc_hotpath_finalize();
    int rc = cmcg_create(&g_ctx, device, n_cap, p);
    if (rc) return rc;
    rc = cmcg_upload(g_ctx, stars);
    if (rc) return rc;
    cmc_hp.N_STAR = (long)p->n_star;
    cmc_hp.madhoc = 1.0 / ((double)p->n_star);
    cmc_hp.tcount = 1;
    g_step = 0;
    refresh();
    return CMCG_OK;
}

void cmc_hotpath_finalize(void)
{
    if (g_ctx) cmcg_destroy(g_ctx);
    g_ctx = NULL;
}

cmcg_ctx *cmc_hotpath_ctx(void) { return g_ctx; }

int cmc_hotpath_download(cmcg_host_soa *h) { return cmcg_download(g_ctx, h); }

/* save_restart_file (src/cmc/cmc_io.c:3829-3884) for the hot path's state: one file
 * holding the device checkpoint (fields, scalars, step counter, parameters). */
int cmc_hotpath_save_restart(const char *path)
{
    int64_t bytes = 0;
    int rc = cmcg_checkpoint_size(g_ctx, &bytes);
    if (rc) return rc;
    void *buf = malloc((size_t)bytes);
    if (!buf) return CMCG_E_NOMEM;
    rc = cmcg_checkpoint_save(g_ctx, buf, bytes);
    if (!rc) {
        FILE *f = fopen(path, "wb");
        if (!f) rc = CMCG_E_ARG;
        else {
            if (fwrite(buf, 1, (size_t)bytes, f) != (size_t)bytes) rc = CMCG_E_ARG;
            if (fclose(f) != 0) rc = CMCG_E_ARG;
        }
    }
    free(buf);
    return rc;
}

/* load_restart_file (src/cmc/cmc_io.c:3886-4031): creates the context from the
 * parameters stored in the file (capacity max(n_cap, N)) and restores the state;
 * the next step continues exactly where the saved run stopped. */
int cmc_hotpath_load_restart(const char *path, int64_t n_cap, int device)
{
    FILE *f = fopen(path, "rb");
    if (!f) return CMCG_E_ARG;
    int rc = CMCG_OK;
    long size = -1;
    if (fseek(f, 0, SEEK_END) == 0) size = ftell(f);
    void *buf = size > 0 ? malloc((size_t)size) : NULL;
    if (!buf || fseek(f, 0, SEEK_SET) != 0 || fread(buf, 1, (size_t)size, f) != (size_t)size) rc = buf ? CMCG_E_ARG : CMCG_E_NOMEM;
    fclose(f);
    cmcg_params p;
    int64_t n = 0;
    if (!rc) rc = cmcg_checkpoint_info(buf, size, &p, &n);
    if (!rc) {
        if (g_ctx) cmc_hotpath_finalize();
        rc = cmcg_create(&g_ctx, device, n_cap > n ? n_cap : n, &p);
    }
    if (!rc) rc = cmcg_checkpoint_load(g_ctx, buf, size);
    free(buf);
    if (rc) return rc;
    cmcg_scalars s;
    cmcg_get_scalars(g_ctx, &s);
    cmc_hp.N_STAR = (long)p.n_star;
    cmc_hp.madhoc = 1.0 / ((double)p.n_star);
    g_step = s.step;
    cmc_hp.tcount = (long)s.step + 1;
    refresh();
    return CMCG_OK;
}

void qsorts_new(void)
{
    int64_t n;
    check(cmcg_sort_by_r(g_ctx, &n), "qsorts_new");
    refresh();
}

/* On one device the sorted r and m already are the global arrays. */
void post_sort_comm(void) {}

long potential_calculate(void)
{
    int64_t n;
    check(cmcg_potential(g_ctx, &n, NULL, NULL), "potential_calculate");
    refresh();
    return (long)n;
}

void calc_potential_new(void) { (void)potential_calculate(); }

double potential(double r)
{
    double out = 0.0;
    check(cmcg_potential_at(g_ctx, 1, &r, &out), "potential");
    return out;
}

void calc_sigma_r_all(void) { check(cmcg_sigma_r(g_ctx), "calc_sigma_r"); }

double simul_relax_new(void)
{
    double dt = 0.0;
    check(cmcg_simul_relax(g_ctx, &dt), "simul_relax_new");
    return dt;
}

void calc_timestep(void *rng)
{
    (void)rng;
    double dt;
    check(cmcg_calc_timestep(g_ctx, &dt), "calc_timestep");
    refresh();
}

void dynamics_apply(double dt, void *rng)
{
    (void)rng;   /* the reference's gsl_rng argument; the path draws from curr_st */
    check(cmcg_relax(g_ctx, dt, g_step), "dynamics_apply");
    refresh();
}

void energy_conservation1(void) { check(cmcg_energy_conservation1(g_ctx), "energy_conservation1"); }

double get_positions(void)
{
    cmcg_orbit_stats st;
    check(cmcg_orbits(g_ctx, g_step, &st), "get_positions");
    refresh();
    return st.max_r;
}

void new_orbits_calculate(void) { (void)get_positions(); }

void energy_conservation2(void) { check(cmcg_energy_conservation2(g_ctx), "energy_conservation2"); }

void tidally_strip_stars(void)
{
    check(cmcg_tidally_strip_stars(g_ctx), "tidally_strip_stars");
    refresh();
}

void ComputeIntermediateEnergy(void)
{
    check(cmcg_compute_intermediate_energy(g_ctx), "ComputeIntermediateEnergy");
}

void energy_conservation3(void) { check(cmcg_set_velocities3(g_ctx), "energy_conservation3"); }

void compute_energy_new(void)
{
    check(cmcg_compute_energy(g_ctx), "compute_energy_new");
    refresh();
    g_step++;
    cmc_hp.tcount++;
}

void comp_mass_percent(void)
{
    int64_t k[256];
    cmcg_mass_pc o = {cmc_lag.mass_r, cmc_lag.ave_mass_r, cmc_lag.densities_r, cmc_lag.ke_rad_r, cmc_lag.ke_tan_r,
                      cmc_lag.v2_rad_r, cmc_lag.v2_tan_r, k, 0, 0};
    int n = cmc_lag.MASS_PC_COUNT < 256 ? cmc_lag.MASS_PC_COUNT : 256;
    check(cmcg_comp_mass_percent(g_ctx, cmc_lag.mass_pc, n, cmc_lag.MASS_PC_BH_INCLUDE, cmc_lag.initial_total_mass,
                                 cmc_lag.MINIMUM_R, &o), "comp_mass_percent");
    for (int i = 0; i < o.reached; i++) cmc_lag.no_star_r[i] = (long)k[i];
}

void central_calculate(void)
{
    cmcg_central c;
    check(cmcg_central_calculate(g_ctx, &c), "central_calculate");
    cmc_cen.rho = cmc_cen.rho_core = c.rho;
    cmc_cen.v_rms = cmc_cen.v_core = c.v_rms;
    cmc_cen.rc = cmc_cen.core_radius = c.rc;
    cmc_cen.m_ave = c.m_ave;
    cmc_cen.n = c.n;
    cmc_cen.rc_spitzer = c.rc_spitzer;
    cmc_cen.N_core = c.N_core;
    cmc_cen.N_core_nb = c.N_core_nb;
    cmc_cen.rc_nb = c.rc_nb;
    cmc_cen.Trc = c.Trc;
    cmc_cen.rh = c.rh;
    cmc_cen.nave = (long)c.nave;
}

/* clusdyn.rh comes from the same half-mass loop; one device call serves both. */
void clusdyn_calculate(void) { central_calculate(); }

const char *cmc_hotpath_dyn_header(void)
{
    return "# Dynamical information [code units]\n"
           "#1:t #2:Dt #3:tcount #4:N #5:M #6:VR #7:N_c #8:r_c #9:r_max #10:Etot #11:KE #12:PE #13:Etot_int "
           "#14:Etot_bin #15:E_cenma #16:Eesc #17:Ebesc #18:Eintesc #19:Eoops #20:Etot+Eoops #21:r_h #22:rho_0 "
           "#23:rc_spitzer #24:v0_rms #25:rc_nb #26.DMse(MSUN) #27.DMrejuv(MSUN) #28.N_c_nb\n";
}

int cmc_hotpath_dyn_line(char *buf, size_t len)
{
    const double Eint = 0.0, Eb = 0.0, cenma_E = 0.0, Ebesc = 0.0, Eintesc = 0.0, DMse = 0.0, DMrejuv = 0.0;
    return snprintf(buf, len,
                    "%.8g %.8g %ld %ld %.8g %.8g %.8g %.8g %.8g %.8g %.8g %.8g %.8g %.8g %.8g %.8g %.8g %.8g "
                    "%.8g %.8g %.8g %.8g %.8g %.8g %.8g %.8g %.8g %.8g\n",
                    cmc_hp.TotalTime, cmc_hp.Dt, cmc_hp.tcount, cmc_hp.N_MAX, cmc_hp.Mtotal,
                    -2.0 * cmc_hp.Ekin / cmc_hp.Epot, cmc_cen.N_core, cmc_cen.core_radius, cmc_hp.max_r,
                    cmc_hp.Etot, cmc_hp.Ekin, cmc_hp.Epot, Eint, Eb, cenma_E, cmc_hp.Eescaped, Ebesc, Eintesc,
                    cmc_hp.Eoops, cmc_hp.Etot + cmc_hp.Eoops, cmc_cen.rh, cmc_cen.rho, cmc_cen.rc_spitzer,
                    cmc_cen.v_rms, cmc_cen.rc_nb, DMse, DMrejuv, cmc_cen.N_core_nb);
}

cmc_orbit_rs_t calc_orbit_new(long index, double E, double J)
{
    cmc_orbit_rs_t o;
    double rp, ra, dp, da;
    int64_t kmin, kmax;
    int32_t circ;
    check(cmcg_calc_pericenter_and_apocenter(g_ctx, index, 1, &E, &J, &rp, &ra, &dp, &da, &kmin, &kmax, &circ),
          "calc_orbit_new");
    o.rp = rp; o.ra = ra; o.dQdrp = dp; o.dQdra = da;
    o.kmin = (long)kmin; o.kmax = (long)kmax; o.circular_flag = circ;
    return o;
}

cmc_orbit_rs_t calc_pericenter_and_apocenter(long index, double E, double J) { return calc_orbit_new(index, E, J); }

long find_zero_Q(long j, long kmin, long kmax, long double E, long double J)
{
    int64_t jj = j, a = kmin, b = kmax, out = -1;
    double e = (double)E, jv = (double)J;
    check(cmcg_find_zero_Q(g_ctx, 1, &jj, &a, &b, &e, &jv, &out), "find_zero_Q");
    return (long)out;
}

double find_root_vr(long index, long k, double E, double J)
{
    int64_t jj = index, kk = k;
    double out = 0.0;
    check(cmcg_find_root_vr(g_ctx, 1, &jj, &kk, &E, &J, &out), "find_root_vr");
    return out;
}

int cmc_hotpath_step(cmcg_orbit_stats *stats)
{
    int rc = cmcg_step(g_ctx, stats);
    check(rc, "henon step");
    refresh();
    g_step++;
    cmc_hp.tcount++;
    return rc;
}
