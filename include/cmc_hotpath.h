/*
 * C host side of the MI355X hot path, keeping the reference's call surface.
 *
 * The reference's hot-path functions are `void` functions over globals
 * (include/cmc/cmc_vars.h: star[], star_r/star_m/star_phi, clus, Dt, Mtotal,
 * Rtidal, curr_st ...).  This header gives the same names and argument meanings;
 * their bodies call the cmcg_* C-ABI (include/cmcg.h) on one process-wide device
 * context created by cmc_hotpath_init.  A maintainer links this library in place
 * of the reference objects that define these names.
 *
 * Failure semantics follow exit_cleanly (src/cmc/cmc_utils.c:208-220): on a
 * negative cmcg code the installed handler is called; the default handler prints
 * the error and the offending star index and exits with the code.
 */
#ifndef CMC_HOTPATH_H
#define CMC_HOTPATH_H
#include <stdint.h>
#include "cmcg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* orbit_rs_t, include/cmc/cmc.h:1621-1650 */
typedef struct {
    double rp, ra, dQdrp, dQdra;
    long kmax, kmin;
    int circular_flag;
} cmc_orbit_rs_t;

/* Mirror of the reference globals the path reads or writes (refreshed after every call).
 * When the reference's own globals are linked in (clus, Etotal, Dt, Mtotal, Rtidal ...,
 * include/cmc/cmc_vars.h), the library writes those too, through weak references. */
typedef struct {
    long N_STAR, N_MAX, N_MAX_NEW;      /* clus.* (include/cmc/cmc.h:1690-1711) */
    double Mtotal, Rtidal, max_r, Dt, TotalTime, madhoc;
    double TidalMassLoss, Eescaped, Jescaped, Etidal;
    double Etot, Ekin, Epot;
    long tcount;
    double Eoops;
} cmc_hotpath_globals;

extern cmc_hotpath_globals cmc_hp;

/* Inputs and outputs of comp_mass_percent (the reference's globals mass_pc,
 * MASS_PC_COUNT, MASS_PC_BH_INCLUDE, MINIMUM_R, initial_total_mass and mass_r,
 * ave_mass_r, no_star_r, densities_r, ke_rad_r, ke_tan_r, v2_rad_r, v2_tan_r,
 * include/cmc/cmc_vars.h).  At most 256 fractions, ascending. */
typedef struct {
    int MASS_PC_COUNT, MASS_PC_BH_INCLUDE;
    double MINIMUM_R, initial_total_mass;
    double mass_pc[256];
    double mass_r[256], ave_mass_r[256], densities_r[256];
    double ke_rad_r[256], ke_tan_r[256], v2_rad_r[256], v2_tan_r[256];
    long no_star_r[256];
} cmc_hotpath_lagrange;
extern cmc_hotpath_lagrange cmc_lag;

/* Outputs of central_calculate and clusdyn_calculate (the reference's globals
 * central.rho/.v_rms/.rc/.m_ave/.n/.rc_spitzer, rho_core, v_core, core_radius,
 * N_core, N_core_nb, rc_nb, Trc and clusdyn.rh; include/cmc/cmc_vars.h). */
typedef struct {
    double rho, v_rms, rc, m_ave, n, rc_spitzer;     /* central.* */
    double rho_core, v_core, core_radius, N_core, N_core_nb, rc_nb, Trc;
    double rh;                                       /* clusdyn.rh */
    long nave;
} cmc_hotpath_central;
extern cmc_hotpath_central cmc_cen;

typedef void (*cmc_hotpath_error_handler)(int code, int64_t index, const char *where);
void cmc_hotpath_set_error_handler(cmc_hotpath_error_handler h);

/* Creates the device context and uploads the cluster (replaces get_star_data's
 * array set-up, src/cmc/cmc_io.c:3384-3440, and set_rng_states). */
int  cmc_hotpath_init(const cmcg_params *p, int64_t n_cap, const cmcg_host_soa *stars, int device);
void cmc_hotpath_finalize(void);
cmcg_ctx *cmc_hotpath_ctx(void);
/* Copies the device state back into host arrays (for output / checkpoints). */
int  cmc_hotpath_download(cmcg_host_soa *h);

/* Restart files of the hot path's state (the counterpart of save_restart_file and
 * load_restart_file, src/cmc/cmc_io.c:3829-3884, 3886-4031): per-star fields, device
 * scalars, step counter and parameters.  _load creates the context from the file's
 * parameters with capacity max(n_cap, N); the run then continues bit for bit. */
int  cmc_hotpath_save_restart(const char *path);
int  cmc_hotpath_load_restart(const char *path, int64_t n_cap, int device);

/* --- the reference's names ------------------------------------------------ */
void   qsorts_new(void);                                   /* src/cmc/cmc_utils.c:2232 */
void   post_sort_comm(void);                               /* src/cmc/cmc_utils.c:2259 */
void   calc_potential_new(void);                           /* src/cmc/cmc_utils.c:2006 */
long   potential_calculate(void);                          /* src/cmc/cmc_utils.c:675  */
double potential(double r);                                /* src/cmc/cmc_utils.c:112  */
void   calc_sigma_r_all(void);                             /* src/cmc/cmc_dynamics_helper.c:2821 */
double simul_relax_new(void);                              /* src/cmc/cmc_dynamics_helper.c:2526 */
void   calc_timestep(void *rng);                           /* src/cmc/cmc_utils.c:2132 */
void   dynamics_apply(double dt, void *rng);               /* src/cmc/cmc_dynamics.c:17 */
void   energy_conservation1(void);                         /* src/cmc/cmc_utils.c:2153 */
void   new_orbits_calculate(void);                         /* src/cmc/cmc_utils.c:2194 */
double get_positions(void);                                /* src/cmc/cmc_evolution_thr.c:944 */
void   energy_conservation2(void);                         /* src/cmc/cmc_utils.c:2217 */
void   tidally_strip_stars(void);                          /* src/cmc/cmc_evolution_thr.c:443 */
void   ComputeIntermediateEnergy(void);                    /* src/cmc/cmc_utils.c:368 */
void   energy_conservation3(void);                         /* src/cmc/cmc_utils.c:360 */
void   compute_energy_new(void);                           /* src/cmc/cmc_utils.c:2017 */
void   comp_mass_percent(void);                            /* src/cmc/cmc_utils.c:872  */
void   central_calculate(void);                            /* src/cmc/cmc_utils.c:1239 */
void   clusdyn_calculate(void);                            /* src/cmc/cmc_utils.c:1614 */
/* One row of the reference's <output>.dyn.dat (the fprintf of src/cmc/cmc_io.c:320-323,
 * same columns and %.8g format, newline included) for the current state, from cmc_hp and
 * cmc_cen: call central_calculate() and compute_energy_new() first.  Columns of physics
 * outside the hot path (internal and binary energies, cenma.E, binary escapes, DMse,
 * DMrejuv) are 0, as they are in the reference with those modules off.  Returns what
 * snprintf returns. */
int    cmc_hotpath_dyn_line(char *buf, size_t len);
/* The two header lines of that file (src/cmc/cmc_io.c:2346-2347). */
const char *cmc_hotpath_dyn_header(void);
cmc_orbit_rs_t calc_orbit_new(long index, double E, double J);                 /* src/cmc/cmc_orbit.c:413 */
cmc_orbit_rs_t calc_pericenter_and_apocenter(long index, double E, double J);  /* calc_orbit_new under its descriptive name */
long   find_zero_Q(long j, long kmin, long kmax, long double E, long double J); /* src/cmc/cmc_nr.c:316 */
double find_root_vr(long index, long k, double E, double J);                  /* src/cmc/cmc_nr.c:500 */
/* One Henon step in the order of src/cmc/cmc.c:240-407 (configs' switch set). */
int    cmc_hotpath_step(cmcg_orbit_stats *stats);

/* --- initial conditions (seeded generators; host side) -------------------- */
/* Outputs are sorted by r with ids 1..n, stored masses m*N_STAR, sentinels at 0 and
 * n+1, Henon units scaled exactly to PE = -1/2, KE = 1/4.  Arrays need n+2 entries;
 * r, m, vr, vt and id must be non-NULL.  *rtid receives the tidal radius to use as
 * cmcg_params.orbit_r. */
int cmc_ic_plummer(int64_t n, uint64_t seed, int kroupa, cmcg_host_soa *out, double *rtid);
int cmc_ic_king(int64_t n, double w0, uint64_t seed, int kroupa, int segregate, cmcg_host_soa *out, double *rtid);
/* Kroupa (2001) two-segment IMF, alpha = 1.3 on [0.08, 0.5] and 2.3 on [0.5, 150]
 * solar masses; n draws into m[0..n-1] (not normalised). */
int cmc_ic_kroupa(int64_t n, uint64_t seed, double *m);

#ifdef __cplusplus
}
#endif
#endif
