/*
 * CPU oracle for the CMC Henon-step hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C over SoA arrays, the reference algorithms of
 * ClusterMonteCarlo/CMC-COSMIC for the per-timestep star loop.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker (or the timed CPU baseline).  The product path never links it.
 *
 * Every function cites the reference file:line it follows.  Index conventions are
 * the reference's: global arrays r/m/phi hold N+2 entries, index 0 is the central
 * sentinel (r[0] = ZERO = 1e-20) and index N+1 the outer one (r = SF_INFINITY).
 * At one MPI rank the reference's local index i and global index g_i coincide
 * (cmc_utils.c:2384-2405 with myid == 0), so one set of arrays serves both.
 *
 * Pinning status of each function is stated next to it ("pinned by ..." or
 * "parity unpinned").
 */
#ifndef CMC_ORACLE_H
#define CMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* constants: include/cmc/cmc.h:58,60,62,64 */
#define ORC_SF_INFINITY 1.0e10
#define ORC_ZERO 1.0e-20
#define ORC_PI 3.14159265358979323
#define ORC_N_TRY 50000

/* ---------------------------------------------------------------- RNG */
/* taus113 state: src/libs/taus113-v2.c (struct rng_t113_state holds 4 words). */
typedef struct { uint32_t z[4]; } orc_t113;

void     orc_t113_seed(uint64_t s, orc_t113 *st);          /* taus113-v2.c:197-216 */
uint32_t orc_t113_int(orc_t113 *st);                       /* taus113-v2.c:166-177 */
double   orc_t113_dbl(orc_t113 *st);                       /* taus113-v2.c:187-189 */
orc_t113 orc_t113_jump_poly(orc_t113 st, const uint32_t jpoly[4]); /* taus113-v2.c:253-270 */
/* Jump by an arbitrary 128-bit step count (lo + 2^64 hi) with GF(2) matrix powers. */
orc_t113 orc_t113_jump(orc_t113 st, uint64_t lo, uint64_t hi);

/* Stream layout shared with the device path ("per-star counter-based" streams).
 *   relaxation, step s, pair p (0-based, stars 2p+1,2p+2): offset s*2^72 + 3p
 *   orbits,     step s, star slot i (1-based):             offset s*2^72 + 2^70 + i*2^20
 * Offsets are counted from the seeded state (after the 2^80*rank jump of
 * set_rng_states, cmc_utils.c:2364-2374; rank 0 here).  Step 0's relaxation draws
 * therefore coincide with the reference's single sequential stream. */
orc_t113 orc_t113_jump_pow2(orc_t113 st, int b);
orc_t113 orc_stream_relax(const orc_t113 *seeded, uint64_t step, uint64_t pair);
orc_t113 orc_stream_orbit(const orc_t113 *seeded, uint64_t step, uint64_t slot);

enum { ORC_RNG_REF_SEQ = 0, ORC_RNG_PER_STAR = 1 };
enum { ORC_Q_LONG_DOUBLE = 0, ORC_Q_FP64 = 1 };
enum { ORC_LIBM_SYSTEM = 0, ORC_LIBM_LONG_DOUBLE = 1, ORC_LIBM_DD = 2, ORC_LIBM_SINCOS = 3 };

/* ---------------------------------------------------------------- state */
typedef struct orc_state {
    /* sizes */
    int64_t N_STAR;      /* clus.N_STAR: initial star count; madhoc = 1/N_STAR */
    int64_t N_MAX;       /* clus.N_MAX: live stars */
    int64_t N_MAX_NEW;   /* clus.N_MAX_NEW */
    int64_t cap;         /* allocated entries per array (>= N_MAX + 2) */
    /* global (replicated in the reference) arrays, index 0..N_MAX+1 */
    double *r, *m, *phi;
    /* per-star fields, index 1..N_MAX_NEW */
    double *vr, *vt, *E, *J;
    int64_t *id;
    double *rnew, *vrnew, *vtnew, *X, *Y, *r_peri, *r_apo;
    double *rOld, *Uoldrold, *Uoldrnew, *vrold, *vtold, *EI;
    double *pot_rnew;    /* potential(rnew) as seen by the orbit phase (diagnostic) */
    double *sigma;       /* sigma_array.sigma, index 1..N; sigma_array.r aliases r */
    int32_t *ntry;       /* rejection tries per star (0 for circular / removed) */
    int32_t *flags;      /* per-star outcome bits, ORC_F_* */
    /* scalars */
    double madhoc, cenma_m, Mtotal, Rtidal, orbit_r, max_r;
    double Dt, Prev_Dt, TotalTime;
    double TidalMassLoss, TidalMassLoss_old, OldTidalMassLoss;
    double Eescaped, Jescaped, Etidal, Eoops;
    double Etot, Ekin, Epot;
    double theta_se_max, gamma;  /* THETASEMAX, GAMMA (Coulomb log factor) */
    int64_t avekernel;           /* AVEKERNEL */
    /* RNG */
    orc_t113 seeded;             /* state after reset_rng_t113_new(IDUM) */
    orc_t113 seq;                /* REF_SEQ cursor (curr_st) */
    int32_t rng_mode;            /* ORC_RNG_* */
    int32_t q_mode;              /* ORC_Q_* */
    uint64_t step;               /* 0-based Henon step counter for the stream layout */
    /* statistics */
    int64_t n_brent;             /* find_root_vr calls */
    int64_t n_circular, n_removed_E, n_removed_tidal, n_removed_mass;
    int64_t n_linear_scan;
    int32_t error;               /* first error code (ORC_E_*), 0 if none */
    int64_t error_index;
    int64_t n_excess_events;     /* set_velocities3 stars with vnew2 <= 0 (last call) */
    int64_t n_excess_dumps;      /* stars that absorbed excess energy (last call) */
    int32_t libm_mode;           /* ORC_LIBM_*: sin/cos of the relaxation */
    /* one rank's share (orc_step_ranks): the relaxation's pairs and the orbit phase's
     * stars are restricted to slots range_lo..range_hi when range_hi > 0, and the
     * relaxation skips calc_sigma_r when skip_sigma is set */
    int64_t range_lo, range_hi;
    int32_t skip_sigma;
} orc_state;

enum {
    ORC_F_CIRCULAR = 1, ORC_F_REMOVED_E = 2, ORC_F_REMOVED_RT = 4, ORC_F_DESTROYED = 8,
    ORC_F_BRENT = 16, ORC_F_LINSCAN = 32, ORC_F_STRIPPED = 64
};
enum { ORC_E_NAN = -1, ORC_E_NTRY = -2, ORC_E_BRACKET = -3, ORC_E_WZERO = -4, ORC_E_WPZERO = -5,
       ORC_E_BRENT = -6 };

/* comp_mass_percent (cmc_utils.c:872-982) for one rank: Lagrange radii and the
 * kinetic-energy and v^2 sums inside them, for the mass fractions pc[0..count-1]
 * (ascending).  Outputs per fraction; entries the loop does not reach keep the
 * values the caller put there. */
typedef struct {
    double *mass_r, *ave_mass_r, *densities_r, *ke_rad_r, *ke_tan_r, *v2_rad_r, *v2_tan_r;
    int64_t *no_star_r;
} orc_mass_pc_out;
void orc_comp_mass_percent(const orc_state *s, const double *pc, int count, int bh_include,
                           double initial_total_mass, double minimum_r, orc_mass_pc_out *o);

/* central_calculate (cmc_utils.c:1239-1345, the Casertano & Hut density-weighted
 * core quantities) and clusdyn_calculate (cmc_utils.c:1614-1624, the half-mass
 * radius) for one rank; the NUM_CENTRAL_STARS block after them (collision and
 * binary rates) is not restated.  Returns 0, or -1 where the reference exits
 * (N_STAR <= 12 or nave >= N_STAR - 6) or the half-mass loop would leave the live
 * stars. */
typedef struct {
    int64_t nave;          /* stars averaged over: the one just beyond the half-mass radius */
    double rh;             /* clusdyn.rh = star_r[nave] */
    double rhojsum, rhoj2sum;
    double rho, v_rms, rc, rc_nb, m_ave, n, rc_spitzer, N_core, N_core_nb, Trc;
} orc_central_out;
int orc_central_calculate(const orc_state *s, orc_central_out *o);

/* double-double sin/cos shared (operation for operation) with the device relaxation */
void orc_dd_sincos(double x, double *s, double *c);
void orc_dd_trig_table(int j, double *out4);   /* sin/cos of j pi/128 (hi, lo), j = 0..32 */

/* ---------------------------------------------------------------- kernels of the path */
double  orc_sqr(double x);                                     /* cmc_utils.c:27-30 */
double  orc_cub(double x);                                     /* cmc_utils.c:39-42 */
int64_t orc_FindZero_r(const orc_state *s, int64_t kmin, int64_t kmax, double r); /* cmc_nr.c:62-82 */
double  orc_potential(orc_state *s, double r);                 /* cmc_utils.c:112-165 */
int64_t orc_potential_calculate(orc_state *s);                 /* cmc_utils.c:675-736 */
double  orc_function_q(const orc_state *s, int64_t j, double r, double pot, double E, double J); /* cmc_utils.c:1654-1662 */
int64_t orc_find_zero_Q(const orc_state *s, int64_t j, int64_t kmin, int64_t kmax, double E, double J); /* cmc_nr.c:316-383 */
int64_t orc_get_positive_Q_index(orc_state *s, int64_t index, double E, double J); /* cmc_orbit.c:30-59 */
void    orc_set_a_b(const orc_state *s, int64_t index, int64_t k, double *a, double *b); /* cmc_orbit.c:722-740 */
double  orc_find_root_vr(orc_state *s, int64_t index, int64_t k, double E, double J);   /* cmc_nr.c:500-603 */

typedef struct { double rp, ra, dQdrp, dQdra; int64_t kmax, kmin; int32_t circular_flag; } orc_orbit_rs;
orc_orbit_rs orc_calc_orbit_new(orc_state *s, int64_t index, double E, double J);       /* cmc_orbit.c:413-540 */

double  orc_get_positions(orc_state *s);                       /* cmc_evolution_thr.c:784-1023 */
double  orc_calc_n_local(const orc_state *s, int64_t k, int64_t p, int64_t N_LIMIT);   /* cmc_dynamics_helper.c:768-786 */
void    orc_calc_sigma_r(orc_state *s);                        /* cmc_dynamics_helper.c:2821-2986 (P=1) */
double  orc_simul_relax_new(const orc_state *s);               /* cmc_dynamics_helper.c:2526-2579 */
void    orc_dynamics_apply(orc_state *s, double dt);           /* cmc_dynamics.c:17-542 (relaxation branch) */
int64_t orc_sort_by_r(orc_state *s);                           /* cmc_sort.c:375-629 + cmc_utils.c:2232-2288 (P=1) */

/* step glue (cmc.c:240-407 order, configs' switch set) */
void    orc_energy_conservation1(orc_state *s);                /* cmc_utils.c:2153-2172 */
void    orc_energy_conservation2(orc_state *s);                /* cmc_utils.c:2217-2227 */
void    orc_tidally_strip_stars(orc_state *s);                 /* cmc_evolution_thr.c:443-646 (TIDAL_TREATMENT 0) */
void    orc_compute_intermediate_energy(orc_state *s);         /* cmc_utils.c:368-391 */
void    orc_set_velocities3(orc_state *s);                     /* cmc_utils.c:254-354 (P=1) */
void    orc_compute_energy(orc_state *s);                      /* cmc_utils.c:604-653 */
double  orc_calc_timestep(orc_state *s);                       /* cmc_utils.c:2132-2148 + GetTimeStep */
void    orc_init(orc_state *s);                                /* cmc.c:151-165 subset: potential, sigma, energy */
void    orc_step(orc_state *s);                                /* one Henon step, cmc.c:240-407 */

/* ---------------------------------------------------------------- P ranks (CPU baseline)
 * The reference's MPI layout on P threads of one host: findLimits chunks
 * (cmc_utils.c:2299-2333, MIN_CHUNK_SIZE = 20), rank i drawing from its own sequential
 * stream jumped i * 2^80 from the seeded state (set_rng_states, :2364-2374), the
 * relaxation and orbit phases run by the ranks on their chunks, the sort as a sample
 * sort, and the rest as at one rank.  At P = 1 it is orc_step in REF_SEQ mode. */
#define ORC_MAX_RANKS 256
typedef struct orc_ranks {
    orc_state *s;
    int P;
    orc_t113 st[ORC_MAX_RANKS];  /* each rank's curr_st */
    int64_t start[ORC_MAX_RANKS], end[ORC_MAX_RANKS];
} orc_ranks;
void orc_find_limits(int64_t N, int64_t blk, int P, int64_t *start, int64_t *end);
int  orc_ranks_init(orc_ranks *R, orc_state *s, int P);
void orc_step_ranks(orc_ranks *R);
int64_t orc_sort_by_r_ranks(orc_state *s, int P);
orc_ranks *orc_ranks_new(orc_state *s, int P);      /* heap-allocated orc_ranks_init */
void orc_ranks_free(orc_ranks *R);

#ifdef __cplusplus
}
#endif
#endif
