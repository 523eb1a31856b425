/*
 * cmcg — MI355X (gfx950) kernels for the CMC-COSMIC Henon-step hot path, behind a
 * plain C ABI.  No torch or HIP types appear in these signatures: contexts are
 * opaque, buffers are host pointers with element counts, errors are negative ints.
 *
 * Index conventions follow the reference (ClusterMonteCarlo/CMC-COSMIC):
 *   - the global arrays star_r / star_m / star_phi (include/cmc/cmc_vars.h:218-226)
 *     hold N+2 entries; index 0 is the central sentinel r = ZERO = 1e-20 and index
 *     N+1 the outer sentinel r = SF_INFINITY = 1e10 (include/cmc/cmc.h:58,60);
 *   - per-star fields (the AoS star_t of include/cmc/cmc.h:409-601) use slots 1..N;
 *   - stored masses are m * N_STAR and madhoc = 1/N_STAR (src/cmc/cmc_fits.c:99,123,
 *     src/cmc/cmc_utils.c:1183).
 *
 * One context = one device stream.  Phase calls are asynchronous on that stream;
 * scalars they return are copied back at the end of the call (which synchronises
 * the stream) unless the *_async variant is used.  A context is not thread-safe.
 *
 * Random numbers: each star (relaxation pair, orbit-sampling star) draws from its
 * own substream of the reference's taus113 generator (src/libs/taus113-v2.c),
 * reached by GF(2) jump-ahead from the state seeded by reset_rng_t113_new(IDUM):
 *   relaxation, step s, pair p (stars 2p+1, 2p+2):  offset s*2^72 + 3p
 *   orbits,     step s, star slot i:                 offset s*2^72 + 2^70 + i*2^20
 * Step 0's relaxation therefore draws exactly the reference's sequential stream.
 */
#ifndef CMCG_H
#define CMCG_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ errors */
/* Error codes map the reference's fail-stop sites (exit_cleanly, src/cmc/cmc_utils.c:208-220). */
#define CMCG_OK            0
#define CMCG_E_NAN        -1   /* NaN in potential or Q (cmc_utils.c:688-691,721-733; cmc_evolution_thr.c:880-883) */
#define CMCG_E_NTRY       -2   /* N_TRY exceeded in rejection sampling (cmc_evolution_thr.c:890-893) */
#define CMCG_E_BRACKET    -3   /* FindZero_r / find_zero_Q bracket failure (cmc_utils.c:144-152) */
#define CMCG_E_WZERO      -4   /* W == 0 in calc_encounter_dyns (cmc_dynamics_helper.c:858-863) */
#define CMCG_E_WPZERO     -5   /* wp == 0 in dynamics_apply (cmc_dynamics.c:450-453) */
#define CMCG_E_BRENT      -6   /* find_root_vr failed (cmc_nr.c:510-517,577-584,597-600) */
#define CMCG_E_TIES       -7   /* (no longer returned: the sort orders runs of equal radii of any length) */
#define CMCG_E_HIP       -10   /* HIP runtime error */
#define CMCG_E_ARG       -11   /* invalid argument */
#define CMCG_E_CAPACITY  -12   /* more stars than the context capacity */
#define CMCG_E_NOMEM     -13   /* device allocation failed */
#define CMCG_E_STATE     -14   /* call out of order (e.g. no bracket model loaded) */
#define CMCG_E_STRIP     -15   /* tidal stripping did not end within the passes of a captured step */

const char *cmcg_error_string(int code);

/* ------------------------------------------------------------------ types */
typedef struct cmcg_ctx cmcg_ctx;

/* Run parameters.  Defaults follow the reference parser (src/cmc/cmc_io.c:1659-2022). */
typedef struct cmcg_params {
    int64_t n_star;        /* clus.N_STAR; madhoc = 1/n_star */
    int64_t avekernel;     /* AVEKERNEL (cmc_io.c:1749), default 20 */
    double  theta_se_max;  /* THETASEMAX (cmc_io.c:1671), default 1.412 */
    double  gamma;         /* CMC_GAMMA Coulomb-log factor (cmc_io.c:1252-1255) */
    double  orbit_r;       /* R_MAX = input RTID (cmc_io.c:2071, cmc.c:145) */
    double  cenma_m;       /* central mass in stored units (0 in these configs) */
    uint64_t seed;         /* IDUM (cmc_io.c:1736); seed 0 means 1 */
    int32_t  tidal_strip;  /* TIDALLY_STRIP_STARS (cmc_io.c:1670), default 1 */
    int32_t  flags;        /* CMCG_PF_* */
} cmcg_params;

#define CMCG_PF_LEARNED_BRACKETS 1   /* use the loaded bracket model in cmcg_orbits */

/* Host-side SoA view used by upload/download.  Any pointer may be NULL (field
 * skipped).  Arrays r, m, phi have n+2 entries (0..n+1); per-star arrays have n+2
 * entries too, slot 0 and n+1 unused.  Borrowed only for the duration of a call. */
typedef struct cmcg_host_soa {
    int64_t n;
    double *r, *m, *phi;
    double *vr, *vt, *E, *J;
    int64_t *id;
    double *rnew, *vrnew, *vtnew, *X, *Y, *r_peri, *r_apo;
    double *rOld, *Uoldrold, *Uoldrnew, *vrold, *vtold, *EI;
    double *sigma;
    double *pot_rnew;      /* potential(rnew) seen by the orbit phase (diagnostic) */
    int32_t *ntry;         /* rejection tries of the last orbit phase */
    int32_t *flags;        /* CMCG_SF_* outcome bits of the last orbit phase */
} cmcg_host_soa;

#define CMCG_SF_CIRCULAR   1
#define CMCG_SF_REMOVED_E  2
#define CMCG_SF_REMOVED_RT 4
#define CMCG_SF_DESTROYED  8
#define CMCG_SF_BRENT     16
#define CMCG_SF_LINSCAN   32
#define CMCG_SF_STRIPPED  64
#define CMCG_SF_BRACKET_HIT 128   /* learned bracket bracketed both roots */

/* Cluster scalars kept on the device (the reference's globals of the same names). */
typedef struct cmcg_scalars {
    int64_t n_max, n_max_new;
    double Mtotal, Rtidal, max_r, Dt, Prev_Dt, TotalTime;
    double TidalMassLoss, TidalMassLoss_old, OldTidalMassLoss;
    double Eescaped, Jescaped, Etidal, Eoops;
    double Etot, Ekin, Epot;
    uint64_t step;
} cmcg_scalars;

typedef struct cmcg_orbit_stats {
    int64_t n_circular, n_removed_E, n_removed_rt, n_destroyed, n_brent, n_linscan;
    int64_t tries_total, tries_max;
    int64_t n_bracket_hit, n_bracket_fallback;   /* learned brackets */
    double  max_r, tidal_mass_loss, E_escaped, J_escaped;
    int32_t error; int32_t pad;
    int64_t error_index;
} cmcg_orbit_stats;

/* ------------------------------------------------------------------ lifetime */
/* Replaces the allocation of the star arrays (cmc_io.c:2099-2104,3434-3436). */
int cmcg_create(cmcg_ctx **out, int device, int64_t n_cap, const cmcg_params *params);
int cmcg_destroy(cmcg_ctx *ctx);
int cmcg_upload(cmcg_ctx *ctx, const cmcg_host_soa *h);      /* sets N_MAX = N_MAX_NEW = h->n */
int cmcg_download(cmcg_ctx *ctx, cmcg_host_soa *h);          /* h->n receives N_MAX_NEW */
int cmcg_get_scalars(cmcg_ctx *ctx, cmcg_scalars *out);
int cmcg_set_scalars(cmcg_ctx *ctx, const cmcg_scalars *in);
int cmcg_sync(cmcg_ctx *ctx);

/* Checkpoints of the device state; replace save_restart_file / load_restart_file
 * (src/cmc/cmc_io.c:3829-3884, 3886-4031) for the hot path's state: every per-star
 * field of cmcg_host_soa, the device scalars and the step counter (which selects the
 * RNG substreams, the counterpart of saving curr_st), and the context parameters.
 * A context created with the same cmcg_params and loaded from a checkpoint continues
 * bit for bit as the saved one.  The buffer is host memory owned by the caller:
 * cmcg_checkpoint_size gives the bytes _save writes (CMCG_E_CAPACITY if cap is
 * smaller); _info reads the parameters and N from a buffer without a context; _load
 * returns CMCG_E_ARG for a malformed buffer or different parameters (flags aside). */
int cmcg_checkpoint_size(cmcg_ctx *ctx, int64_t *bytes);
int cmcg_checkpoint_save(cmcg_ctx *ctx, void *buf, int64_t cap);
int cmcg_checkpoint_info(const void *buf, int64_t size, cmcg_params *params, int64_t *n_max);
int cmcg_checkpoint_load(cmcg_ctx *ctx, const void *buf, int64_t size);

/* ------------------------------------------------------------------ hot-path phases */
/* a1: qsorts_new + remove_stripped_stars + post_sort_comm
 *     (src/cmc/cmc_utils.c:2232-2288, src/cmc/cmc_sort.c:245-250,343-354,375-629).
 *     Sorts slots 1..N_MAX_NEW by (r, id) ascending, drops the r == SF_INFINITY tail. */
int cmcg_sort_by_r(cmcg_ctx *ctx, int64_t *n_max);

/* a2: calc_potential_new -> potential_calculate (src/cmc/cmc_utils.c:675-736,2006-2012). */
int cmcg_potential(cmcg_ctx *ctx, int64_t *n_max, double *mtotal, double *rtidal);

/* a3: potential(r) (src/cmc/cmc_utils.c:112-165) for a batch of radii (host arrays). */
int cmcg_potential_at(cmcg_ctx *ctx, int64_t count, const double *r, double *out);

/* a10: simul_relax_new (src/cmc/cmc_dynamics_helper.c:2526-2579). */
int cmcg_simul_relax(cmcg_ctx *ctx, double *dtrel);
/* calc_timestep (src/cmc/cmc_utils.c:2132-2148 + GetTimeStep cmc_evolution_thr.c:317-438):
 * sets Dt and TotalTime on the device. */
int cmcg_calc_timestep(cmcg_ctx *ctx, double *dt);

/* calc_sigma_r (src/cmc/cmc_dynamics_helper.c:2821-2986) into the sigma array. */
int cmcg_sigma_r(cmcg_ctx *ctx);

/* a4: dynamics_apply relaxation branch (src/cmc/cmc_dynamics.c:17-542). step selects
 * the random substreams. Includes calc_sigma_r as the reference does (:37). */
int cmcg_relax(cmcg_ctx *ctx, double dt, uint64_t step);

/* a5: new_orbits_calculate -> get_positions (src/cmc/cmc_utils.c:2194-2211,
 *     src/cmc/cmc_evolution_thr.c:784-1023). */
int cmcg_orbits(cmcg_ctx *ctx, uint64_t step, cmcg_orbit_stats *stats);

/* a6: calc_orbit_new (src/cmc/cmc_orbit.c:413-540; also known as
 *     calc_pericenter_and_apocenter) for slots first..first+count-1 with the given
 *     E (already including the self term, cmc_evolution_thr.c:806) and J.
 *     Outputs may be NULL. */
int cmcg_calc_pericenter_and_apocenter(cmcg_ctx *ctx, int64_t first, int64_t count,
                                       const double *E, const double *J,
                                       double *rp, double *ra, double *dQdrp, double *dQdra,
                                       int64_t *kmin, int64_t *kmax, int32_t *circular);

/* The same as cmcg_calc_pericenter_and_apocenter, without any device or host
 * allocation: no counterpart in the reference.  For callers that cannot allocate
 * (tight loops that query apsides, code that must not call hipMalloc).  Its inputs
 * and outputs are staged in one fixed device array per GPU (about 2.4 MB, held until
 * the process ends), so count may be at most CMCG_APSIDES_NOALLOC_MAX
 * (CMCG_E_CAPACITY beyond), and calls that overlap on one GPU race for it.  Returns
 * CMCG_E_HIP if the array could not be located when the first context was created
 * on that device. */
#define CMCG_APSIDES_NOALLOC_MAX 32768
int cmcg_calc_pericenter_and_apocenter_noalloc(cmcg_ctx *ctx, int64_t first, int64_t count,
                                               const double *E, const double *J,
                                               double *rp, double *ra, double *dQdrp, double *dQdra,
                                               int64_t *kmin, int64_t *kmax, int32_t *circular);

/* a7: find_zero_Q (src/cmc/cmc_nr.c:316-383) for a batch of (j, kmin, kmax, E, J). */
int cmcg_find_zero_Q(cmcg_ctx *ctx, int64_t count, const int64_t *j, const int64_t *kmin,
                     const int64_t *kmax, const double *E, const double *J, int64_t *out);

/* a8: find_root_vr (src/cmc/cmc_nr.c:500-603, with calc_vr_within_interval :393-403 and
 *     calc_pot_within_interval :239-269) for a batch of (index j, interval k, E, J):
 *     the root of v_r^2 on [r_k, r_{k+1}] by the Brent-Dekker method to the reference's
 *     APSIDES_PRECISION, then its residual fix-up.  CMCG_E_BRENT where the reference
 *     exits (same signs at the ends, no convergence, negative residual). */
int cmcg_find_root_vr(cmcg_ctx *ctx, int64_t count, const int64_t *j, const int64_t *k, const double *E,
                      const double *J, double *out);

/* ------------------------------------------------------------------ step glue */
int cmcg_energy_conservation1(cmcg_ctx *ctx);        /* cmc_utils.c:2153-2172 */
int cmcg_energy_conservation2(cmcg_ctx *ctx);        /* cmc_utils.c:2217-2227 */
int cmcg_tidally_strip_stars(cmcg_ctx *ctx);         /* cmc_evolution_thr.c:443-646 */
int cmcg_compute_intermediate_energy(cmcg_ctx *ctx); /* cmc_utils.c:368-391 */
int cmcg_set_velocities3(cmcg_ctx *ctx);             /* cmc_utils.c:254-354 */
int cmcg_compute_energy(cmcg_ctx *ctx);              /* cmc_utils.c:604-653, 2017-2028 */
/* Initial potential, sigma and energies (cmc.c:151-165). */
int cmcg_init_cluster(cmcg_ctx *ctx);
/* One whole Henon step in the order of src/cmc/cmc.c:240-407.  Every field the
 * reference carries from one step to the next (r, m, phi, vr, vt, E, J, id, r_peri,
 * r_apo) ends in the reference's order and values.  The step's own scratch, which the
 * next step rewrites before any read (rnew, vrnew, vtnew, X, Y, rOld, Uoldrold,
 * Uoldrnew, vrold, vtold, EI, sigma and the per-star counters), is not carried
 * through the step's sort and is left in the order the step began with. */
int cmcg_step(cmcg_ctx *ctx, cmcg_orbit_stats *stats);
/* The same step replayed from a HIP graph (captured on first use; the step
 * alternates between two buffer states, so two graphs are kept).  The tidal
 * stripping loop runs inside the graph as three guarded grid passes and a one-block
 * tail that plays the grid's blocks in turn, so the sums add in the eager loop's
 * order (CMCG_E_STRIP if it had not ended after 64 further passes).  Results are identical to cmcg_step.  No counterpart in the reference:
 * for callers that step many small clusters, where launch overhead dominates. */
int cmcg_step_graph(cmcg_ctx *ctx, cmcg_orbit_stats *stats);
/* One step of each of `count` clusters (distinct contexts on one device) as one
 * batched step: every cluster's step is recorded once per buffer state, the launches
 * of each kernel are merged over the clusters into one segmented launch (the
 * clusters' grids laid end to end; a block finds its cluster in a table of block
 * offsets), and the merged sequence is replayed from one HIP graph on the first
 * context's stream, ordered after every context's earlier work.  Each cluster's
 * results are identical to its own cmcg_step.  Plans for the last few cluster sets
 * and buffer states are kept (rebuilt after anything that drops a context's step
 * graphs).  At most 65536 clusters.  stats/errors (optional) receive one entry per
 * cluster; returns the first error.  No counterpart in the reference (an ensemble
 * driver). */
int cmcg_ensemble_step(cmcg_ctx **ctxs, int count, cmcg_orbit_stats *stats, int32_t *errors);
/* The same ensemble step with each cluster's own step graph launched on its own
 * stream before any wait (the clusters run concurrently, kernel by kernel). */
int cmcg_ensemble_step_concurrent(cmcg_ctx **ctxs, int count, cmcg_orbit_stats *stats, int32_t *errors);
/* Of the last cmcg_ensemble_step: the merged launches its graph holds and the
 * kernel launches recorded over all its clusters. */
int cmcg_ensemble_stats(int64_t *launches, int64_t *recorded);

/* ------------------------------------------------------------------ RNG access */
/* Draw `ndraw` consecutive doubles of the substream of each of `count` consecutive
 * pairs/stars starting at `first` (phase 0 = relaxation pair index, 1 = orbit slot).
 * out is count*ndraw doubles, pair/star-major. */
int cmcg_rng_draws(cmcg_ctx *ctx, int phase, uint64_t step, int64_t first, int64_t count,
                   int ndraw, double *out);

/* ------------------------------------------------------------------ autotuning */
/* Time every launch shape of the orbit sampler (threads per block x LDS samples of
 * the (r, phi) search table) `reps` times on the context's current state and keep
 * the fastest for later steps.  The state going into the next step is unchanged
 * (the phase is rerun on the same inputs, its scalars are restored, and only its own
 * outputs - r_peri, r_apo, X, rnew, ... - are left from the last trial, which the next
 * step recomputes); per-star results do not depend on the shape.  No counterpart in
 * the reference. */
int cmcg_autotune(cmcg_ctx *ctx, int reps, int32_t *variant, int32_t *threads, int32_t *coarse,
                  double *ms_best, double *ms_default);
/* The same for the relaxation pass (with energy_conservation1 fused, as in the step):
 * times its block sizes (128, 256, 512, 1024 threads) on the context's current state,
 * restoring the fields and scalars it rewrites after every trial, and keeps the
 * fastest for the following steps; per-pair results do not depend on it.  No
 * counterpart in the reference. */
int cmcg_autotune_relax(cmcg_ctx *ctx, int reps, int32_t *threads, double *ms_best, double *ms_default);
/* The orbit sampler's launch variants (threads per block, LDS samples, most lanes one
 * star's rejection tries get per round, blocks per CU its register allocation must
 * allow: 1 = unconstrained) and their times in the last cmcg_autotune
 * (0 before one ran); fills at most max entries, *count = number of variants.
 * No counterpart in the reference. */
/* comp_mass_percent (src/cmc/cmc_utils.c:872-982) on one device: for the mass
 * fractions mass_pc[0..count-1] (ascending, count <= 256) the Lagrange radius, mean
 * mass, star count, density and the kinetic-energy and v^2 sums inside it, with
 * the reference's sequential sums and its rule of at most one fraction per star.
 * Host arrays of count entries (any may be NULL); entries the reference's loop would
 * not reach are left unchanged, and `reached` tells how many were written.
 * bh_include is MASS_PC_BH_INCLUDE, minimum_r is MINIMUM_R (both only matter with a
 * central mass), initial_total_mass scales ave_mass_r.  densities_r uses x^3 rounded
 * once where the reference calls pow(r, 3). */
typedef struct cmcg_mass_pc {
    double *mass_r, *ave_mass_r, *densities_r, *ke_rad_r, *ke_tan_r, *v2_rad_r, *v2_tan_r;
    int64_t *no_star_r;
    int32_t reached, pad;
} cmcg_mass_pc;
int cmcg_comp_mass_percent(cmcg_ctx *ctx, const double *mass_pc, int32_t count, int32_t bh_include,
                           double initial_total_mass, double minimum_r, cmcg_mass_pc *out);

/* Core quantities; replaces central_calculate (src/cmc/cmc_utils.c:1239-1345, the
 * Casertano & Hut (1985) density-weighted averages over the stars out to the
 * half-mass radius, J = 6) and clusdyn_calculate (:1614-1624, rh) for one rank.
 * Every sum is the reference's sequential recurrence, bit for bit.  The block after
 * them (NUM_CENTRAL_STARS averages for collision and binary rates) belongs to the
 * physics modules outside the hot path and is not computed.  Returns CMCG_E_STATE
 * where the reference exits (N_STAR <= 12 or nave >= N_STAR - 6, or the half-mass
 * star's window passes the live stars), CMCG_E_NAN for an infinite, NaN or negative
 * density term (r not ascending). */
typedef struct cmcg_central {
    int64_t nave;          /* stars averaged: the one just beyond the half-mass radius */
    double rh;             /* clusdyn.rh = star_r[nave] */
    double rhojsum, rhoj2sum;
    double rho;            /* central.rho (rho_core) */
    double v_rms;          /* central.v_rms (v_core) */
    double rc;             /* central.rc (core_radius) */
    double rc_nb;          /* rc_nb */
    double m_ave, n, rc_spitzer;
    double N_core, N_core_nb, Trc;
} cmcg_central;
int cmcg_central_calculate(cmcg_ctx *ctx, cmcg_central *out);

/* Test hooks (no counterpart in the reference; for tests that force rare paths).
 * CMCG_HOOK_CHAIN_K_DROP: widen the error bound with which the exact-chain emulation
 * (potential, comp_mass_percent, central_calculate) classifies its elements by 2^value
 * (0..92; 0 = normal).  At 30 or more nearly every element is then treated as
 * sensitive and takes the fallback paths (ordered compaction, list overflow,
 * sequential repair); the results must still equal the reference's recurrences bit
 * for bit, only slower. */
#define CMCG_HOOK_CHAIN_K_DROP 1
/* CMCG_HOOK_ORBIT_VARIANT: run the orbit sampler with launch variant `value`
 * (0 .. count-1 of cmcg_orbit_variants), as cmcg_autotune would choose it. */
#define CMCG_HOOK_ORBIT_VARIANT 2
/* CMCG_HOOK_MLP_MFMA: value 1 makes cmcg_mlp_predict run the network on the matrix
 * cores in f16 (the kernel that proposes the orbit phase's learned brackets), 0 the
 * per-thread FP32 kernel. */
#define CMCG_HOOK_MLP_MFMA 3
/* CMCG_HOOK_PERFECT_BRACKETS: value 1 copies the apsides that the last orbit phase
 * wrote (r_peri, r_apo) into the learned-bracket proposals and keeps them (no network
 * inference) for the following orbit phases with learned brackets on: a perfect
 * predictor, to measure the most any learned proposal could save.  0 returns to the
 * network.  Needs cmcg_mlp_init. */
#define CMCG_HOOK_PERFECT_BRACKETS 4
/* CMCG_HOOK_COLD_PHASES: value > 0 makes every timed phase of an eager step start
 * with cold caches: before the phase's first event a kernel reads and rewrites a
 * `value` MiB device buffer (allocated here), evicting the step's data from the L2s
 * and the MALL; the flush lies outside the phase's events.  0 frees the buffer. */
#define CMCG_HOOK_COLD_PHASES 5
/* CMCG_HOOK_RELAX_SHAPE: run the relaxation pass with `value` threads per block
 * (128, 256, 512 or 1024), as cmcg_autotune_relax would choose it. */
#define CMCG_HOOK_RELAX_SHAPE 6
int cmcg_test_hook(cmcg_ctx *ctx, int32_t hook, int64_t value);

int cmcg_orbit_variants(cmcg_ctx *ctx, int32_t max, int32_t *count, int32_t *threads, int32_t *coarse,
                        int32_t *cap, int32_t *min_blocks, double *ms_last);

/* ------------------------------------------------------------------ timing */
/* Average device time (ms) of the named kernel family over the calls since the last
 * reset, measured with HIP events on the context stream. names: "sort", "potential",
 * "relax", "orbits", "glue", "step". */
int cmcg_timer_enable(cmcg_ctx *ctx, int on);
int cmcg_timer_get(cmcg_ctx *ctx, const char *name, double *ms_total, int64_t *calls);
int cmcg_timer_reset(cmcg_ctx *ctx);

/* ------------------------------------------------------------------ learned brackets */
/* A 6 -> 64 -> 64 -> 2 SiLU network (FP32) predicting (log r_peri, log r_apo) of a star
 * from (E, J, log r, phi(r), log(r^2 dphi/dr), m).  No counterpart in the reference:
 * it proposes the starting bracket of find_zero_Q (cmc_nr.c:316-383); the exact index
 * bisection decides the result and the full bracket is used whenever the proposal
 * does not bracket the sign change.
 * Parameters: W1[64][6], b1[64], W2[64][64], b2[64], W3[2][64], b3[2] (4738 floats);
 * norm: input means[6], input std[6], target means[2], target std[2] (16 floats).
 * Training rows: 6 inputs then 2 targets (8 floats). */
#define CMCG_MLP_NPARAM 4738
#define CMCG_MLP_NNORM 16
/* Allocate the network (uniform(+-1/sqrt(fan_in)) initial weights from `seed`), its
 * Adam state and a ring buffer of `max_rows` training rows. */
int cmcg_mlp_init(cmcg_ctx *ctx, uint64_t seed, int64_t max_rows);
int cmcg_mlp_get_params(cmcg_ctx *ctx, float *theta, float *norm);
int cmcg_mlp_set_params(cmcg_ctx *ctx, const float *theta, const float *norm);
/* on != 0: after every orbit phase, append each sampled star's inputs and exact
 * apsides to the ring buffer. */
int cmcg_mlp_harvest(cmcg_ctx *ctx, int on);
/* Number of stored rows; when rows != NULL also copies them (count x 8 floats). */
int cmcg_mlp_rows(cmcg_ctx *ctx, int64_t *count, float *rows);
/* Fit the input/target standardisation to the stored rows. */
int cmcg_mlp_fit_norm(cmcg_ctx *ctx);
/* Mean squared error (standardised targets, summed over the two outputs) of the n
 * given rows and its gradient with respect to the parameters (one batch). */
int cmcg_mlp_loss_grad(cmcg_ctx *ctx, const float *rows, int64_t n, double *loss, float *grad);
/* `iters` Adam steps (lr, beta1 0.9, beta2 0.999) on minibatches of `batch` stored rows
 * in a shuffled order drawn from `seed`; *loss = loss of the last minibatch. */
int cmcg_mlp_train(cmcg_ctx *ctx, int64_t iters, int64_t batch, float lr, uint64_t seed, double *loss);
/* y[2n] = predicted (log r_peri, log r_apo) for x[6n] raw inputs. */
int cmcg_mlp_predict(cmcg_ctx *ctx, const float *x, int64_t n, float *y);
/* on != 0: the orbit phase starts find_zero_Q from the predicted apsides widened by
 * +-margin in log r (0 < margin < 10). */
int cmcg_set_learned_brackets(cmcg_ctx *ctx, int on, double margin);

#ifdef __cplusplus
}
#endif
#endif
