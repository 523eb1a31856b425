// Host-side launch entry points of the cmcg kernels (one translation unit per
// phase).  All launches are asynchronous on the given stream; they read the live
// star count from DevScalars on the device, and `n_ub` is only an upper bound used
// to size grids (stars beyond the device-side count exit immediately), so a whole
// step can be enqueued without host round trips.
#pragma once
#include "cmcg_internal.hpp"
#include "mlp.hpp"

// orbit sampler: most LDS samples of any launch variant; lane jumps l 2^20 by nibble
// tables of the powers 2^ORB_NIB_B0 .. 2^(ORB_NIB_B0 + ORB_NIB_POWERS - 1)
constexpr int ORB_MAX_COARSE = 2048;
// sort payload records: 16 words (128 B, one cache line) per star
constexpr int SORT_AOS_WORDS = 16;
constexpr int SORT_RUN_SHORT = 64;   // runs of equal sort keys up to this length are ordered in place
constexpr int SORT_TILE = 4096;      // words per tile of the radix passes (2048: 4% slower)
constexpr int SORT_RUN_LONG = 4096;  // ... up to this length in LDS; longer ones by a merge sort
constexpr int SORT_MM_PARTS = 1024;   // blocks of the sort's min/max pass, one (min, max) pair each
constexpr int RED2_BLOCKS = 64;
constexpr int ORB_NIB_B0 = 20, ORB_NIB_POWERS = 6;

constexpr int RNG_WB_POWERS = 10;   // nibble-table powers of the wave-base jumps
constexpr int RIDX_TILE = 4096;     // buckets per block of the radial index's scan
__host__ __device__ constexpr int64_t ridx_entries(int bits) { return ((int64_t)1 << bits) + 2; }
__host__ __device__ constexpr int64_t ridx_tiles(int bits) { return (ridx_entries(bits) + RIDX_TILE - 1) / RIDX_TILE; }

struct Scratch {
    // potential chains: per-tile int128 sums and prefixes plus a boundary value
    // (48 B per tile), chain increments / iterates / enclosed masses per star
    double *tile_sums;         // 6 words per tile (potential chains)
    int64_t max_tiles;
    double *chain_x, *chain_g, *chain_mk;
    double *chain_scal;         // 4 doubles
    int64_t *chain_len;         // 4
    int64_t *chain_fbad;        // 6
    // potential chains: list of sensitive elements (index, int128 prefix, value)
    int64_t *esa_idx;
    double *esa_racc;           // regular sum before each listed element
    double *esa_v;
    int64_t *esa_n;             // [0] sensitive count, [1] appended count, [2] two arrival counters
    int64_t *esa_raw_idx;       // appended sensitive elements (unordered), 2048
    double *esa_raw_part;       // their in-tile regular sums (int64), binades (int32) and s x
    int64_t esa_max;
    // deterministic block partials: 8 doubles per block
    double *partials;
    int64_t max_blocks;
    // sort
    uint64_t *keys_a, *keys_b;  // sort words: 32-bit normalised key of r above the 32-bit slot
    uint32_t *vals_a;           // sorted permutation: slot of the star now at each position
    uint64_t *sort_mm;          // live key (min, max) per block of the min/max pass: 2 * SORT_MM_PARTS
    uint32_t *block_hist;       // 256 * max_sort_tiles
    uint32_t *digit_totals;     // 8 * 256
    int64_t max_sort_tiles;
    uint64_t *sort_aos;         // packed payload records, SORT_AOS_WORDS words per star
    int64_t *sort_long;         // starts of runs of equal key prefixes longer than SORT_RUN_SHORT
    int64_t *sort_nlong;        // number of listed runs
    int64_t sort_long_cap;
    double2 *mid;               // (r, phi) at every 16th node (search middle level)
    double2 *nodes;             // packed (r, phi) of every node, for the searches
    float2 *pred;               // learned apsides proposals (log r_peri, log r_apo) per star
    uint32_t *ridx;             // radial index: nodes below each bucket of r (ridx_nb + 2 entries)
    uint32_t *ridx_tb;          // its value at every RIDX_TILE-th bucket
    int ridx_bits;              // log2 of the bucket count
    double2 *u_old;             // (new potential, radius) at each old sorted radius (set_velocities3)
    int64_t *merge_ends;        // set_velocities3: node index of both ends of each merge block
    int64_t merge_ends_cap;
    double *sv3_dpre, *sv3_dpost; // vold^2 + 2(1-q)(Uoldrold - Uoldrnew) per star, before and after the step's sort
    unsigned long long *dt_part;  // the step's time step: 64 slots, 128 bytes apart, of complemented dt minima
    double2 *coarse;            // orbit sampler's (r, phi) samples, copied into LDS by every block
    void *orb_nib;              // orbit sampler's GF(2) nibble tables (constant per context)
    void *relax_nib;            // relaxation's nibble tables, powers 2^0 .. 2^7 (constant per context)
    void *wb_nib;               // wave-base nibble tables: powers 2^6 .. 2^15 and 2^26 .. 2^35
    double *red2;               // second level of the block-partial reductions (RED2_BLOCKS x 16)
    int64_t sort_aos_cap;
    // set_velocities3 excess chain: per-star excess, ordered event list, per-event
    // walk ends and merge flags, per-block counts, fallback flag
    double *excess_c;
    int64_t *excess_list;
    int64_t *sv3_wend;
    int32_t *sv3_merged;
    int64_t *sv3_blk;
    int64_t *sv3_fallback;      // first presumed head that failed the check
    // tidal stripping loop flag
    int32_t *strip_done;
    // orbit stars deferred to the linear-scan path (65536 slots)
    int64_t *defer_list, *defer_kfirst;   // KParams::defer_cap entries
    // RNG bases of the current step: [0] relaxation, [1] orbits; seeded state
    Rng *rng_base;
    Rng *rng_seeded;
    Rng *rng_wave;              // per-wave substream starts (relaxation pairs / orbit slots)
    int64_t max_waves;
};

hipError_t launch_potential(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub,
                            const Scratch &sc, hipStream_t st);
hipError_t launch_mass_percent(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub, const Scratch &sc,
                               const double *pc, int count, int bh_include, double itm, double minimum_r,
                               void *out, hipStream_t st);
size_t mass_percent_out_bytes();
hipError_t launch_central(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub, const Scratch &sc,
                          void *out, hipStream_t st);
size_t central_out_bytes();
hipError_t launch_potential_at(const DevFields &f, DevScalars *s, const KParams &P,
                               const double *x, double *out, int64_t count, hipStream_t st);

hipError_t launch_sigma(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub, hipStream_t st);
hipError_t launch_simul_relax(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub, hipStream_t st,
                              bool init_min = true);
hipError_t launch_timestep_rng_bases(const RngTable *T, DevScalars *s, const Scratch &sc, hipStream_t st);
hipError_t launch_timestep_sigma(const DevFields &f, DevScalars *s, const KParams &P, const RngTable *T,
                                 const Scratch &sc, int64_t n_ub, hipStream_t st);
hipError_t launch_timestep_finalize(DevScalars *s, hipStream_t st);
// threads: block size of the relaxation kernel, one of RELAX_SHAPES (an autotuned
// launch shape; results do not depend on it)
constexpr int RELAX_SHAPES[] = {256, 128, 512, 1024};
constexpr int RELAX_NSHAPES = 4;
hipError_t launch_relax(const DevFields &f, DevScalars *s, const KParams &P, const RngTable *T,
                        const Scratch &sc, int64_t n_ub, hipStream_t st,
                        bool with_ec1 = false, int threads = 256);
hipError_t launch_rng_bases(const RngTable *T, DevScalars *s, const Scratch &sc, hipStream_t st);
hipError_t launch_rng_bases_step(const RngTable *T, DevScalars *s, const Scratch &sc, int64_t step, hipStream_t st);
// substream start of every 64-lane wave: relaxation (phase 0, pairs 64w..) or orbits
// (phase 1, slots 1 + 64w..), for `count` pairs / slots
hipError_t launch_rng_wave_bases(const RngTable *T, const Scratch &sc, int phase, int64_t count, hipStream_t st);
// the packed node table and the radial index in one pass over r and phi
hipError_t launch_node_index(const DevFields &f, DevScalars *s, const Scratch &sc, int64_t n_ub, hipStream_t st);
// middle level of the (r, phi) search tables, from the current arrays
hipError_t launch_mid_build(const DevFields &f, DevScalars *s, const Scratch &sc, int64_t n_ub, hipStream_t st);
hipError_t launch_rng_draws(const RngTable *T, const Rng *base, int phase, int64_t first, int64_t count,
                            int ndraw, double *out, hipStream_t st);

// mlp != nullptr: learned brackets (the network's proposals for every star first, on
// the matrix cores, into sc.pred)
hipError_t launch_orbits(const DevFields &f, DevScalars *s, const KParams &P, const RngTable *T,
                         const Scratch &sc, int64_t n_ub, hipStream_t st, const MlpState *mlp = nullptr,
                         float margin = 0.f, hipEvent_t ev_a = nullptr, hipEvent_t ev_b = nullptr, int variant = 0,
                         bool pred_fixed = false);
// proposals equal to the apsides the last orbit phase wrote (log r_peri, log r_apo): a
// perfect predictor, for measuring what any learned proposal could save
hipError_t launch_pred_from_apsides(const DevFields &f, DevScalars *s, const Scratch &sc, int64_t n_ub,
                                    hipStream_t st);
int orbit_variant_count();
void orbit_variant_shape(int v, int *threads, int *coarse, int *cap = nullptr, int *minb = nullptr);
// Device arrays receiving calc_orbit_new's results, one per output.
struct ApsidesOut {
    double *rp, *ra, *dQdrp, *dQdra;
    int64_t *kmin, *kmax;
    int32_t *circular;
};
hipError_t launch_apsides_soa(const DevFields &f, DevScalars *s, const KParams &P, int64_t first, int64_t count,
                              const double *E, const double *J, const ApsidesOut &o, hipStream_t st);
hipError_t launch_apsides(const DevFields &f, DevScalars *s, const KParams &P, int64_t first, int64_t count,
                          const double *E, const double *J, double *out4, int64_t *kout, int32_t *circ,
                          hipStream_t st);
hipError_t launch_find_root_vr(const DevFields &f, DevScalars *s, const KParams &P, int64_t count,
                               const int64_t *j, const int64_t *k, const double *E, const double *J, double *out,
                               hipStream_t st);
hipError_t launch_find_zero_Q(const DevFields &f, DevScalars *s, const KParams &P, int64_t count,
                              const int64_t *j, const int64_t *kmin, const int64_t *kmax,
                              const double *E, const double *J, int64_t *out, hipStream_t st);

hipError_t launch_sort_fields(const double *key_src, double *r_dst, const int64_t *id_src, int nf,
                              const uint64_t *const *src, uint64_t *const *dst, DevScalars *s, const Scratch &sc,
                              int64_t n_ub, hipStream_t st);

hipError_t launch_step_begin(DevScalars *s, hipStream_t st);
hipError_t launch_after_dynamics(DevScalars *s, hipStream_t st);
hipError_t launch_energy_conservation1(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub, hipStream_t st);
hipError_t launch_energy_conservation2(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub, hipStream_t st,
                                       bool with_ie = false, double *dpre = nullptr);
hipError_t launch_strip(const DevFields &f, DevScalars *s, const KParams &P, const Scratch &sc, int32_t *done,
                        int64_t n_ub, int passes, bool begin, hipStream_t st);
hipError_t launch_strip_tail(const DevFields &f, DevScalars *s, const KParams &P, const Scratch &sc, hipStream_t st);
hipError_t launch_intermediate_energy(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub, hipStream_t st);
hipError_t launch_ie_transfer(const DevFields &f, DevScalars *s, int64_t n_ub, hipStream_t st);
// r_old_sorted: the radii the sort of this step read (old sorted order), or nullptr
hipError_t launch_set_velocities3(const DevFields &f, DevScalars *s, const KParams &P, const Scratch &sc,
                                  int64_t n_ub, hipStream_t st, const double *r_old_sorted = nullptr,
                                  const double *dpost = nullptr);
hipError_t launch_compute_energy(const DevFields &f, DevScalars *s, const KParams &P, const Scratch &sc,
                                 int64_t n_ub, hipStream_t st);

// learned apsides brackets (k_mlp.hip)
int mlp_grid_blocks(int64_t batch);
hipError_t launch_mlp_loss_grad(const MlpState &m, const float *rows, int64_t count, int64_t first, int64_t batch,
                                uint64_t mult, uint64_t off, hipStream_t st);
hipError_t launch_mlp_adam(const MlpState &m, int64_t step, float lr, hipStream_t st);
hipError_t launch_mlp_predict(const MlpState &m, const float *x, int64_t n, float *y, hipStream_t st);
// the network on the matrix cores: every star's proposal (pred[j], j = 1..n), or given rows
hipError_t launch_mlp_infer_stars(const DevFields &f, DevScalars *s, const KParams &P, const MlpState &m,
                                  float2 *pred, int64_t n_ub, hipStream_t st);
hipError_t launch_mlp_infer_rows(const MlpState &m, const float *x, int64_t n, float *y, hipStream_t st);
hipError_t launch_mlp_fit_norm(const MlpState &m, int64_t count, hipStream_t st);
hipError_t launch_mlp_harvest(const DevFields &f, DevScalars *s, const KParams &P, const MlpState &m, int64_t n_ub,
                              hipStream_t st);
