// C-ABI of the cmcg library: context lifetime, host<->device transfers, the
// hot-path phase entry points and the whole Henon step (src/cmc/cmc.c:240-407).
#include "kernels.hpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

void cmcg_report_hip(hipError_t e, const char *expr, const char *file, int line)
{
    if (getenv("CMCG_VERBOSE")) fprintf(stderr, "cmcg: HIP error %s (%d) in %s at %s:%d\n", hipGetErrorString(e), (int)e, expr, file, line);
}

struct cmcg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n_cap = 0;
    cmcg_params params{};
    KParams kp{};
    DevFields f{}, alt{};
    DevScalars *s = nullptr;
    RngTable *T = nullptr;
    Scratch sc{};
    std::vector<void *> allocs;
    int64_t n_ub = 0;
    // timing
    bool timing = false;
    std::map<std::string, std::pair<double, int64_t>> timers;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    // host-side copy of the seeded RNG state
    Rng seeded{};
    // learned brackets
    MlpState mlp{};
    bool mlp_ready = false, harvest = false, learned = false;
    bool mlp_mfma = false;       // cmcg_mlp_predict on the matrix cores (test hook)
    bool pred_fixed = false;     // learned brackets from fixed proposals (test hook)
    float margin = 0.05f;        // half-width of the learned brackets in log r
    int64_t adam_t = 0;
    float *io = nullptr;
    int64_t io_cap = 0;          // floats
    // step graphs: the fused sort swaps buffer pointers, so steps alternate between
    // two pointer states; graph[p] is the step captured in state p
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int parity = 0;
    uint64_t serial = 0;         // unique per context (ensemble plans name their clusters by it)
    uint64_t gen = 0;            // bumped whenever the graphs are dropped
    int orb_variant = 0;         // launch shape of the orbit sampler (autotuned)
    int relax_threads = 256;     // block size of the relaxation kernel (autotuned)
    std::vector<double> tune_ms; // per-variant time of the last cmcg_autotune
    uint4 *flush_buf = nullptr;  // CMCG_HOOK_COLD_PHASES: evicted through before each timed phase
    size_t flush_words = 0;
    void *mpc_buf = nullptr;     // cmcg_comp_mass_percent results (allocated on first use)
    void *cen_buf = nullptr;     // cmcg_central_calculate results (allocated on first use)
};

// t[b][c][q][v] = XOR of the columns of matrix power b0 + b selected by nibble v at
// bits 4q..4q+3 of state word c (the device's build_nibble_tables, done once here)
template <int NB>
static void fill_nibble_tables(const RngTable &ht, int b0, NibbleTables<NB> &nt)
{
    for (int b = 0; b < NB; b++)
        for (int cc = 0; cc < 4; cc++)
            for (int q = 0; q < 8; q++)
                for (int v = 0; v < 16; v++) {
                    uint32_t o = 0;
                    for (int i = 0; i < 4; i++)
                        if ((v >> i) & 1) o ^= ht.col[b0 + b][cc][4 * q + i];
                    nt.t[b][cc][q][v] = o;
                }
}

static void invalidate_graphs(cmcg_ctx *c)
{
    for (auto &g : c->graph)
        if (g) { hipGraphExecDestroy(g); g = nullptr; }
    c->parity = 0;
    c->gen++;
}

namespace {

// ------------------------------------------------------------------ host RNG setup
// taus113 seeding, src/libs/taus113-v2.c:197-216
Rng host_seed(uint64_t s)
{
    auto step = [](Rng &st) {
        uint32_t b;
        b = ((st.z0 << 6) ^ st.z0) >> 13;  st.z0 = ((st.z0 & 4294967294u) << 18) ^ b;
        b = ((st.z1 << 2) ^ st.z1) >> 27;  st.z1 = ((st.z1 & 4294967288u) << 2) ^ b;
        b = ((st.z2 << 13) ^ st.z2) >> 21; st.z2 = ((st.z2 & 4294967280u) << 7) ^ b;
        b = ((st.z3 << 3) ^ st.z3) >> 12;  st.z3 = ((st.z3 & 4294967168u) << 13) ^ b;
    };
    if (s == 0) s = 1;
    Rng st;
    uint64_t z;
    z = (69069ull * s) & 0xffffffffull;     if (z < 2) z += 2;     st.z0 = (uint32_t)z;
    z = (69069ull * st.z0) & 0xffffffffull; if (z < 8) z += 8;     st.z1 = (uint32_t)z;
    z = (69069ull * st.z1) & 0xffffffffull; if (z < 16) z += 16;   st.z2 = (uint32_t)z;
    z = (69069ull * st.z2) & 0xffffffffull; if (z < 128) z += 128; st.z3 = (uint32_t)z;
    for (int i = 0; i < 10; i++) step(st);
    return st;
}

// Column i of the one-step map of component j applied to e_i.
uint32_t comp_step(int j, uint32_t z)
{
    switch (j) {
    case 0: return ((z & 4294967294u) << 18) ^ (((z << 6) ^ z) >> 13);
    case 1: return ((z & 4294967288u) << 2) ^ (((z << 2) ^ z) >> 27);
    case 2: return ((z & 4294967280u) << 7) ^ (((z << 13) ^ z) >> 21);
    default: return ((z & 4294967168u) << 13) ^ (((z << 3) ^ z) >> 12);
    }
}

uint32_t host_matvec(const uint32_t *col, uint32_t v)
{
    uint32_t o = 0;
    for (int i = 0; i < 32; i++)
        if ((v >> i) & 1u) o ^= col[i];
    return o;
}

void build_table(RngTable &t)
{
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 32; i++) t.col[0][j][i] = comp_step(j, 1u << i);
    for (int b = 1; b < 128; b++)
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 32; i++) t.col[b][j][i] = host_matvec(t.col[b - 1][j], t.col[b - 1][j][i]);
}

template <typename T>
int dalloc(cmcg_ctx *c, T **p, size_t count)
{
    void *q = nullptr;
    if (hipMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 8) != hipSuccess) return CMCG_E_NOMEM;
    if (hipMemset(q, 0, count * sizeof(T) > 0 ? count * sizeof(T) : 8) != hipSuccess) return CMCG_E_HIP;
    c->allocs.push_back(q);
    *p = (T *)q;
    return CMCG_OK;
}

int alloc_fields(cmcg_ctx *c, DevFields &f, bool full)
{
    size_t n = (size_t)c->n_cap + 6;     // slots 0..n_cap+1 plus 4 entries of padding
    double **dfields_sorted[] = {&f.r, &f.m, &f.vr, &f.vt, &f.E, &f.J, &f.rnew, &f.vrnew, &f.vtnew, &f.X, &f.Y,
                                 &f.r_peri, &f.r_apo, &f.rOld, &f.Uoldrold, &f.Uoldrnew, &f.vrold, &f.vtold, &f.EI};
    for (auto pp : dfields_sorted)
        if (int e = dalloc(c, pp, n)) return e;
    if (int e = dalloc(c, &f.id, n)) return e;
    if (full) {
        if (int e = dalloc(c, &f.phi, n)) return e;
        if (int e = dalloc(c, &f.sigma, n)) return e;
        if (int e = dalloc(c, &f.pot_rnew, n)) return e;
        if (int e = dalloc(c, &f.ntry, n)) return e;
        if (int e = dalloc(c, &f.flags, n)) return e;
    }
    return CMCG_OK;
}

int check_error(cmcg_ctx *c)
{
    int32_t err = 0;
    if (hipMemcpy(&err, &c->s->error, sizeof(err), hipMemcpyDeviceToHost) != hipSuccess) return CMCG_E_HIP;
    return err;
}

hipError_t sync(cmcg_ctx *c) { return hipStreamSynchronize(c->stream); }

hipEvent_t get_event(cmcg_ctx *c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
}

// reads and rewrites a buffer larger than the L2s and the MALL together
struct flush_kernel {
    static constexpr int LB = 256, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, uint4 *__restrict__ buf, size_t words)
    {
        for (size_t i = (size_t)kg_.bid * 256 + threadIdx.x; i < words; i += (size_t)kg_.nb * 256) {
            uint4 v = buf[i];
            v.x += 1u;
            buf[i] = v;
        }
    }
};

struct PhaseTimer {
    cmcg_ctx *c;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    PhaseTimer(cmcg_ctx *c_, const char *n) : c(c_), name(n)
    {
        if (c->timing) {
            if (c->flush_buf) kl<flush_kernel>(dim3(8192), dim3(256), 0, c->stream, c->flush_buf, c->flush_words);
            a = get_event(c);
            hipEventRecord(a, c->stream);
        }
    }
    ~PhaseTimer()
    {
        if (c->timing) {
            b = get_event(c);
            hipEventRecord(b, c->stream);
            c->pending.push_back({name, {a, b}});
        }
    }
};

// a pair of events bracketing one kernel launch, collected like a phase timer
void kernel_events(cmcg_ctx *c, const char *name, hipEvent_t &a, hipEvent_t &b)
{
    a = b = nullptr;
    if (!c->timing) return;
    a = get_event(c);
    b = get_event(c);
    c->pending.push_back({name, {a, b}});
}

void collect_timers(cmcg_ctx *c)
{
    for (auto &p : c->pending) {
        hipEventSynchronize(p.second.second);
        float ms = 0.f;
        hipEventElapsedTime(&ms, p.second.first, p.second.second);
        auto &acc = c->timers[p.first];
        acc.first += ms;
        acc.second += 1;
        c->event_pool.push_back(p.second.first);
        c->event_pool.push_back(p.second.second);
    }
    c->pending.clear();
}

struct set_dt_step_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s, double dt, int64_t step)
{
    if (dt == dt) s->Dt = dt;
    if (step >= 0) s->step = (uint64_t)step;
}
};
struct step_end_kernel { static constexpr int LB = 1024, MB = 0; static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s) { s->step += 1; } };

// payload field lists for the sort
struct SortSpec {
    int nf = 0;
    const uint64_t *src[24];
    uint64_t *dst[24];
    double **swap_a[24];
    double **swap_b[24];
    int64_t **swap_ia = nullptr, **swap_ib = nullptr;
};

// fa, fb: the live/alternate pointer pair swapped after the sort (null for a scratch
// array that the sort writes into a fixed destination)
void add_field(SortSpec &sp, double *src, double *dst, double **fa, double **fb)
{
    sp.src[sp.nf] = (const uint64_t *)src;
    sp.dst[sp.nf] = (uint64_t *)dst;
    sp.swap_a[sp.nf] = fa;
    sp.swap_b[sp.nf] = fb;
    sp.nf++;
}

// The payload fields of the sort and the live/alternate pointer pairs they swap.
static void sort_spec(cmcg_ctx *c, bool fused, SortSpec &sp)
{
    DevFields &f = c->f, &a = c->alt;
    if (fused) {
        // ComputeIntermediateEnergy's transfer (cmc_utils.c:382-390) fused into the gather:
        // r <- rnew (the key), rOld <- r, vr <- vrnew, vt <- vtnew.
        // The step's own scratch is not carried: set_velocities3 reads rOld through the
        // permutation and, of vtold, vrold, Uoldrold and Uoldrnew, only their
        // combination D, which energy_conservation2 formed and the sort moves (8-word
        // records).
        add_field(sp, f.m, a.m, &f.m, &a.m);
        add_field(sp, f.vrnew, a.vr, &f.vr, &a.vr);
        add_field(sp, f.vtnew, a.vt, &f.vt, &a.vt);
        add_field(sp, c->sc.sv3_dpre, c->sc.sv3_dpost, nullptr, nullptr);
        add_field(sp, f.r_apo, a.r_apo, &f.r_apo, &a.r_apo);
        add_field(sp, f.r_peri, a.r_peri, &f.r_peri, &a.r_peri);
    } else {
        double *DevFields::*flds[] = {&DevFields::m, &DevFields::vr, &DevFields::vt, &DevFields::E,
                                      &DevFields::J, &DevFields::rnew, &DevFields::vrnew, &DevFields::vtnew,
                                      &DevFields::X, &DevFields::Y, &DevFields::r_peri, &DevFields::r_apo,
                                      &DevFields::rOld, &DevFields::Uoldrold, &DevFields::Uoldrnew,
                                      &DevFields::vrold, &DevFields::vtold, &DevFields::EI};
        for (auto fp : flds) add_field(sp, f.*fp, a.*fp, &(f.*fp), &(a.*fp));
    }
    // id travels as 8-byte payload too
    sp.src[sp.nf] = (const uint64_t *)f.id;
    sp.dst[sp.nf] = (uint64_t *)a.id;
    sp.nf++;
}

static void swap_sorted(cmcg_ctx *c, const SortSpec &sp)
{
    std::swap(c->f.r, c->alt.r);
    for (int i = 0; i < sp.nf - 1; i++)
        if (sp.swap_a[i]) std::swap(*sp.swap_a[i], *sp.swap_b[i]);
    std::swap(c->f.id, c->alt.id);
}

// Runs the sort and swaps every destination field into the live set.
int run_sort(cmcg_ctx *c, bool fused)
{
    SortSpec sp;
    sort_spec(c, fused, sp);
    const double *key = fused ? c->f.rnew : c->f.r;
    hipError_t e = launch_sort_fields(key, c->alt.r, c->f.id, sp.nf, sp.src, sp.dst, c->s, c->sc, c->n_ub, c->stream);
    if (e != hipSuccess) { cmcg_report_hip(e, "launch_sort_fields", __FILE__, __LINE__); return CMCG_E_HIP; }
    swap_sorted(c, sp);
    // the old vr/vt buffers are now stale copies; vrnew/vtnew keep their values
    return CMCG_OK;
}

int read_scalars(cmcg_ctx *c, DevScalars &h)
{
    if (sync(c) != hipSuccess) return CMCG_E_HIP;
    if (hipMemcpy(&h, c->s, sizeof(DevScalars), hipMemcpyDeviceToHost) != hipSuccess) return CMCG_E_HIP;
    collect_timers(c);
    return CMCG_OK;
}

#define LAUNCH(expr)                                                                 \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) { cmcg_report_hip(_e, #expr, __FILE__, __LINE__); return CMCG_E_HIP; } \
    } while (0)

} // namespace

thread_local SegRecorder *g_seg_rec = nullptr;

static void ens_forget(uint64_t serial);

// ======================================================================= C ABI
extern "C" {

const char *cmcg_error_string(int code)
{
    switch (code) {
    case CMCG_OK: return "ok";
    case CMCG_E_NAN: return "NaN detected";
    case CMCG_E_NTRY: return "N_TRY exceeded";
    case CMCG_E_BRACKET: return "bisection bracket failure";
    case CMCG_E_WZERO: return "W = 0 in encounter";
    case CMCG_E_WPZERO: return "wp = 0 in relaxation";
    case CMCG_E_BRENT: return "apsis root finder failed";
    case CMCG_E_TIES: return "run of equal radii too long";
    case CMCG_E_HIP: return "HIP runtime error";
    case CMCG_E_ARG: return "invalid argument";
    case CMCG_E_CAPACITY: return "capacity exceeded";
    case CMCG_E_NOMEM: return "device allocation failed";
    case CMCG_E_STATE: return "call out of order";
    case CMCG_E_STRIP: return "tidal stripping did not end within the captured passes";
    default: return "unknown error";
    }
}

// ---- staging of cmcg_calc_pericenter_and_apocenter_noalloc ----
// One file-scope device array, `scratch`, that each GPU holds until the process
// ends: E and J in, the seven outputs back, for at most CMCG_APSIDES_NOALLOC_MAX
// stars.  Its host-side name is not a device address; host code uses only the
// address hipGetSymbolAddress returns on each device (looked up once per device
// when a context is created there).  Calls that overlap on one GPU share it and race.
constexpr size_t NOALLOC_SLOT = (size_t)CMCG_APSIDES_NOALLOC_MAX * 8;   // one 8-byte array
constexpr size_t NOALLOC_BYTES = 8 * NOALLOC_SLOT + (size_t)CMCG_APSIDES_NOALLOC_MAX * 4;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 256;
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
int g_scratch_rc[NOALLOC_MAX_DEVICES];          // 0 = not looked up, 1 = ready, CMCG_E_HIP = failed

// With `device` current: look up this device's copy of `scratch` once.
void noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_rc[device] != 0) return;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = ptr;
    g_scratch_rc[device] = (e == hipSuccess && ptr) ? 1 : CMCG_E_HIP;
}
}  // namespace

int cmcg_create(cmcg_ctx **out, int device, int64_t n_cap, const cmcg_params *params)
{
    if (!out || !params || n_cap < 1 || params->n_star < 1 || params->avekernel < 1) return CMCG_E_ARG;
    if (n_cap > (int64_t)0x7fffffff) return CMCG_E_CAPACITY;
    *out = nullptr;
    static std::atomic<uint64_t> serials{0};
    cmcg_ctx *c = new cmcg_ctx();
    c->serial = ++serials;
    c->device = device;
    c->n_cap = n_cap;
    c->params = *params;
    if (hipSetDevice(device) != hipSuccess) { delete c; return CMCG_E_HIP; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return CMCG_E_HIP; }
    noalloc_setup(device);
    KParams &P = c->kp;
    P.n_star = params->n_star;
    P.inv_nstar = 1.0 / params->n_star;
    P.madhoc = 1.0 / ((double)params->n_star);
    P.cenma_m = params->cenma_m;
    P.orbit_r = params->orbit_r;
    P.theta_se_max = params->theta_se_max;
    P.gamma = params->gamma;
    P.avekernel = params->avekernel;
    P.defer_cap = 1;
    while (P.defer_cap < n_cap + 2) P.defer_cap <<= 1;      // every star could be deferred
    int e;
    if ((e = alloc_fields(c, c->f, true)) || (e = alloc_fields(c, c->alt, false))) { cmcg_destroy(c); return e; }
    c->alt.phi = c->f.phi; c->alt.sigma = c->f.sigma; c->alt.pot_rnew = c->f.pot_rnew;
    c->alt.ntry = c->f.ntry; c->alt.flags = c->f.flags;
    if ((e = dalloc(c, &c->s, 1))) { cmcg_destroy(c); return e; }
    Scratch &sc = c->sc;
    sc.max_tiles = (n_cap + 2047) / 2048 + 1;
    sc.esa_max = 1 << 16;
    sc.max_waves = (n_cap + 2) / 64 + 2;
    sc.max_blocks = (n_cap + 127) / 128 + 2;      // the smallest block of any launch shape: 128 threads
    if (sc.max_blocks < 1024) sc.max_blocks = 1024;   // the stripping passes' fixed grid
    sc.max_sort_tiles = (n_cap + SORT_TILE - 1) / SORT_TILE + 1;
    sc.merge_ends_cap = 2 * (n_cap / 256 + 4);       // two per merge block of set_velocities3 (>= 256 stars)
    sc.sort_aos_cap = n_cap + 4;
    // about two radial-index buckets per star: 2^10 .. 2^23 buckets (at N = 4e6 one per
    // star makes the orbit sampler 3% slower; four per star cost more to build than the
    // sampler gains)
    sc.ridx_bits = 10;
    while (sc.ridx_bits < 23 && ((int64_t)1 << sc.ridx_bits) < 2 * (n_cap + 2)) sc.ridx_bits++;
    if ((e = dalloc(c, &sc.ridx, (size_t)ridx_entries(sc.ridx_bits))) ||
        (e = dalloc(c, &sc.ridx_tb, (size_t)ridx_tiles(sc.ridx_bits) + 1))) { cmcg_destroy(c); return e; }
    if ((e = dalloc(c, &sc.tile_sums, (size_t)sc.max_tiles * 6)) || (e = dalloc(c, &sc.partials, (size_t)sc.max_blocks * 16)) ||
        (e = dalloc(c, &sc.keys_a, (size_t)n_cap + 4)) || (e = dalloc(c, &sc.sort_aos, (size_t)(n_cap + 4) * SORT_AOS_WORDS)) || (e = dalloc(c, &sc.sort_long, (size_t)(sc.sort_long_cap = n_cap / (SORT_RUN_SHORT + 1) + 2))) ||
        (e = dalloc(c, &sc.sort_nlong, 1)) ||
        (e = dalloc(c, &sc.mid, (size_t)(n_cap + 2) / MID_STEP + 4)) || (e = dalloc(c, &sc.nodes, (size_t)n_cap + 8)) || (e = dalloc(c, &sc.u_old, (size_t)n_cap + 8)) ||
        (e = dalloc(c, &sc.merge_ends, (size_t)sc.merge_ends_cap)) || (e = dalloc(c, &sc.sv3_dpre, (size_t)n_cap + 6)) || (e = dalloc(c, &sc.sv3_dpost, (size_t)n_cap + 6)) || (e = dalloc(c, &sc.dt_part, (size_t)64 * 16)) || (e = dalloc(c, &sc.coarse, (size_t)ORB_MAX_COARSE + 8)) || (e = dalloc(c, &sc.red2, (size_t)RED2_BLOCKS * 16)) ||
        (e = dalloc(c, (char **)&sc.orb_nib, sizeof(NibbleTables<ORB_NIB_POWERS>))) ||
        (e = dalloc(c, (char **)&sc.relax_nib, sizeof(NibbleTables<8>))) ||
        (e = dalloc(c, (char **)&sc.wb_nib, 2 * sizeof(NibbleTables<RNG_WB_POWERS>))) || (e = dalloc(c, &sc.keys_b, (size_t)n_cap + 4)) ||
        (e = dalloc(c, &sc.vals_a, (size_t)n_cap + 4)) || (e = dalloc(c, &sc.sort_mm, 2 * SORT_MM_PARTS)) ||
        (e = dalloc(c, &sc.block_hist, (size_t)sc.max_sort_tiles * 256)) || (e = dalloc(c, &sc.digit_totals, 8 * 256)) ||
        (e = dalloc(c, &sc.excess_c, (size_t)n_cap + 6)) || (e = dalloc(c, &sc.excess_list, (size_t)n_cap + 2)) ||
        (e = dalloc(c, &sc.sv3_wend, (size_t)n_cap + 2)) || (e = dalloc(c, &sc.sv3_merged, (size_t)n_cap + 2)) ||
        (e = dalloc(c, &sc.sv3_blk, (size_t)sc.max_blocks)) || (e = dalloc(c, &sc.sv3_fallback, 1)) ||
        (e = dalloc(c, &sc.strip_done, 1)) || (e = dalloc(c, &sc.chain_x, (size_t)n_cap + 8)) ||
        (e = dalloc(c, &sc.chain_g, (size_t)n_cap + 8)) || (e = dalloc(c, &sc.chain_mk, (size_t)n_cap + 8)) ||
        (e = dalloc(c, &sc.defer_list, (size_t)P.defer_cap)) || (e = dalloc(c, &sc.defer_kfirst, (size_t)P.defer_cap)) ||
        (e = dalloc(c, &sc.chain_scal, 4)) || (e = dalloc(c, &sc.chain_len, 4)) || (e = dalloc(c, &sc.chain_fbad, 8)) ||
        (e = dalloc(c, &sc.esa_idx, (size_t)sc.esa_max)) || (e = dalloc(c, &sc.esa_racc, (size_t)sc.esa_max)) ||
        (e = dalloc(c, &sc.esa_v, (size_t)sc.esa_max)) || (e = dalloc(c, &sc.esa_n, 4)) || (e = dalloc(c, &sc.esa_raw_idx, 2048)) || (e = dalloc(c, &sc.esa_raw_part, 3 * 2048)) || (e = dalloc(c, &sc.rng_base, 2)) || (e = dalloc(c, &sc.rng_wave, (size_t)sc.max_waves)) || (e = dalloc(c, &sc.rng_seeded, 1)) ||
        (e = dalloc(c, &c->T, 1))) {
        cmcg_destroy(c);
        return e;
    }
    RngTable *ht = new RngTable;
    build_table(*ht);
    c->seeded = host_seed(params->seed);
    // the orbit sampler's per-lane jumps by l 2^20 (l < 64): nibble tables of the
    // matrix powers 2^20 .. 2^25, t[b][c][q][v] = column XOR for nibble v at bits 4q..
    auto *nib = new NibbleTables<ORB_NIB_POWERS>;
    fill_nibble_tables(*ht, ORB_NIB_B0, *nib);
    // the relaxation's per-lane jumps by 3 l (l < 64): matrix powers 2^0 .. 2^7
    auto *rnib = new NibbleTables<8>;
    fill_nibble_tables(*ht, 0, *rnib);
    // the wave-base jumps (offsets 3 t 2^6 and t 2^26 within a block)
    auto *wnib = new NibbleTables<RNG_WB_POWERS>[2];
    fill_nibble_tables(*ht, 6, wnib[0]);
    fill_nibble_tables(*ht, 26, wnib[1]);
    bool ok = hipMemcpy(c->T, ht, sizeof(RngTable), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(sc.rng_seeded, &c->seeded, sizeof(Rng), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(sc.orb_nib, nib, sizeof(*nib), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(sc.relax_nib, rnib, sizeof(*rnib), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(sc.wb_nib, wnib, 2 * sizeof(NibbleTables<RNG_WB_POWERS>), hipMemcpyHostToDevice) == hipSuccess;
    delete[] wnib;
    delete rnib;
    delete nib;
    delete ht;
    if (!ok) { cmcg_destroy(c); return CMCG_E_HIP; }
    *out = c;
    return CMCG_OK;
}

int cmcg_destroy(cmcg_ctx *c)
{
    if (!c) return CMCG_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    ens_forget(c->serial);
    hipSetDevice(c->device);
    invalidate_graphs(c);
    for (auto &p : c->pending) { hipEventDestroy(p.second.first); hipEventDestroy(p.second.second); }
    for (auto e : c->event_pool) hipEventDestroy(e);
    for (void *p : c->allocs) hipFree(p);
    if (c->io) hipFree(c->io);
    if (c->flush_buf) hipFree(c->flush_buf);
    if (c->mpc_buf) hipFree(c->mpc_buf);
    if (c->cen_buf) hipFree(c->cen_buf);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return CMCG_OK;
}

int cmcg_sync(cmcg_ctx *c)
{
    if (!c) return CMCG_E_ARG;
    if (sync(c) != hipSuccess) return CMCG_E_HIP;
    collect_timers(c);
    return check_error(c);
}

#define H2D(dst, srcp, count)                                                                     \
    if (srcp) {                                                                                   \
        if (hipMemcpyAsync(dst, srcp, (count) * sizeof(*(srcp)), hipMemcpyHostToDevice, c->stream) != hipSuccess) \
            return CMCG_E_HIP;                                                                    \
    }
#define D2H(dstp, src, count)                                                                     \
    if (dstp) {                                                                                   \
        if (hipMemcpyAsync(dstp, src, (count) * sizeof(*(dstp)), hipMemcpyDeviceToHost, c->stream) != hipSuccess) \
            return CMCG_E_HIP;                                                                    \
    }

int cmcg_upload(cmcg_ctx *c, const cmcg_host_soa *h)
{
    if (!c || !h || h->n < 0) return CMCG_E_ARG;
    if (h->n > c->n_cap) return CMCG_E_CAPACITY;
    size_t n = (size_t)h->n + 2;
    DevFields &f = c->f;
    H2D(f.r, h->r, n); H2D(f.m, h->m, n); H2D(f.phi, h->phi, n);
    H2D(f.vr, h->vr, n); H2D(f.vt, h->vt, n); H2D(f.E, h->E, n); H2D(f.J, h->J, n);
    H2D(f.id, h->id, n);
    H2D(f.rnew, h->rnew, n); H2D(f.vrnew, h->vrnew, n); H2D(f.vtnew, h->vtnew, n);
    H2D(f.X, h->X, n); H2D(f.Y, h->Y, n); H2D(f.r_peri, h->r_peri, n); H2D(f.r_apo, h->r_apo, n);
    H2D(f.rOld, h->rOld, n); H2D(f.Uoldrold, h->Uoldrold, n); H2D(f.Uoldrnew, h->Uoldrnew, n);
    H2D(f.vrold, h->vrold, n); H2D(f.vtold, h->vtold, n); H2D(f.EI, h->EI, n);
    H2D(f.sigma, h->sigma, n); H2D(f.pot_rnew, h->pot_rnew, n);
    H2D(f.ntry, h->ntry, n); H2D(f.flags, h->flags, n);
    DevScalars hs;
    if (hipMemcpyAsync(&hs, c->s, sizeof(hs), hipMemcpyDeviceToHost, c->stream) != hipSuccess) return CMCG_E_HIP;
    if (sync(c) != hipSuccess) return CMCG_E_HIP;
    hs.n_max = h->n;
    hs.n_max_new = h->n;
    if (hipMemcpy(c->s, &hs, sizeof(hs), hipMemcpyHostToDevice) != hipSuccess) return CMCG_E_HIP;
    c->n_ub = h->n;
    return CMCG_OK;
}

int cmcg_download(cmcg_ctx *c, cmcg_host_soa *h)
{
    if (!c || !h) return CMCG_E_ARG;
    DevScalars hs;
    if (int e = read_scalars(c, hs)) return e;
    size_t n = (size_t)hs.n_max_new + 2;
    DevFields &f = c->f;
    D2H(h->r, f.r, n); D2H(h->m, f.m, n); D2H(h->phi, f.phi, n);
    D2H(h->vr, f.vr, n); D2H(h->vt, f.vt, n); D2H(h->E, f.E, n); D2H(h->J, f.J, n);
    D2H(h->id, f.id, n);
    D2H(h->rnew, f.rnew, n); D2H(h->vrnew, f.vrnew, n); D2H(h->vtnew, f.vtnew, n);
    D2H(h->X, f.X, n); D2H(h->Y, f.Y, n); D2H(h->r_peri, f.r_peri, n); D2H(h->r_apo, f.r_apo, n);
    D2H(h->rOld, f.rOld, n); D2H(h->Uoldrold, f.Uoldrold, n); D2H(h->Uoldrnew, f.Uoldrnew, n);
    D2H(h->vrold, f.vrold, n); D2H(h->vtold, f.vtold, n); D2H(h->EI, f.EI, n);
    D2H(h->sigma, f.sigma, n); D2H(h->pot_rnew, f.pot_rnew, n);
    D2H(h->ntry, f.ntry, n); D2H(h->flags, f.flags, n);
    if (sync(c) != hipSuccess) return CMCG_E_HIP;
    h->n = hs.n_max_new;
    return CMCG_OK;
}

int cmcg_get_scalars(cmcg_ctx *c, cmcg_scalars *o)
{
    if (!c || !o) return CMCG_E_ARG;
    DevScalars h;
    if (int e = read_scalars(c, h)) return e;
    o->n_max = h.n_max; o->n_max_new = h.n_max_new;
    o->Mtotal = h.Mtotal; o->Rtidal = h.Rtidal; o->max_r = h.max_r; o->Dt = h.Dt; o->Prev_Dt = h.Prev_Dt;
    o->TotalTime = h.TotalTime; o->TidalMassLoss = h.TidalMassLoss; o->TidalMassLoss_old = h.TidalMassLoss_old;
    o->OldTidalMassLoss = h.OldTidalMassLoss; o->Eescaped = h.Eescaped; o->Jescaped = h.Jescaped;
    o->Etidal = h.Etidal; o->Eoops = h.Eoops; o->Etot = h.Etot; o->Ekin = h.Ekin; o->Epot = h.Epot;
    o->step = h.step;
    return h.error;
}

int cmcg_set_scalars(cmcg_ctx *c, const cmcg_scalars *in)
{
    if (!c || !in) return CMCG_E_ARG;
    if (in->n_max > c->n_cap || in->n_max_new > c->n_cap) return CMCG_E_CAPACITY;
    DevScalars h;
    if (int e = read_scalars(c, h)) { if (e < 0 && e == CMCG_E_HIP) return e; }
    h.n_max = in->n_max; h.n_max_new = in->n_max_new;
    h.Mtotal = in->Mtotal; h.Rtidal = in->Rtidal; h.max_r = in->max_r; h.Dt = in->Dt; h.Prev_Dt = in->Prev_Dt;
    h.TotalTime = in->TotalTime; h.TidalMassLoss = in->TidalMassLoss; h.TidalMassLoss_old = in->TidalMassLoss_old;
    h.OldTidalMassLoss = in->OldTidalMassLoss; h.Eescaped = in->Eescaped; h.Jescaped = in->Jescaped;
    h.Etidal = in->Etidal; h.Eoops = in->Eoops; h.Etot = in->Etot; h.Ekin = in->Ekin; h.Epot = in->Epot;
    h.step = in->step;
    h.error = 0; h.error_index = 0;
    if (hipMemcpy(c->s, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) return CMCG_E_HIP;
    c->n_ub = in->n_max_new > in->n_max ? in->n_max_new : in->n_max;
    return CMCG_OK;
}

// ------------------------------------------------------------ checkpoints
// Layout of a checkpoint buffer (little-endian, this library's struct layouts):
//   CkptHeader | DevScalars | per-star field 0 .. CKPT_NFIELDS-1, each (slots) entries
// where slots = max(N_MAX, N_MAX_NEW) + 2 (sentinels included).  The step counter in
// DevScalars selects the RNG substreams of the next step, so a context loaded from a
// checkpoint continues exactly as the saved one would have.
namespace {
constexpr uint64_t CKPT_MAGIC = 0x31544b50474d43ull;    // "CMGPKT1"
constexpr int CKPT_NFIELDS = 25;
struct CkptHeader {
    uint64_t magic;
    uint32_t version, scalars_bytes;
    int64_t slots;
    int64_t bytes;
    cmcg_params params;
};

struct CkptField { void *dev; size_t elem; };

void ckpt_fields(const DevFields &f, CkptField *o)
{
    void *p[CKPT_NFIELDS] = {f.r, f.m, f.phi, f.vr, f.vt, f.E, f.J, f.id, f.rnew, f.vrnew, f.vtnew, f.X, f.Y,
                             f.r_peri, f.r_apo, f.rOld, f.Uoldrold, f.Uoldrnew, f.vrold, f.vtold, f.EI, f.sigma,
                             f.pot_rnew, f.ntry, f.flags};
    for (int k = 0; k < CKPT_NFIELDS; k++) o[k] = {p[k], k >= CKPT_NFIELDS - 2 ? 4u : 8u};
}

int64_t ckpt_bytes(int64_t slots)
{
    return (int64_t)sizeof(CkptHeader) + (int64_t)sizeof(DevScalars) + slots * (23 * 8 + 2 * 4);
}

int64_t ckpt_slots(const DevScalars &h) { return (h.n_max > h.n_max_new ? h.n_max : h.n_max_new) + 2; }
}  // namespace

int cmcg_checkpoint_size(cmcg_ctx *c, int64_t *bytes)
{
    if (!c || !bytes) return CMCG_E_ARG;
    DevScalars h;
    if (hipMemcpyAsync(&h, c->s, sizeof(h), hipMemcpyDeviceToHost, c->stream) != hipSuccess || sync(c) != hipSuccess)
        return CMCG_E_HIP;
    *bytes = ckpt_bytes(ckpt_slots(h));
    return CMCG_OK;
}

int cmcg_checkpoint_save(cmcg_ctx *c, void *buf, int64_t cap)
{
    if (!c || !buf) return CMCG_E_ARG;
    DevScalars h;
    if (hipMemcpyAsync(&h, c->s, sizeof(h), hipMemcpyDeviceToHost, c->stream) != hipSuccess || sync(c) != hipSuccess)
        return CMCG_E_HIP;
    const int64_t slots = ckpt_slots(h), bytes = ckpt_bytes(slots);
    if (cap < bytes) return CMCG_E_CAPACITY;
    CkptHeader hd{};
    hd.magic = CKPT_MAGIC;
    hd.version = 1;
    hd.scalars_bytes = (uint32_t)sizeof(DevScalars);
    hd.slots = slots;
    hd.bytes = bytes;
    hd.params = c->params;
    char *o = (char *)buf;
    memcpy(o, &hd, sizeof(hd));
    memcpy(o + sizeof(hd), &h, sizeof(h));
    o += sizeof(hd) + sizeof(h);
    CkptField fl[CKPT_NFIELDS];
    ckpt_fields(c->f, fl);
    for (int k = 0; k < CKPT_NFIELDS; k++) {
        if (hipMemcpyAsync(o, fl[k].dev, slots * fl[k].elem, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            return CMCG_E_HIP;
        o += slots * fl[k].elem;
    }
    return sync(c) == hipSuccess ? CMCG_OK : CMCG_E_HIP;
}

int cmcg_checkpoint_info(const void *buf, int64_t size, cmcg_params *params, int64_t *n_max)
{
    if (!buf || size < (int64_t)(sizeof(CkptHeader) + sizeof(DevScalars))) return CMCG_E_ARG;
    CkptHeader hd;
    DevScalars h;
    memcpy(&hd, buf, sizeof(hd));
    memcpy(&h, (const char *)buf + sizeof(hd), sizeof(h));
    if (hd.magic != CKPT_MAGIC || hd.version != 1 || hd.scalars_bytes != sizeof(DevScalars) || hd.slots < 2 ||
        hd.bytes != ckpt_bytes(hd.slots) || size < hd.bytes || hd.slots != ckpt_slots(h))
        return CMCG_E_ARG;
    if (params) *params = hd.params;
    if (n_max) *n_max = hd.slots - 2;
    return CMCG_OK;
}

int cmcg_checkpoint_load(cmcg_ctx *c, const void *buf, int64_t size)
{
    if (!c) return CMCG_E_ARG;
    if (int e = cmcg_checkpoint_info(buf, size, nullptr, nullptr)) return e;
    CkptHeader hd;
    DevScalars h;
    memcpy(&hd, buf, sizeof(hd));
    memcpy(&h, (const char *)buf + sizeof(hd), sizeof(h));
    const cmcg_params &a = hd.params, &b = c->params;    // flags do not change results
    if (a.n_star != b.n_star || a.avekernel != b.avekernel || a.theta_se_max != b.theta_se_max ||
        a.gamma != b.gamma || a.orbit_r != b.orbit_r || a.cenma_m != b.cenma_m || a.seed != b.seed ||
        a.tidal_strip != b.tidal_strip)
        return CMCG_E_ARG;
    if (hd.slots - 2 > c->n_cap) return CMCG_E_CAPACITY;
    h.error = 0;
    h.error_index = 0;
    const char *o = (const char *)buf + sizeof(hd) + sizeof(h);
    CkptField fl[CKPT_NFIELDS];
    ckpt_fields(c->f, fl);
    for (int k = 0; k < CKPT_NFIELDS; k++) {
        if (hipMemcpyAsync(fl[k].dev, o, hd.slots * fl[k].elem, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return CMCG_E_HIP;
        o += hd.slots * fl[k].elem;
    }
    if (hipMemcpyAsync(c->s, &h, sizeof(h), hipMemcpyHostToDevice, c->stream) != hipSuccess || sync(c) != hipSuccess)
        return CMCG_E_HIP;
    c->n_ub = hd.slots - 2;
    return CMCG_OK;
}

int cmcg_sort_by_r(cmcg_ctx *c, int64_t *n_max)
{
    if (!c) return CMCG_E_ARG;
    {
        PhaseTimer t(c, "sort");
        invalidate_graphs(c);
        if (int e = run_sort(c, false)) return e;
    }
    DevScalars h;
    if (int e = read_scalars(c, h)) return e;
    c->n_ub = h.n_max;
    if (n_max) *n_max = h.n_max;
    return h.error;
}

int cmcg_potential(cmcg_ctx *c, int64_t *n_max, double *mtotal, double *rtidal)
{
    if (!c) return CMCG_E_ARG;
    {
        PhaseTimer t(c, "potential");
        LAUNCH(launch_potential(c->f, c->s, c->kp, c->n_ub, c->sc, c->stream));
    }
    DevScalars h;
    if (int e = read_scalars(c, h)) return e;
    c->n_ub = h.n_max;
    if (n_max) *n_max = h.n_max;
    if (mtotal) *mtotal = h.Mtotal;
    if (rtidal) *rtidal = h.Rtidal;
    return h.error;
}

int cmcg_potential_at(cmcg_ctx *c, int64_t count, const double *r, double *out)
{
    if (!c || count < 0 || (count > 0 && (!r || !out))) return CMCG_E_ARG;
    if (count == 0) return CMCG_OK;
    double *dx = nullptr, *dy = nullptr;
    if (hipMalloc(&dx, count * sizeof(double)) != hipSuccess) return CMCG_E_NOMEM;
    if (hipMalloc(&dy, count * sizeof(double)) != hipSuccess) { hipFree(dx); return CMCG_E_NOMEM; }
    int rc = CMCG_OK;
    if (hipMemcpyAsync(dx, r, count * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        launch_potential_at(c->f, c->s, c->kp, dx, dy, count, c->stream) != hipSuccess ||
        hipMemcpyAsync(out, dy, count * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        sync(c) != hipSuccess)
        rc = CMCG_E_HIP;
    hipFree(dx);
    hipFree(dy);
    if (rc) return rc;
    return check_error(c);
}

int cmcg_simul_relax(cmcg_ctx *c, double *dtrel)
{
    if (!c) return CMCG_E_ARG;
    LAUNCH(launch_simul_relax(c->f, c->s, c->kp, c->n_ub, c->stream));
    DevScalars h;
    if (int e = read_scalars(c, h)) return e;
    if (dtrel) *dtrel = h.dtrel;
    return h.error;
}

int cmcg_calc_timestep(cmcg_ctx *c, double *dt)
{
    if (!c) return CMCG_E_ARG;
    LAUNCH(launch_simul_relax(c->f, c->s, c->kp, c->n_ub, c->stream));
    LAUNCH(launch_timestep_finalize(c->s, c->stream));
    DevScalars h;
    if (int e = read_scalars(c, h)) return e;
    if (dt) *dt = h.Dt;
    return h.error;
}

int cmcg_sigma_r(cmcg_ctx *c)
{
    if (!c) return CMCG_E_ARG;
    LAUNCH(launch_sigma(c->f, c->s, c->kp, c->n_ub, c->stream));
    return cmcg_sync(c);
}

int cmcg_relax(cmcg_ctx *c, double dt, uint64_t step)
{
    if (!c) return CMCG_E_ARG;
    {
        PhaseTimer t(c, "relax");
        kl<set_dt_step_kernel>(dim3(1), dim3(1), 0, c->stream, c->s, dt, (int64_t)step);
        LAUNCH(launch_rng_bases(c->T, c->s, c->sc, c->stream));
        LAUNCH(launch_sigma(c->f, c->s, c->kp, c->n_ub, c->stream));
        LAUNCH(launch_relax(c->f, c->s, c->kp, c->T, c->sc, c->n_ub, c->stream));
    }
    return cmcg_sync(c);
}

static void fill_stats(const DevScalars &h, cmcg_orbit_stats *st)
{
    if (!st) return;
    st->n_circular = h.n_circular; st->n_removed_E = h.n_removed_E; st->n_removed_rt = h.n_removed_rt;
    st->n_destroyed = h.n_destroyed; st->n_brent = h.n_brent; st->n_linscan = h.n_linscan;
    st->tries_total = h.tries_total; st->tries_max = h.tries_max;
    st->n_bracket_hit = h.n_bracket_hit; st->n_bracket_fallback = h.n_bracket_fallback;
    st->max_r = h.max_r; st->tidal_mass_loss = h.TidalMassLoss; st->E_escaped = h.Eescaped; st->J_escaped = h.Jescaped;
    st->error = h.error; st->error_index = h.error_index;
}

int cmcg_orbits(cmcg_ctx *c, uint64_t step, cmcg_orbit_stats *stats)
{
    if (!c) return CMCG_E_ARG;
    {
        PhaseTimer t(c, "orbits");
        kl<set_dt_step_kernel>(dim3(1), dim3(1), 0, c->stream, c->s, __builtin_nan(""), (int64_t)step);
        LAUNCH(launch_rng_bases(c->T, c->s, c->sc, c->stream));
        hipEvent_t ka = nullptr, kb = nullptr;
        kernel_events(c, "orbits_kernel", ka, kb);
        LAUNCH(launch_orbits(c->f, c->s, c->kp, c->T, c->sc, c->n_ub, c->stream, c->learned ? &c->mlp : nullptr,
                             c->margin, ka, kb, c->orb_variant, c->pred_fixed));
        if (c->harvest) LAUNCH(launch_mlp_harvest(c->f, c->s, c->kp, c->mlp, c->n_ub, c->stream));
    }
    DevScalars h;
    if (int e = read_scalars(c, h)) return e;
    fill_stats(h, stats);
    return h.error;
}

int cmcg_calc_pericenter_and_apocenter(cmcg_ctx *c, int64_t first, int64_t count, const double *E, const double *J,
                                       double *rp, double *ra, double *dQdrp, double *dQdra, int64_t *kmin,
                                       int64_t *kmax, int32_t *circular)
{
    if (!c || count < 0 || !E || !J) return CMCG_E_ARG;
    if (count == 0) return CMCG_OK;
    if (first < 1 || first + count - 1 > c->n_ub) return CMCG_E_ARG;
    double *dE, *dJ, *d4;
    int64_t *dk;
    int32_t *dc;
    if (hipMalloc(&dE, count * 8) || hipMalloc(&dJ, count * 8) || hipMalloc(&d4, count * 32) ||
        hipMalloc(&dk, count * 16) || hipMalloc(&dc, count * 4))
        return CMCG_E_NOMEM;
    std::vector<double> h4(4 * count);
    std::vector<int64_t> hk(2 * count);
    std::vector<int32_t> hc(count);
    int rc = CMCG_OK;
    if (hipMemcpyAsync(dE, E, count * 8, hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(dJ, J, count * 8, hipMemcpyHostToDevice, c->stream) ||
        launch_apsides(c->f, c->s, c->kp, first, count, dE, dJ, d4, dk, dc, c->stream) ||
        hipMemcpyAsync(h4.data(), d4, count * 32, hipMemcpyDeviceToHost, c->stream) ||
        hipMemcpyAsync(hk.data(), dk, count * 16, hipMemcpyDeviceToHost, c->stream) ||
        hipMemcpyAsync(hc.data(), dc, count * 4, hipMemcpyDeviceToHost, c->stream) || sync(c))
        rc = CMCG_E_HIP;
    hipFree(dE); hipFree(dJ); hipFree(d4); hipFree(dk); hipFree(dc);
    if (rc) return rc;
    for (int64_t i = 0; i < count; i++) {
        if (rp) rp[i] = h4[4 * i];
        if (ra) ra[i] = h4[4 * i + 1];
        if (dQdrp) dQdrp[i] = h4[4 * i + 2];
        if (dQdra) dQdra[i] = h4[4 * i + 3];
        if (kmin) kmin[i] = hk[2 * i];
        if (kmax) kmax[i] = hk[2 * i + 1];
        if (circular) circular[i] = hc[i];
    }
    return check_error(c);
}

int cmcg_calc_pericenter_and_apocenter_noalloc(cmcg_ctx *c, int64_t first, int64_t count, const double *E,
                                               const double *J, double *rp, double *ra, double *dQdrp,
                                               double *dQdra, int64_t *kmin, int64_t *kmax, int32_t *circular)
{
    if (!c || count < 0 || !E || !J) return CMCG_E_ARG;
    if (count == 0) return CMCG_OK;
    if (first < 1 || first + count - 1 > c->n_ub) return CMCG_E_ARG;
    if (count > CMCG_APSIDES_NOALLOC_MAX) return CMCG_E_CAPACITY;
    if (c->device < 0 || c->device >= NOALLOC_MAX_DEVICES || g_scratch_rc[c->device] != 1) return CMCG_E_HIP;
    char *base = (char *)g_scratch_ptr[c->device];
    double *dE = (double *)(base + 0 * NOALLOC_SLOT), *dJ = (double *)(base + 1 * NOALLOC_SLOT);
    ApsidesOut o;
    o.rp = (double *)(base + 2 * NOALLOC_SLOT);
    o.ra = (double *)(base + 3 * NOALLOC_SLOT);
    o.dQdrp = (double *)(base + 4 * NOALLOC_SLOT);
    o.dQdra = (double *)(base + 5 * NOALLOC_SLOT);
    o.kmin = (int64_t *)(base + 6 * NOALLOC_SLOT);
    o.kmax = (int64_t *)(base + 7 * NOALLOC_SLOT);
    o.circular = (int32_t *)(base + 8 * NOALLOC_SLOT);
    const size_t b8 = (size_t)count * 8;
    hipStream_t st = c->stream;
    bool bad = hipMemcpyAsync(dE, E, b8, hipMemcpyHostToDevice, st) ||
               hipMemcpyAsync(dJ, J, b8, hipMemcpyHostToDevice, st) ||
               launch_apsides_soa(c->f, c->s, c->kp, first, count, dE, dJ, o, st);
    if (!bad && rp) bad = hipMemcpyAsync(rp, o.rp, b8, hipMemcpyDeviceToHost, st);
    if (!bad && ra) bad = hipMemcpyAsync(ra, o.ra, b8, hipMemcpyDeviceToHost, st);
    if (!bad && dQdrp) bad = hipMemcpyAsync(dQdrp, o.dQdrp, b8, hipMemcpyDeviceToHost, st);
    if (!bad && dQdra) bad = hipMemcpyAsync(dQdra, o.dQdra, b8, hipMemcpyDeviceToHost, st);
    if (!bad && kmin) bad = hipMemcpyAsync(kmin, o.kmin, b8, hipMemcpyDeviceToHost, st);
    if (!bad && kmax) bad = hipMemcpyAsync(kmax, o.kmax, b8, hipMemcpyDeviceToHost, st);
    if (!bad && circular) bad = hipMemcpyAsync(circular, o.circular, (size_t)count * 4, hipMemcpyDeviceToHost, st);
    if (sync(c) || bad) return CMCG_E_HIP;
    return check_error(c);
}

int cmcg_find_zero_Q(cmcg_ctx *c, int64_t count, const int64_t *j, const int64_t *kmin, const int64_t *kmax,
                     const double *E, const double *J, int64_t *out)
{
    if (!c || count < 0 || (count > 0 && (!j || !kmin || !kmax || !E || !J || !out))) return CMCG_E_ARG;
    if (count == 0) return CMCG_OK;
    for (int64_t i = 0; i < count; i++)
        if (j[i] < 0 || j[i] > c->n_ub + 1 || kmin[i] < 0 || kmax[i] > c->n_ub + 1 || kmin[i] > kmax[i])
            return CMCG_E_ARG;
    int64_t *dj, *dmin, *dmax, *dout;
    double *dE, *dJ;
    if (hipMalloc(&dj, count * 8) || hipMalloc(&dmin, count * 8) || hipMalloc(&dmax, count * 8) ||
        hipMalloc(&dout, count * 8) || hipMalloc(&dE, count * 8) || hipMalloc(&dJ, count * 8))
        return CMCG_E_NOMEM;
    int rc = CMCG_OK;
    if (hipMemcpyAsync(dj, j, count * 8, hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(dmin, kmin, count * 8, hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(dmax, kmax, count * 8, hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(dE, E, count * 8, hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(dJ, J, count * 8, hipMemcpyHostToDevice, c->stream) ||
        launch_find_zero_Q(c->f, c->s, c->kp, count, dj, dmin, dmax, dE, dJ, dout, c->stream) ||
        hipMemcpyAsync(out, dout, count * 8, hipMemcpyDeviceToHost, c->stream) || sync(c))
        rc = CMCG_E_HIP;
    hipFree(dj); hipFree(dmin); hipFree(dmax); hipFree(dout); hipFree(dE); hipFree(dJ);
    if (rc) return rc;
    return check_error(c);
}

int cmcg_find_root_vr(cmcg_ctx *c, int64_t count, const int64_t *j, const int64_t *k, const double *E,
                      const double *J, double *out)
{
    if (!c || count < 0 || (count > 0 && (!j || !k || !E || !J || !out))) return CMCG_E_ARG;
    if (count == 0) return CMCG_OK;
    for (int64_t i = 0; i < count; i++)
        if (j[i] < 1 || j[i] > c->n_ub || k[i] < 0 || k[i] > c->n_ub) return CMCG_E_ARG;
    int64_t *dj, *dk;
    double *dE, *dJ, *dout;
    if (hipMalloc(&dj, count * 8) || hipMalloc(&dk, count * 8) || hipMalloc(&dE, count * 8) ||
        hipMalloc(&dJ, count * 8) || hipMalloc(&dout, count * 8))
        return CMCG_E_NOMEM;
    int rc = CMCG_OK;
    if (hipMemcpyAsync(dj, j, count * 8, hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(dk, k, count * 8, hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(dE, E, count * 8, hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(dJ, J, count * 8, hipMemcpyHostToDevice, c->stream) ||
        launch_find_root_vr(c->f, c->s, c->kp, count, dj, dk, dE, dJ, dout, c->stream) ||
        hipMemcpyAsync(out, dout, count * 8, hipMemcpyDeviceToHost, c->stream) || sync(c))
        rc = CMCG_E_HIP;
    hipFree(dj); hipFree(dk); hipFree(dE); hipFree(dJ); hipFree(dout);
    if (rc) return rc;
    return check_error(c);
}

int cmcg_energy_conservation1(cmcg_ctx *c)
{
    if (!c) return CMCG_E_ARG;
    LAUNCH(launch_energy_conservation1(c->f, c->s, c->kp, c->n_ub, c->stream));
    return cmcg_sync(c);
}

int cmcg_energy_conservation2(cmcg_ctx *c)
{
    if (!c) return CMCG_E_ARG;
    LAUNCH(launch_energy_conservation2(c->f, c->s, c->kp, c->n_ub, c->stream));
    return cmcg_sync(c);
}

int cmcg_tidally_strip_stars(cmcg_ctx *c)
{
    if (!c) return CMCG_E_ARG;
    LAUNCH(launch_strip(c->f, c->s, c->kp, c->sc, c->sc.strip_done, c->n_ub, 1, true, c->stream));
    for (int guard = 0; guard < 100000; guard++) {
        int32_t done = 0;
        if (sync(c) != hipSuccess) return CMCG_E_HIP;
        if (hipMemcpy(&done, c->sc.strip_done, 4, hipMemcpyDeviceToHost) != hipSuccess) return CMCG_E_HIP;
        if (done) break;
        LAUNCH(launch_strip(c->f, c->s, c->kp, c->sc, c->sc.strip_done, c->n_ub, 1, false, c->stream));
    }
    return cmcg_sync(c);
}

int cmcg_compute_intermediate_energy(cmcg_ctx *c)
{
    if (!c) return CMCG_E_ARG;
    LAUNCH(launch_intermediate_energy(c->f, c->s, c->kp, c->n_ub, c->stream));
    LAUNCH(launch_ie_transfer(c->f, c->s, c->n_ub, c->stream));
    return cmcg_sync(c);
}

int cmcg_set_velocities3(cmcg_ctx *c)
{
    if (!c) return CMCG_E_ARG;
    LAUNCH(launch_set_velocities3(c->f, c->s, c->kp, c->sc, c->n_ub, c->stream));
    return cmcg_sync(c);
}

int cmcg_compute_energy(cmcg_ctx *c)
{
    if (!c) return CMCG_E_ARG;
    LAUNCH(launch_compute_energy(c->f, c->s, c->kp, c->sc, c->n_ub, c->stream));
    return cmcg_sync(c);
}

int cmcg_init_cluster(cmcg_ctx *c)
{
    if (!c) return CMCG_E_ARG;
    LAUNCH(launch_potential(c->f, c->s, c->kp, c->n_ub, c->sc, c->stream));
    LAUNCH(launch_sigma(c->f, c->s, c->kp, c->n_ub, c->stream));
    LAUNCH(launch_compute_energy(c->f, c->s, c->kp, c->sc, c->n_ub, c->stream));
    DevScalars h;
    if (int e = read_scalars(c, h)) return e;
    c->n_ub = h.n_max;
    return h.error;
}

// Enqueues one Henon step (src/cmc/cmc.c:240-407 order).  With `graph` the tidal
// stripping loop is a few guarded grid passes and a one-block tail that finishes a
// longer loop (no host round trip), so the whole step can be captured; nub bounds
// the grids (the kernels read N_MAX on the device).
constexpr int STRIP_GRAPH_PASSES = 3;     // grid passes; a one-block tail kernel finishes rarer longer loops

static int enqueue_step(cmcg_ctx *c, int64_t nub, bool graph)
{
    {
        PhaseTimer t(c, "timestep");
        LAUNCH(launch_step_begin(c->s, c->stream));                 // also starts the time step's minimum
        // calc_sigma_r with simul_relax_new, then Dt and the RNG bases
        LAUNCH(launch_timestep_sigma(c->f, c->s, c->kp, c->T, c->sc, nub, c->stream));
    }
    {
        // relaxation with energy_conservation1 fused in
        PhaseTimer t(c, "relax");
        LAUNCH(launch_relax(c->f, c->s, c->kp, c->T, c->sc, nub, c->stream, true, c->relax_threads));
        LAUNCH(launch_after_dynamics(c->s, c->stream));
    }
    {
        PhaseTimer t(c, "orbits");
        hipEvent_t ka = nullptr, kb = nullptr;
        kernel_events(c, "orbits_kernel", ka, kb);
        LAUNCH(launch_orbits(c->f, c->s, c->kp, c->T, c->sc, nub, c->stream, c->learned ? &c->mlp : nullptr,
                             c->margin, ka, kb, c->orb_variant, c->pred_fixed));
        if (c->harvest) LAUNCH(launch_mlp_harvest(c->f, c->s, c->kp, c->mlp, nub, c->stream));
    }
    {
        // energy_conservation2 with ComputeIntermediateEnergy's EI fused in
        PhaseTimer t(c, "glue_pre_sort");
        LAUNCH(launch_energy_conservation2(c->f, c->s, c->kp, nub, c->stream, true, c->sc.sv3_dpre));
        if (c->params.tidal_strip)
            LAUNCH(launch_strip(c->f, c->s, c->kp, c->sc, c->sc.strip_done, nub, graph ? STRIP_GRAPH_PASSES : 2, true,
                                c->stream));
    }
    if (c->params.tidal_strip) {
        if (graph) {
            LAUNCH(launch_strip_tail(c->f, c->s, c->kp, c->sc, c->stream));
        } else {
            // the stripping loop normally ends within two passes; otherwise continue it here
            for (int guard = 0; guard < 100000; guard++) {
                int32_t done = 0;
                if (sync(c) != hipSuccess) return CMCG_E_HIP;
                if (hipMemcpy(&done, c->sc.strip_done, 4, hipMemcpyDeviceToHost) != hipSuccess) return CMCG_E_HIP;
                if (done) break;
                LAUNCH(launch_strip(c->f, c->s, c->kp, c->sc, c->sc.strip_done, nub, 1, false, c->stream));
            }
        }
    }
    {
        PhaseTimer t(c, "sort");
        SortSpec sp;
        sort_spec(c, true, sp);
        hipError_t e = launch_sort_fields(c->f.rnew, c->alt.r, c->f.id, sp.nf, sp.src, sp.dst, c->s, c->sc, nub,
                                          c->stream);
        if (e != hipSuccess) { cmcg_report_hip(e, "launch_sort_fields", __FILE__, __LINE__); return CMCG_E_HIP; }
        swap_sorted(c, sp);
    }
    {
        PhaseTimer t(c, "potential");
        LAUNCH(launch_potential(c->f, c->s, c->kp, nub, c->sc, c->stream));
    }
    {
        PhaseTimer t(c, "glue_post_sort");
        // the sort above read the old sorted radii from what is now the alternate r
        LAUNCH(launch_set_velocities3(c->f, c->s, c->kp, c->sc, nub, c->stream, c->alt.r, c->sc.sv3_dpost));
        LAUNCH(launch_compute_energy(c->f, c->s, c->kp, c->sc, nub, c->stream));
        kl<step_end_kernel>(dim3(1), dim3(1), 0, c->stream, c->s);
    }
    c->parity ^= 1;
    return CMCG_OK;
}

static int finish_from(cmcg_ctx *c, const DevScalars &h, cmcg_orbit_stats *stats)
{
    c->n_ub = h.n_max;
    fill_stats(h, stats);
    return h.error;
}

static int finish_step(cmcg_ctx *c, cmcg_orbit_stats *stats)
{
    DevScalars h;
    if (int e = read_scalars(c, h)) return e;
    return finish_from(c, h, stats);
}

int cmcg_step(cmcg_ctx *c, cmcg_orbit_stats *stats)
{
    if (!c) return CMCG_E_ARG;
    if (int e = enqueue_step(c, c->n_ub, false)) return e;
    return finish_step(c, stats);
}

// Enqueue one step as a graph launch (captured on first use in each pointer state).
static int launch_step_graph(cmcg_ctx *c)
{
    const int p = c->parity;
    if (!c->graph[p]) {
        const bool timing = c->timing;
        c->timing = false;
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            c->timing = timing;
            return CMCG_E_HIP;
        }
        int e = enqueue_step(c, c->n_cap, true);       // swaps the host pointers, flips parity
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(c->stream, &g);
        c->timing = timing;
        if (e || ce != hipSuccess) {
            if (g) hipGraphDestroy(g);
            invalidate_graphs(c);
            return e ? e : CMCG_E_HIP;
        }
        hipError_t ie = hipGraphInstantiate(&c->graph[p], g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ie != hipSuccess) { c->graph[p] = nullptr; invalidate_graphs(c); return CMCG_E_HIP; }
    } else {
        SortSpec sp;
        sort_spec(c, true, sp);
        swap_sorted(c, sp);                             // the pointer state the graph leaves
        c->parity ^= 1;
    }
    if (hipGraphLaunch(c->graph[p], c->stream) != hipSuccess) return CMCG_E_HIP;
    return CMCG_OK;
}

int cmcg_step_graph(cmcg_ctx *c, cmcg_orbit_stats *stats)
{
    if (!c) return CMCG_E_ARG;
    if (int e = launch_step_graph(c)) return e;
    return finish_step(c, stats);
}

int cmcg_ensemble_step_concurrent(cmcg_ctx **ctxs, int count, cmcg_orbit_stats *stats, int32_t *errors)
{
    if (!ctxs || count < 1) return CMCG_E_ARG;
    for (int i = 0; i < count; i++)
        if (!ctxs[i] || ctxs[i]->device != ctxs[0]->device) return CMCG_E_ARG;
    int first_err = CMCG_OK;
    // every cluster's step graph goes to its own stream before any wait
    std::vector<int> launched(count, 0);
    for (int i = 0; i < count; i++) {
        int e = launch_step_graph(ctxs[i]);
        launched[i] = e == CMCG_OK;
        if (errors) errors[i] = e;
        if (e && !first_err) first_err = e;
    }
    for (int i = 0; i < count; i++) {
        if (!launched[i]) continue;
        int e = finish_step(ctxs[i], stats ? &stats[i] : nullptr);
        if (errors) errors[i] = e;
        if (e && !first_err) first_err = e;
    }
    return first_err;
}

} // extern "C"

// ------------------------------------------------------------------ merged ensemble
// One plan per pointer state of a set of clusters: every cluster's step is recorded
// (not launched), the recorded launches of all clusters are merged into one
// segmented launch per kernel, and the merged sequence is captured in one graph.
namespace {

// copies every cluster's scalars into one array (block c: cluster c)
struct ens_scalars_kernel {
    static constexpr int LB = 64, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const DevScalars *const *__restrict__ src,
                                               DevScalars *__restrict__ dst)
    {
        static_assert(sizeof(DevScalars) % 8 == 0, "word copy");
        const uint64_t *a = (const uint64_t *)src[kg_.bid];
        uint64_t *b = (uint64_t *)(dst + kg_.bid);
        for (unsigned w = threadIdx.x; w < sizeof(DevScalars) / 8; w += 64) b[w] = a[w];
    }
};

struct EnsGroup {
    SegLaunchFn fn;
    uint32_t blocks, gy;
    dim3 block;
    size_t lds;
    size_t ent_off, chunk_off, arg_off;   // into the plan's device buffer
};

struct EnsPlan {
    std::vector<uint64_t> key;   // (serial, gen, parity) per cluster
    int device = 0;
    void *dmem = nullptr;        // block tables and argument bytes of every merged launch
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> ev;  // one per cluster: that cluster's stream has reached the step
    DevScalars *host_scal = nullptr;   // pinned: every cluster's scalars after the step
    int64_t launches = 0, recorded = 0;
    ~EnsPlan()
    {
        hipSetDevice(device);
        if (exec) hipGraphExecDestroy(exec);
        if (host_scal) hipHostFree(host_scal);
        for (hipEvent_t e : ev)
            if (e) hipEventDestroy(e);
        if (dmem) hipFree(dmem);
    }
};

constexpr size_t ENS_MAX_PLANS = 4;     // both pointer states of two ensembles
std::mutex g_ens_mu;
std::vector<std::unique_ptr<EnsPlan>> g_ens_plans;   // most recently used last
int64_t g_ens_last_launches = 0, g_ens_last_recorded = 0;

std::vector<uint64_t> ens_key(cmcg_ctx **ctxs, int count)
{
    std::vector<uint64_t> k;
    k.reserve(3 * (size_t)count);
    for (int i = 0; i < count; i++) {
        k.push_back(ctxs[i]->serial);
        k.push_back(ctxs[i]->gen);
        k.push_back((uint64_t)ctxs[i]->parity);
    }
    return k;
}

// the host-side pointer swap and parity flip the step leaves behind
void apply_step_swap(cmcg_ctx *c)
{
    SortSpec sp;
    sort_spec(c, true, sp);
    swap_sorted(c, sp);
    c->parity ^= 1;
}

struct CtxPointers { DevFields f, alt; Scratch sc; int parity; };

// Merges the recorded launches and captures the merged sequence in the plan's graph.
int ens_assemble(EnsPlan *plan, const std::vector<SegRecorder> &rec, cmcg_ctx **ctxs, int count)
{
    // Alignment: each cluster's launches are matched to the longest cluster's sequence
    // (a longest common subsequence of kernel keys).  Launches are issued slot by slot:
    // slot 2j+1 holds the launches matched to position j of that sequence, slot 2j the
    // unmatched launches a cluster has before that match (another launch shape of the
    // orbit sampler, a kernel only some clusters run).  Within a slot, each round takes
    // the kernel that most clusters have next (ties: the cluster with most launches
    // left in the slot) and launches it for all of them.  Each cluster's launches keep
    // their order.
    auto same = [](const SegRec &a, const SegRec &b) {
        return a.fn == b.fn && a.gy == b.gy && a.block.x == b.block.x && a.block.y == b.block.y &&
               a.block.z == b.block.z && a.lds == b.lds && a.args.size() == b.args.size();
    };
    std::vector<std::vector<std::pair<int, size_t>>> members;   // (cluster, its record) per merged launch
    int refi = 0;
    for (int i = 0; i < count; i++) {
        plan->recorded += (int64_t)rec[i].recs.size();
        if (rec[i].recs.size() > rec[refi].recs.size()) refi = i;
    }
    const std::vector<SegRec> &R = rec[refi].recs;
    const size_t L = R.size();
    std::vector<std::vector<std::vector<size_t>>> slot(count, std::vector<std::vector<size_t>>(2 * L + 1));
    std::vector<uint32_t> dp;
    for (int i = 0; i < count; i++) {
        const std::vector<SegRec> &A = rec[i].recs;
        const size_t n = A.size(), W = L + 1;
        dp.assign((n + 1) * W, 0);                       // dp[a][r]: LCS of A[a..] and R[r..]
        for (size_t x = n; x-- > 0;)
            for (size_t y = L; y-- > 0;)
                dp[x * W + y] = same(A[x], R[y]) ? dp[(x + 1) * W + y + 1] + 1
                                                 : std::max(dp[(x + 1) * W + y], dp[x * W + y + 1]);
        size_t x = 0, y = 0;
        while (x < n) {
            if (y < L && same(A[x], R[y]) && dp[x * W + y] == dp[(x + 1) * W + y + 1] + 1) {
                slot[i][2 * y + 1].push_back(x);
                x++, y++;
            } else if (y < L && dp[x * W + y + 1] >= dp[(x + 1) * W + y]) {
                y++;
            } else {
                slot[i][2 * y].push_back(x);
                x++;
            }
        }
    }
    for (size_t sl = 0; sl < 2 * L + 1; sl++) {
        std::vector<size_t> cur(count, 0);
        for (;;) {
            int best = -1, best_n = 0;
            size_t best_left = 0;
            for (int i = 0; i < count; i++) {
                const std::vector<size_t> &li = slot[i][sl];
                if (cur[i] >= li.size()) continue;
                const SegRec &r = rec[i].recs[li[cur[i]]];
                int nm = 0;
                for (int j = 0; j < count; j++) {
                    const std::vector<size_t> &lj = slot[j][sl];
                    if (cur[j] < lj.size() && same(rec[j].recs[lj[cur[j]]], r)) nm++;
                }
                const size_t left = li.size() - cur[i];
                if (nm > best_n || (nm == best_n && left > best_left)) { best = i; best_n = nm; best_left = left; }
            }
            if (best < 0) break;
            const SegRec &r0 = rec[best].recs[slot[best][sl][cur[best]]];
            std::vector<std::pair<int, size_t>> m;
            for (int j = 0; j < count; j++) {
                const std::vector<size_t> &lj = slot[j][sl];
                if (cur[j] < lj.size() && same(rec[j].recs[lj[cur[j]]], r0)) m.emplace_back(j, lj[cur[j]]);
            }
            for (auto &jm : m) cur[jm.first]++;
            members.push_back(std::move(m));
        }
    }
    // device buffer: per merged launch the members' grids (each starting at a multiple
    // of 8 blocks), the owner of every 8 blocks and the members' argument bytes, each
    // 256-byte aligned
    std::vector<unsigned char> host;
    auto reserve = [&](size_t bytes) {
        const size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + bytes);
        return off;
    };
    std::vector<EnsGroup> groups;
    for (const auto &m : members) {
        const SegRec &r0 = rec[m[0].first].recs[m[0].second];
        const size_t asz = r0.args.size();
        EnsGroup g{r0.fn, 0, r0.gy, r0.block, r0.lds, 0, 0, 0};
        // cluster k's grid starts at a multiple of 8 blocks
        std::vector<uint32_t> first(m.size());
        uint64_t end = 0;
        for (size_t k = 0; k < m.size(); k++) {
            first[k] = (uint32_t)((end + 7) & ~(uint64_t)7);
            end = first[k] + (uint64_t)rec[m[k].first].recs[m[k].second].gx;
            if (end > 0x7fffffffull) return CMCG_E_ARG;
        }
        const size_t nchunk = (size_t)((end + 7) >> 3);
        g.ent_off = reserve(8 * m.size());
        g.chunk_off = reserve(2 * nchunk);
        g.arg_off = reserve(asz * m.size());
        uint16_t *chunk = (uint16_t *)(host.data() + g.chunk_off);
        for (size_t k = 0; k < m.size(); k++) {
            const SegRec &r = rec[m[k].first].recs[m[k].second];
            const uint32_t e[2] = {first[k], r.gx};
            memcpy(host.data() + g.ent_off + 8 * k, e, 8);
            memcpy(host.data() + g.arg_off + asz * k, r.args.data(), asz);
            for (size_t q = first[k] >> 3; q < ((size_t)first[k] + r.gx + 7) >> 3; q++) chunk[q] = (uint16_t)k;
        }
        g.blocks = (uint32_t)end;
        groups.push_back(g);
    }
    const size_t ptr_off = reserve(sizeof(DevScalars *) * (size_t)count);
    for (int i = 0; i < count; i++) memcpy(host.data() + ptr_off + sizeof(DevScalars *) * i, &ctxs[i]->s, sizeof(DevScalars *));
    const size_t scal_off = reserve(sizeof(DevScalars) * (size_t)count);
    plan->launches = (int64_t)groups.size();
    if (hipHostMalloc((void **)&plan->host_scal, sizeof(DevScalars) * (size_t)count, hipHostMallocDefault) != hipSuccess) {
        plan->host_scal = nullptr;
        return CMCG_E_NOMEM;
    }
    hipStream_t st = ctxs[0]->stream;
    if (hipMalloc(&plan->dmem, host.size() ? host.size() : 256) != hipSuccess) return CMCG_E_NOMEM;
    if (hipMemcpy(plan->dmem, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) return CMCG_E_HIP;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return CMCG_E_HIP;
    for (const EnsGroup &g : groups)
        g.fn(g.blocks, g.gy, g.block, g.lds, st, (const uint2 *)((char *)plan->dmem + g.ent_off),
             (const uint16_t *)((char *)plan->dmem + g.chunk_off), (char *)plan->dmem + g.arg_off);
    // every cluster's scalars to the host in one copy
    DevScalars *scal = (DevScalars *)((char *)plan->dmem + scal_off);
    kl<ens_scalars_kernel>(dim3((unsigned)count), dim3(64), 0, st,
                           (const DevScalars *const *)((char *)plan->dmem + ptr_off), scal);
    hipMemcpyAsync(plan->host_scal, scal, sizeof(DevScalars) * (size_t)count, hipMemcpyDeviceToHost, st);
    hipGraph_t graph = nullptr;
    hipError_t ce = hipStreamEndCapture(st, &graph);
    if (ce != hipSuccess || !graph) {
        if (graph) hipGraphDestroy(graph);
        return CMCG_E_HIP;
    }
    hipError_t ie = hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ie != hipSuccess) { plan->exec = nullptr; return CMCG_E_HIP; }
    plan->ev.assign(count, nullptr);
    for (int i = 0; i < count; i++)
        if (hipEventCreateWithFlags(&plan->ev[i], hipEventDisableTiming) != hipSuccess) return CMCG_E_HIP;
    return CMCG_OK;
}


// Records each cluster's step and builds its plan.  On success the clusters are left
// in the pointer state after the step (as enqueue_step leaves them); on failure in
// the state they had.
int ens_build(cmcg_ctx **ctxs, int count, std::unique_ptr<EnsPlan> &out)
{
    auto plan = std::make_unique<EnsPlan>();
    plan->key = ens_key(ctxs, count);
    plan->device = ctxs[0]->device;
    std::vector<SegRecorder> rec(count);
    std::vector<CtxPointers> saved(count);
    int err = CMCG_OK, done_ctx = 0;
    for (int i = 0; i < count && !err; i++, done_ctx++) {
        cmcg_ctx *c = ctxs[i];
        saved[i] = CtxPointers{c->f, c->alt, c->sc, c->parity};
        const bool timing = c->timing;
        c->timing = false;
        // the capture proves the step issued nothing but kernel launches (all recorded)
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            c->timing = timing;
            err = CMCG_E_HIP;
            break;
        }
        g_seg_rec = &rec[i];
        int e = enqueue_step(c, c->n_cap, true);
        g_seg_rec = nullptr;
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(c->stream, &g);
        c->timing = timing;
        size_t nodes = 0;
        if (g) {
            if (hipGraphGetNodes(g, nullptr, &nodes) != hipSuccess) ce = hipErrorUnknown;
            hipGraphDestroy(g);
        }
        if (e) err = e;
        else if (ce != hipSuccess || nodes != 0 || rec[i].bad) err = CMCG_E_HIP;
    }
    if (!err) err = ens_assemble(plan.get(), rec, ctxs, count);
    if (err) {
        for (int i = 0; i < done_ctx; i++) {
            ctxs[i]->f = saved[i].f;
            ctxs[i]->alt = saved[i].alt;
            ctxs[i]->sc = saved[i].sc;
            ctxs[i]->parity = saved[i].parity;
        }
        return err;
    }
    out = std::move(plan);
    return CMCG_OK;
}

// Launches a plan on the first cluster's stream, ordered after every cluster's
// earlier work (the caller waits for it before any later work).
int ens_launch(EnsPlan &p, cmcg_ctx **ctxs, int count)
{
    hipStream_t st = ctxs[0]->stream;
    for (int i = 1; i < count; i++) {
        const hipError_t q = hipStreamQuery(ctxs[i]->stream);
        if (q == hipSuccess) continue;                       // nothing pending there
        if (q != hipErrorNotReady) return CMCG_E_HIP;
        if (hipEventRecord(p.ev[i], ctxs[i]->stream) != hipSuccess) return CMCG_E_HIP;
        if (hipStreamWaitEvent(st, p.ev[i], 0) != hipSuccess) return CMCG_E_HIP;
    }
    if (hipGraphLaunch(p.exec, st) != hipSuccess) return CMCG_E_HIP;
    return CMCG_OK;
}

} // namespace

// drops the ensemble plans that step a cluster (before its context goes away)
static void ens_forget(uint64_t serial)
{
    std::lock_guard<std::mutex> lk(g_ens_mu);
    auto has = [serial](const std::unique_ptr<EnsPlan> &p) {
        for (size_t i = 0; i < p->key.size(); i += 3)
            if (p->key[i] == serial) return true;
        return false;
    };
    g_ens_plans.erase(std::remove_if(g_ens_plans.begin(), g_ens_plans.end(), has), g_ens_plans.end());
}

extern "C" {

int cmcg_ensemble_step(cmcg_ctx **ctxs, int count, cmcg_orbit_stats *stats, int32_t *errors)
{
    if (!ctxs || count < 1 || count > 65536) return CMCG_E_ARG;   // block owners are 16-bit
    for (int i = 0; i < count; i++) {
        if (!ctxs[i] || ctxs[i]->device != ctxs[0]->device) return CMCG_E_ARG;
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return CMCG_E_ARG;
    }
    hipSetDevice(ctxs[0]->device);
    {
        std::lock_guard<std::mutex> lk(g_ens_mu);
        const std::vector<uint64_t> key = ens_key(ctxs, count);
        auto it = std::find_if(g_ens_plans.begin(), g_ens_plans.end(),
                               [&](const std::unique_ptr<EnsPlan> &p) { return p->key == key; });
        std::unique_ptr<EnsPlan> plan;
        if (it != g_ens_plans.end()) {
            plan = std::move(*it);
            g_ens_plans.erase(it);
            for (int i = 0; i < count; i++) apply_step_swap(ctxs[i]);   // the state the plan's step leaves
        } else if (int e = ens_build(ctxs, count, plan)) {
            return e;
        }
        int e = ens_launch(*plan, ctxs, count);
        g_ens_last_launches = plan->launches;
        g_ens_last_recorded = plan->recorded;
        g_ens_plans.push_back(std::move(plan));
        if (g_ens_plans.size() > ENS_MAX_PLANS) g_ens_plans.erase(g_ens_plans.begin());
        if (e) return e;
        if (hipStreamSynchronize(ctxs[0]->stream) != hipSuccess) return CMCG_E_HIP;
        const DevScalars *h = g_ens_plans.back()->host_scal;
        int first_err = CMCG_OK;
        for (int i = 0; i < count; i++) {
            int e = finish_from(ctxs[i], h[i], stats ? &stats[i] : nullptr);
            if (errors) errors[i] = e;
            if (e && !first_err) first_err = e;
        }
        return first_err;
    }
}

int cmcg_ensemble_stats(int64_t *launches, int64_t *recorded)
{
    std::lock_guard<std::mutex> lk(g_ens_mu);
    if (launches) *launches = g_ens_last_launches;
    if (recorded) *recorded = g_ens_last_recorded;
    return CMCG_OK;
}

int cmcg_rng_draws(cmcg_ctx *c, int phase, uint64_t step, int64_t first, int64_t count, int ndraw, double *out)
{
    if (!c || (phase != 0 && phase != 1) || first < 0 || count < 0 || ndraw < 1 || (count > 0 && !out))
        return CMCG_E_ARG;
    if (count == 0) return CMCG_OK;
    double *d = nullptr;
    if (hipMalloc(&d, (size_t)count * ndraw * 8) != hipSuccess) return CMCG_E_NOMEM;
    int rc = CMCG_OK;
    if (launch_rng_bases_step(c->T, c->s, c->sc, (int64_t)step, c->stream) ||
        launch_rng_draws(c->T, c->sc.rng_base, phase, first, count, ndraw, d, c->stream) ||
        hipMemcpyAsync(out, d, (size_t)count * ndraw * 8, hipMemcpyDeviceToHost, c->stream) || sync(c))
        rc = CMCG_E_HIP;
    hipFree(d);
    return rc;
}

// ------------------------------------------------------------------ learned brackets
static uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static int io_reserve(cmcg_ctx *c, int64_t floats)
{
    if (floats <= c->io_cap) return CMCG_OK;
    if (c->io) { hipFree(c->io); c->io = nullptr; c->io_cap = 0; }
    if (hipMalloc((void **)&c->io, (size_t)floats * sizeof(float)) != hipSuccess) return CMCG_E_NOMEM;
    c->io_cap = floats;
    return CMCG_OK;
}

int cmcg_mlp_init(cmcg_ctx *c, uint64_t seed, int64_t max_rows)
{
    if (!c || max_rows < 1) return CMCG_E_ARG;
    MlpState &m = c->mlp;
    if (!c->mlp_ready) {
        int e;
        if ((e = dalloc(c, &m.theta, MLP_NPARAM)) || (e = dalloc(c, &m.grad, MLP_NPARAM)) ||
            (e = dalloc(c, &m.m, MLP_NPARAM)) || (e = dalloc(c, &m.v, MLP_NPARAM)) || (e = dalloc(c, &m.norm, 1)) ||
            (e = dalloc(c, &m.gpart, (size_t)MLP_MAX_BLOCKS * MLP_NPARAM)) || (e = dalloc(c, &m.lpart, MLP_MAX_BLOCKS)) ||
            (e = dalloc(c, &m.loss, 1)) || (e = dalloc(c, &m.rows, (size_t)max_rows * 8)) ||
            (e = dalloc(c, &m.counter, 1)) || (e = dalloc(c, &m.colsum, 256 * 16)) ||
            (e = dalloc(c, &c->sc.pred, (size_t)c->n_cap + 2)))
            return e;
        m.cap = max_rows;
        c->mlp_ready = true;
    } else if (max_rows > m.cap) {
        return CMCG_E_CAPACITY;
    }
    std::vector<float> th(MLP_NPARAM);
    uint64_t st = seed;
    auto uni = [&](double a) { return (float)(((splitmix64(st) >> 11) * 0x1.0p-53 * 2.0 - 1.0) * a); };
    const int fan[6][2] = {{MLP_NIN, MLP_NH * MLP_NIN}, {MLP_NIN, MLP_NH}, {MLP_NH, MLP_NH * MLP_NH},
                           {MLP_NH, MLP_NH}, {MLP_NH, MLP_NOUT * MLP_NH}, {MLP_NH, MLP_NOUT}};
    int p = 0;
    for (auto &fb : fan)
        for (int i = 0; i < fb[1]; i++) th[p++] = uni(1.0 / sqrt((double)fb[0]));
    MlpNorm nrm;
    for (int i = 0; i < MLP_NIN; i++) { nrm.xm[i] = 0.f; nrm.xs[i] = 1.f; }
    for (int i = 0; i < MLP_NOUT; i++) { nrm.ym[i] = 0.f; nrm.ys[i] = 1.f; }
    std::vector<float> zero(MLP_NPARAM, 0.f);
    unsigned long long z0 = 0;
    H2D(m.theta, th.data(), MLP_NPARAM);
    H2D(m.m, zero.data(), MLP_NPARAM);
    H2D(m.v, zero.data(), MLP_NPARAM);
    H2D(m.norm, &nrm, 1);
    H2D(m.counter, &z0, 1);
    c->adam_t = 0;
    return sync(c) == hipSuccess ? CMCG_OK : CMCG_E_HIP;
}

int cmcg_mlp_get_params(cmcg_ctx *c, float *theta, float *norm)
{
    if (!c || !c->mlp_ready) return CMCG_E_STATE;
    D2H(theta, c->mlp.theta, MLP_NPARAM);
    D2H(norm, (const float *)c->mlp.norm, CMCG_MLP_NNORM);
    return sync(c) == hipSuccess ? CMCG_OK : CMCG_E_HIP;
}

int cmcg_mlp_set_params(cmcg_ctx *c, const float *theta, const float *norm)
{
    if (!c || !c->mlp_ready) return CMCG_E_STATE;
    H2D(c->mlp.theta, theta, MLP_NPARAM);
    H2D((float *)c->mlp.norm, norm, CMCG_MLP_NNORM);
    return sync(c) == hipSuccess ? CMCG_OK : CMCG_E_HIP;
}

int cmcg_mlp_harvest(cmcg_ctx *c, int on)
{
    if (!c) return CMCG_E_ARG;
    if (on && !c->mlp_ready) return CMCG_E_STATE;
    if (c->harvest != (on != 0)) invalidate_graphs(c);
    c->harvest = on != 0;
    return CMCG_OK;
}

static int mlp_count(cmcg_ctx *c, int64_t *count)
{
    unsigned long long n = 0;
    if (hipMemcpyAsync(&n, c->mlp.counter, sizeof(n), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        sync(c) != hipSuccess)
        return CMCG_E_HIP;
    *count = (int64_t)n < c->mlp.cap ? (int64_t)n : c->mlp.cap;
    return CMCG_OK;
}

int cmcg_mlp_rows(cmcg_ctx *c, int64_t *count, float *rows)
{
    if (!c || !count) return CMCG_E_ARG;
    if (!c->mlp_ready) return CMCG_E_STATE;
    if (int e = mlp_count(c, count)) return e;
    D2H(rows, c->mlp.rows, *count * 8);
    return sync(c) == hipSuccess ? CMCG_OK : CMCG_E_HIP;
}

int cmcg_mlp_fit_norm(cmcg_ctx *c)
{
    if (!c || !c->mlp_ready) return CMCG_E_STATE;
    int64_t n;
    if (int e = mlp_count(c, &n)) return e;
    if (n < 2) return CMCG_E_STATE;
    LAUNCH(launch_mlp_fit_norm(c->mlp, n, c->stream));
    return sync(c) == hipSuccess ? CMCG_OK : CMCG_E_HIP;
}

int cmcg_mlp_loss_grad(cmcg_ctx *c, const float *rows, int64_t n, double *loss, float *grad)
{
    if (!c || !rows || n < 1) return CMCG_E_ARG;
    if (!c->mlp_ready) return CMCG_E_STATE;
    if (int e = io_reserve(c, 8 * n)) return e;
    H2D(c->io, rows, 8 * n);
    LAUNCH(launch_mlp_loss_grad(c->mlp, c->io, n, 0, n, 1, 0, c->stream));
    D2H(loss, c->mlp.loss, 1);
    D2H(grad, c->mlp.grad, MLP_NPARAM);
    return sync(c) == hipSuccess ? CMCG_OK : CMCG_E_HIP;
}

static uint64_t gcd64(uint64_t a, uint64_t b)
{
    while (b) { uint64_t t = a % b; a = b; b = t; }
    return a;
}

int cmcg_mlp_train(cmcg_ctx *c, int64_t iters, int64_t batch, float lr, uint64_t seed, double *loss)
{
    if (!c || iters < 0 || batch < 1 || !(lr > 0.f)) return CMCG_E_ARG;
    if (!c->mlp_ready) return CMCG_E_STATE;
    int64_t n;
    if (int e = mlp_count(c, &n)) return e;
    if (n < 1) return CMCG_E_STATE;
    uint64_t st = seed;
    uint64_t mult = splitmix64(st) % (uint64_t)n;
    while (mult == 0 || gcd64(mult, (uint64_t)n) != 1) mult = (mult + 1) % (uint64_t)n + (n == 1);
    uint64_t off = splitmix64(st) % (uint64_t)n;
    {
        PhaseTimer t(c, "mlp_train");
        for (int64_t it = 0; it < iters; it++) {
            LAUNCH(launch_mlp_loss_grad(c->mlp, c->mlp.rows, n, (it * batch) % n, batch, mult, off, c->stream));
            LAUNCH(launch_mlp_adam(c->mlp, ++c->adam_t, lr, c->stream));
        }
    }
    D2H(loss, c->mlp.loss, 1);
    return sync(c) == hipSuccess ? CMCG_OK : CMCG_E_HIP;
}

int cmcg_mlp_predict(cmcg_ctx *c, const float *x, int64_t n, float *y)
{
    if (!c || !x || !y || n < 1) return CMCG_E_ARG;
    if (!c->mlp_ready) return CMCG_E_STATE;
    if (int e = io_reserve(c, 8 * n)) return e;
    H2D(c->io, x, MLP_NIN * n);
    if (c->mlp_mfma) LAUNCH(launch_mlp_infer_rows(c->mlp, c->io, n, c->io + MLP_NIN * n, c->stream));
    else LAUNCH(launch_mlp_predict(c->mlp, c->io, n, c->io + MLP_NIN * n, c->stream));
    D2H(y, c->io + MLP_NIN * n, MLP_NOUT * n);
    return sync(c) == hipSuccess ? CMCG_OK : CMCG_E_HIP;
}

int cmcg_set_learned_brackets(cmcg_ctx *c, int on, double margin)
{
    if (!c || (on && !(margin > 0.0 && margin < 10.0))) return CMCG_E_ARG;
    if (on && !c->mlp_ready) return CMCG_E_STATE;
    if (c->learned != (on != 0) || (on && c->margin != (float)margin)) invalidate_graphs(c);
    c->learned = on != 0;
    if (on) c->margin = (float)margin;
    return CMCG_OK;
}

// ------------------------------------------------------------------ autotuning
int cmcg_autotune(cmcg_ctx *c, int reps, int32_t *variant, int32_t *threads, int32_t *coarse, double *ms_best,
                  double *ms_default)
{
    if (!c || reps < 1) return CMCG_E_ARG;
    DevScalars saved;
    if (sync(c) != hipSuccess || hipMemcpy(&saved, c->s, sizeof(saved), hipMemcpyDeviceToHost) != hipSuccess)
        return CMCG_E_HIP;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return CMCG_E_HIP;
    if (hipEventCreate(&b) != hipSuccess) { hipEventDestroy(a); return CMCG_E_HIP; }
    int best = 0;
    double best_ms = 1e300, def_ms = 0.0;
    int rc = CMCG_OK;
    c->tune_ms.assign(orbit_variant_count(), 0.0);
    for (int v = 0; v < orbit_variant_count() && rc == CMCG_OK; v++) {
        double total = 0.0;
        for (int it = 0; it <= reps; it++) {      // the first run warms up and is not timed
            hipEvent_t ea = it ? a : nullptr, eb = it ? b : nullptr;
            if (launch_rng_bases(c->T, c->s, c->sc, c->stream) != hipSuccess ||
                launch_orbits(c->f, c->s, c->kp, c->T, c->sc, c->n_ub, c->stream, c->learned ? &c->mlp : nullptr,
                              c->margin, ea, eb, v, c->pred_fixed) != hipSuccess ||
                sync(c) != hipSuccess) { rc = CMCG_E_HIP; break; }
            if (it) {
                float ms = 0.f;
                hipEventElapsedTime(&ms, a, b);
                total += ms;
            }
            // the phase is idempotent on the stars; put back the scalars it accumulates
            if (hipMemcpy(c->s, &saved, sizeof(saved), hipMemcpyHostToDevice) != hipSuccess) { rc = CMCG_E_HIP; break; }
        }
        double avg = total / reps;
        c->tune_ms[v] = avg;
        if (v == 0) def_ms = avg;
        if (avg < best_ms) { best_ms = avg; best = v; }
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc) return rc;
    if (best != c->orb_variant) invalidate_graphs(c);
    c->orb_variant = best;
    int th, cs;
    orbit_variant_shape(best, &th, &cs);
    if (variant) *variant = best;
    if (threads) *threads = th;
    if (coarse) *coarse = cs;
    if (ms_best) *ms_best = best_ms;
    if (ms_default) *ms_default = def_ms;
    return CMCG_OK;
}

int cmcg_comp_mass_percent(cmcg_ctx *c, const double *mass_pc, int32_t count, int32_t bh_include,
                           double initial_total_mass, double minimum_r, cmcg_mass_pc *out)
{
    if (!c || !out || count < 0 || count > 256 || (count > 0 && !mass_pc)) return CMCG_E_ARG;
    for (int i = 1; i < count; i++)
        if (!(mass_pc[i] >= mass_pc[i - 1])) return CMCG_E_ARG;
    if (!c->mpc_buf) {
        if (hipMalloc(&c->mpc_buf, mass_percent_out_bytes() + 256 * sizeof(double)) != hipSuccess) {
            c->mpc_buf = nullptr;
            return CMCG_E_NOMEM;
        }
    }
    char *dev = (char *)c->mpc_buf;
    double *dpc = (double *)(dev + mass_percent_out_bytes());
    if (count > 0 && hipMemcpyAsync(dpc, mass_pc, count * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return CMCG_E_HIP;
    if (hipMemsetAsync(dev, 0, mass_percent_out_bytes(), c->stream) != hipSuccess) return CMCG_E_HIP;
    LAUNCH(launch_mass_percent(c->f, c->s, c->kp, c->n_ub, c->sc, dpc, count, bh_include, initial_total_mass,
                               minimum_r, dev, c->stream));
    std::vector<char> host(mass_percent_out_bytes());
    if (hipMemcpyAsync(host.data(), dev, host.size(), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        sync(c) != hipSuccess)
        return CMCG_E_HIP;
    // layout: double v[8][256], int64 k[256], int64 reached
    const double *v = (const double *)host.data();
    const int64_t *k = (const int64_t *)(host.data() + 8 * 256 * sizeof(double));
    const int64_t reached = k[256];
    double *dst[7] = {out->mass_r, out->ave_mass_r, out->densities_r, out->ke_rad_r, out->ke_tan_r, out->v2_rad_r,
                      out->v2_tan_r};
    for (int64_t i = 0; i < reached && i < count; i++) {
        for (int q = 0; q < 7; q++)
            if (dst[q]) dst[q][i] = v[q * 256 + i];
        if (out->no_star_r) out->no_star_r[i] = k[i];
    }
    out->reached = (int32_t)reached;
    return check_error(c);
}

int cmcg_central_calculate(cmcg_ctx *c, cmcg_central *out)
{
    if (!c || !out) return CMCG_E_ARG;
    if (!c->cen_buf) {
        if (hipMalloc(&c->cen_buf, central_out_bytes()) != hipSuccess) {
            c->cen_buf = nullptr;
            return CMCG_E_NOMEM;
        }
    }
    LAUNCH(launch_central(c->f, c->s, c->kp, c->n_ub, c->sc, c->cen_buf, c->stream));
    // layout: int64 nave, int32 error, int32 pad, then 13 doubles in cmcg_central's order
    std::vector<char> host(central_out_bytes());
    if (hipMemcpyAsync(host.data(), c->cen_buf, host.size(), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        sync(c) != hipSuccess)
        return CMCG_E_HIP;
    int32_t err;
    std::memcpy(&out->nave, host.data(), sizeof(int64_t));
    std::memcpy(&err, host.data() + 8, sizeof(int32_t));
    std::memcpy(&out->rh, host.data() + 16, 13 * sizeof(double));
    if (int e = check_error(c)) return e;
    return err == 0 ? CMCG_OK : err == 1 ? CMCG_E_STATE : CMCG_E_NAN;
}

int cmcg_test_hook(cmcg_ctx *c, int32_t hook, int64_t value)
{
    if (!c) return CMCG_E_ARG;
    switch (hook) {
    case CMCG_HOOK_CHAIN_K_DROP:
        if (value < 0 || value > 92) return CMCG_E_ARG;
        c->kp.chain_k_drop = (int32_t)value;
        invalidate_graphs(c);
        return CMCG_OK;
    case CMCG_HOOK_ORBIT_VARIANT:
        if (value < 0 || value >= orbit_variant_count()) return CMCG_E_ARG;
        c->orb_variant = (int)value;
        invalidate_graphs(c);
        return CMCG_OK;
    case CMCG_HOOK_PERFECT_BRACKETS:
        if (value < 0 || value > 1) return CMCG_E_ARG;
        if (value && !c->mlp_ready) return CMCG_E_STATE;
        if (value) LAUNCH(launch_pred_from_apsides(c->f, c->s, c->sc, c->n_ub, c->stream));
        c->pred_fixed = value != 0;
        invalidate_graphs(c);
        return sync(c) == hipSuccess ? CMCG_OK : CMCG_E_HIP;
    case CMCG_HOOK_MLP_MFMA:
        if (value < 0 || value > 1) return CMCG_E_ARG;
        c->mlp_mfma = value != 0;
        return CMCG_OK;
    case CMCG_HOOK_RELAX_SHAPE: {
        bool ok = false;
        for (int v = 0; v < RELAX_NSHAPES; v++) ok = ok || RELAX_SHAPES[v] == value;
        if (!ok) return CMCG_E_ARG;
        if (c->relax_threads != (int)value) invalidate_graphs(c);
        c->relax_threads = (int)value;
        return CMCG_OK;
    }
    case CMCG_HOOK_COLD_PHASES:
        if (value < 0 || value > 65536) return CMCG_E_ARG;
        if (sync(c) != hipSuccess) return CMCG_E_HIP;
        if (c->flush_buf) { hipFree(c->flush_buf); c->flush_buf = nullptr; c->flush_words = 0; }
        if (value) {
            const size_t bytes = (size_t)value << 20;
            if (hipMalloc((void **)&c->flush_buf, bytes) != hipSuccess) { c->flush_buf = nullptr; return CMCG_E_NOMEM; }
            if (hipMemset(c->flush_buf, 0, bytes) != hipSuccess) return CMCG_E_HIP;
            c->flush_words = bytes / sizeof(uint4);
        }
        return CMCG_OK;
    default:
        return CMCG_E_ARG;
    }
}

int cmcg_autotune_relax(cmcg_ctx *c, int reps, int32_t *threads, double *ms_best, double *ms_default)
{
    if (!c || reps < 1) return CMCG_E_ARG;
    DevScalars saved;
    if (sync(c) != hipSuccess || hipMemcpy(&saved, c->s, sizeof(saved), hipMemcpyDeviceToHost) != hipSuccess)
        return CMCG_E_HIP;
    // the fields the fused relaxation pass rewrites, restored after every trial
    double *const fields[] = {c->f.vr, c->f.vt, c->f.E, c->f.J, c->f.Y, c->f.vrold, c->f.vtold, c->f.Uoldrold};
    constexpr int NF = (int)(sizeof(fields) / sizeof(fields[0]));
    const size_t bytes = (size_t)(c->n_cap + 2) * sizeof(double);
    double *keep = nullptr;
    if (hipMalloc(&keep, bytes * NF) != hipSuccess) return CMCG_E_NOMEM;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = CMCG_OK;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) rc = CMCG_E_HIP;
    for (int i = 0; i < NF && !rc; i++)
        if (hipMemcpyAsync(keep + (c->n_cap + 2) * i, fields[i], bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            rc = CMCG_E_HIP;
    int best = c->relax_threads;
    double best_ms = 1e300, def_ms = 0.0;
    for (int v = 0; v < RELAX_NSHAPES && !rc; v++) {
        double total = 0.0;
        for (int it = 0; it <= reps && !rc; it++) {      // the first run warms up and is not timed
            if (it && hipEventRecord(a, c->stream) != hipSuccess) { rc = CMCG_E_HIP; break; }
            if (launch_relax(c->f, c->s, c->kp, c->T, c->sc, c->n_ub, c->stream, true, RELAX_SHAPES[v]) != hipSuccess) {
                rc = CMCG_E_HIP;
                break;
            }
            if (it && hipEventRecord(b, c->stream) != hipSuccess) { rc = CMCG_E_HIP; break; }
            for (int i = 0; i < NF && !rc; i++)
                if (hipMemcpyAsync(fields[i], keep + (c->n_cap + 2) * i, bytes, hipMemcpyDeviceToDevice, c->stream) !=
                    hipSuccess)
                    rc = CMCG_E_HIP;
            if (rc || sync(c) != hipSuccess) { rc = CMCG_E_HIP; break; }
            if (it) {
                float ms = 0.f;
                hipEventElapsedTime(&ms, a, b);
                total += ms;
            }
            if (hipMemcpy(c->s, &saved, sizeof(saved), hipMemcpyHostToDevice) != hipSuccess) rc = CMCG_E_HIP;
        }
        const double avg = total / reps;
        if (v == 0) def_ms = avg;
        if (avg < best_ms) { best_ms = avg; best = RELAX_SHAPES[v]; }
    }
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    hipFree(keep);
    if (rc) return rc;
    if (best != c->relax_threads) invalidate_graphs(c);
    c->relax_threads = best;
    if (threads) *threads = best;
    if (ms_best) *ms_best = best_ms;
    if (ms_default) *ms_default = def_ms;
    return CMCG_OK;
}

int cmcg_orbit_variants(cmcg_ctx *c, int32_t max, int32_t *count, int32_t *threads, int32_t *coarse, int32_t *cap,
                        int32_t *min_blocks, double *ms_last)
{
    if (!c || !count || max < 0) return CMCG_E_ARG;
    const int nv = orbit_variant_count();
    *count = nv;
    for (int v = 0; v < nv && v < max; v++) {
        int th, cs, cp, mb;
        orbit_variant_shape(v, &th, &cs, &cp, &mb);
        if (threads) threads[v] = th;
        if (coarse) coarse[v] = cs;
        if (cap) cap[v] = cp;
        if (min_blocks) min_blocks[v] = mb;
        if (ms_last) ms_last[v] = v < (int)c->tune_ms.size() ? c->tune_ms[v] : 0.0;
    }
    return CMCG_OK;
}

int cmcg_timer_enable(cmcg_ctx *c, int on)
{
    if (!c) return CMCG_E_ARG;
    c->timing = on != 0;
    return CMCG_OK;
}

int cmcg_timer_get(cmcg_ctx *c, const char *name, double *ms_total, int64_t *calls)
{
    if (!c || !name) return CMCG_E_ARG;
    if (sync(c) != hipSuccess) return CMCG_E_HIP;
    collect_timers(c);
    auto it = c->timers.find(name);
    if (ms_total) *ms_total = it == c->timers.end() ? 0.0 : it->second.first;
    if (calls) *calls = it == c->timers.end() ? 0 : it->second.second;
    return CMCG_OK;
}

int cmcg_timer_reset(cmcg_ctx *c)
{
    if (!c) return CMCG_E_ARG;
    if (sync(c) != hipSuccess) return CMCG_E_HIP;
    collect_timers(c);
    c->timers.clear();
    return CMCG_OK;
}

} // extern "C"
