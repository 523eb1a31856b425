// a1: qsorts_new -> sample_sort at one rank (src/cmc/cmc_sort.c:375-629): order the
// stars by r ascending (compare_type, :245-250), ties by id (the reference's libc
// qsort leaves them unordered), drop the r == SF_INFINITY tail
// (remove_stripped_stars, :343-354), then refresh r and m (post_sort_comm,
// src/cmc/cmc_utils.c:2259-2288).
//
// Device algorithm: LSD radix sort of (64-bit order-preserving key of r, 32-bit
// slot) with 8-bit digits, 8 passes of three kernels (tile histogram, per-digit
// tile scan, stable scatter).  The scatter ranks each 4096-key tile with 64-lane
// ballot multisplit (stable in tile order), stages the sorted tile in LDS and
// writes it out digit-run by digit-run, so global writes are coalesced.  A tie
// pass orders runs of equal keys by id, then one gather moves the payload fields.
#include "kernels.hpp"

namespace {
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_MAX_FIELDS = 24;
// LSD passes over the top 40 bits of the key only; keys equal in those bits are then
// put in (key, id) order by the run pass (one thread per run up to RUN_SHORT keys,
// one block sorting in LDS up to RUN_LONG keys).
constexpr int RS_PASSES = 5;
constexpr int RS_SHIFT0 = 64 - 8 * RS_PASSES;
constexpr int RUN_SHORT = 64;
constexpr int RUN_LONG = 4096;
constexpr int AOS_W = SORT_AOS_WORDS;   // words per packed payload record (one 128-B line)
constexpr int AOS_LD = AOS_W + 1;        // odd LDS stride per record: no bank conflicts

__device__ __forceinline__ uint64_t sortable(double x)
{
    uint64_t b = (uint64_t)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double unsortable(uint64_t k)
{
    uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

struct GatherList {
    int nf;
    const uint64_t *src[RS_MAX_FIELDS];
    uint64_t *dst[RS_MAX_FIELDS];
};

__global__ void rs_zero_kernel(uint32_t *__restrict__ a, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// keys/vals from slots 1..n plus the eight global digit histograms; with `aos` the
// payload fields of each slot are also packed into one 96-byte record, so the
// final gather reads one record per star instead of one cache line per field.
__global__ __launch_bounds__(RS_THREADS) void rs_prep_kernel(const double *__restrict__ keysrc, DevScalars *s,
                                                             uint64_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                             uint32_t *__restrict__ digit_totals,
                                                             uint64_t *__restrict__ aos, GatherList g, uint64_t kinf)
{
    constexpr int P0 = 8 - RS_PASSES;      // first byte sorted on
    __shared__ uint32_t h[RS_PASSES][256]; // 40 KB of LDS in all: four blocks per CU
    __shared__ uint64_t recs[RS_THREADS * AOS_LD];     // one chunk of packed records
    const int t = threadIdx.x;
    for (int p = 0; p < RS_PASSES; p++) h[p][t] = 0;
    __syncthreads();
    const int64_t n = s->n_max_new;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int u = 0; u < RS_ITEMS; u++) {
        const int64_t i0 = base + u * RS_THREADS;
        const int64_t i = i0 + t;
        if (i < n) {
            uint64_t k = sortable(keysrc[i + 1]);
            if (k >= kinf) k = ~0ull;      // the SF_INFINITY tail: one key, no prefix shared with a live star
            keys[i] = k;
            vals[i] = (uint32_t)(i + 1);
            for (int p = P0; p < 8; p++) atomicAdd(&h[p - P0][(k >> (8 * p)) & 255], 1u);
        }
        if (aos && i0 < n) {        // block-uniform: pack the chunk through LDS, write it contiguously
            if (i < n)
                for (int f = 0; f < g.nf; f++) recs[AOS_LD * t + f] = g.src[f][i + 1];
            __syncthreads();
            const int64_t cnt = n - i0 < RS_THREADS ? n - i0 : RS_THREADS;
            uint64_t *dst = aos + (int64_t)AOS_W * i0;
            for (int w = t; w < cnt * AOS_W; w += RS_THREADS) dst[w] = recs[AOS_LD * (w / AOS_W) + w % AOS_W];
            __syncthreads();
        }
    }
    __syncthreads();
    for (int p = P0; p < 8; p++)
        if (h[p - P0][t]) atomicAdd(&digit_totals[p * 256 + t], h[p - P0][t]);
}

__global__ __launch_bounds__(RS_THREADS) void rs_hist_kernel(const uint64_t *__restrict__ keys, DevScalars *s,
                                                             int shift, uint32_t *__restrict__ block_hist,
                                                             int64_t ntiles)
{
    __shared__ uint32_t h[256];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const int64_t n = s->n_max_new;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int u = 0; u < RS_ITEMS; u++) {
        int64_t i = base + u * RS_THREADS + t;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1u);
    }
    __syncthreads();
    block_hist[(int64_t)t * ntiles + blockIdx.x] = h[t];
}

// One block per digit: exclusive scan of that digit's tile counts, plus the digit's
// global base (sum of the lower digits' totals for this pass).
__global__ __launch_bounds__(RS_THREADS) void rs_scan_kernel(uint32_t *__restrict__ block_hist, int64_t ntiles,
                                                             const uint32_t *__restrict__ totals)
{
    __shared__ uint32_t red[RS_THREADS];
    const int t = threadIdx.x;
    const int d = blockIdx.x;
    red[t] = t < d ? totals[t] : 0u;
    __syncthreads();
    for (int off = RS_THREADS / 2; off > 0; off >>= 1) {
        if (t < off) red[t] += red[t + off];
        __syncthreads();
    }
    const uint32_t dbase = red[0];
    __syncthreads();
    uint32_t *row = block_hist + (int64_t)d * ntiles;
    const int64_t chunk = (ntiles + RS_THREADS - 1) / RS_THREADS;
    const int64_t b0 = (int64_t)t * chunk;
    uint32_t sum = 0;
    for (int64_t b = b0; b < b0 + chunk && b < ntiles; b++) sum += row[b];
    red[t] = sum;
    __syncthreads();
    // inclusive Hillis-Steele scan of the per-thread sums
    for (int off = 1; off < RS_THREADS; off <<= 1) {
        uint32_t v = t >= off ? red[t - off] : 0u;
        __syncthreads();
        red[t] += v;
        __syncthreads();
    }
    uint32_t run = dbase + red[t] - sum;
    for (int64_t b = b0; b < b0 + chunk && b < ntiles; b++) {
        uint32_t c = row[b];
        row[b] = run;
        run += c;
    }
}

// Stable scatter of one 4096-key tile.  Wave w ranks the contiguous sub-tile
// [w*1024, (w+1)*1024) on its own (ballot matching of the 8 digit bits, a per-wave
// digit counter in LDS updated by the highest lane of each match group), so the
// item loop needs no block barrier; the (digit, wave) counts then give each wave's
// offset inside the tile's digit runs.  Index order (wave, item, lane) is the tile's
// index order, so equal digits keep their order.
__global__ __launch_bounds__(RS_THREADS) void rs_scatter_kernel(const uint64_t *__restrict__ keys_in,
                                                                const uint32_t *__restrict__ vals_in,
                                                                uint64_t *__restrict__ keys_out,
                                                                uint32_t *__restrict__ vals_out, DevScalars *s,
                                                                int shift, const uint32_t *__restrict__ tile_off,
                                                                int64_t ntiles)
{
    constexpr int NW = RS_THREADS / 64;
    __shared__ uint64_t skeys[RS_TILE];
    __shared__ uint32_t svals[RS_TILE];
    __shared__ uint32_t wcnt[NW][256];
    __shared__ uint32_t dstart[256];       // 53 KB in all: three blocks per CU
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t n = s->n_max_new;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    const int64_t wbase = base + (int64_t)w * (64 * RS_ITEMS);
    for (int d = lane; d < 256; d += 64) wcnt[w][d] = 0;
    uint64_t k[RS_ITEMS];
    uint32_t v[RS_ITEMS];
    uint32_t rank[RS_ITEMS];
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    for (int u = 0; u < RS_ITEMS; u++) {
        int64_t i = wbase + u * 64 + lane;
        bool valid = i < n;
        k[u] = valid ? keys_in[i] : 0ull;
        v[u] = valid ? vals_in[i] : 0u;
        uint32_t d = (uint32_t)((k[u] >> shift) & 255);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t before = valid ? wcnt[w][d] : 0u;
        rank[u] = before + (uint32_t)__popcll(peers & lt_mask);
        if (valid && (peers >> lane) == 1ull) wcnt[w][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // per digit (thread t = digit): the waves' offsets and the tile total
    uint32_t tot = 0;
    for (int w2 = 0; w2 < NW; w2++) {
        uint32_t c = wcnt[w2][t];
        wcnt[w2][t] = tot;
        tot += c;
    }
    // exclusive scan of the tile's digit counts
    dstart[t] = tot;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t x = t >= off ? dstart[t - off] : 0u;
        __syncthreads();
        dstart[t] += x;
        __syncthreads();
    }
    uint32_t incl = dstart[t];
    __syncthreads();
    dstart[t] = incl - tot;
    __syncthreads();
    for (int u = 0; u < RS_ITEMS; u++) {
        int64_t i = wbase + u * 64 + lane;
        if (i < n) {
            uint32_t d = (uint32_t)((k[u] >> shift) & 255);
            uint32_t p = dstart[d] + wcnt[w][d] + rank[u];
            skeys[p] = k[u];
            svals[p] = v[u];
        }
    }
    __syncthreads();
    const int64_t cnt_tile = n - base < RS_TILE ? n - base : RS_TILE;
    for (int u = 0; u < RS_ITEMS; u++) {
        int p = u * RS_THREADS + t;
        if (p < cnt_tile) {
            uint64_t key = skeys[p];
            uint32_t d = (uint32_t)((key >> shift) & 255);
            int64_t g = (int64_t)tile_off[(int64_t)d * ntiles + blockIdx.x] + (p - (int64_t)dstart[d]);
            keys_out[g] = key;
            vals_out[g] = svals[p];
        }
    }
}

// Runs of keys equal in the top 40 bits (live stars only: the tail keys are all ones)
// are put in (key, id) order.  Run starts sort runs of up to RUN_SHORT keys
// themselves (insertion sort) and list longer ones for rs_long_runs_kernel.
__device__ __forceinline__ bool key_id_less(uint64_t ka, int64_t ia, uint64_t kb, int64_t ib)
{
    return ka < kb || (ka == kb && ia < ib);
}

__global__ void rs_runs_kernel(uint64_t *__restrict__ keys, uint32_t *__restrict__ vals,
                               const int64_t *__restrict__ id, DevScalars *s, int64_t *__restrict__ long_runs,
                               int64_t *__restrict__ n_long)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = s->n_max_new;
    if (i + 1 >= n) return;
    const uint64_t k = keys[i];
    if (k == ~0ull) return;
    const uint64_t p = k >> RS_SHIFT0;
    if ((keys[i + 1] >> RS_SHIFT0) != p) return;
    if (i > 0 && (keys[i - 1] >> RS_SHIFT0) == p) return;
    int64_t e = i + 1;
    while (e + 1 < n && (keys[e + 1] >> RS_SHIFT0) == p && e - i < RUN_SHORT) e++;
    if (e + 1 < n && (keys[e + 1] >> RS_SHIFT0) == p) {
        unsigned long long slot = atomicAdd((unsigned long long *)n_long, 1ull);
        if ((int64_t)slot < SORT_LONG_RUNS) long_runs[slot] = i;
        else dev_set_error(s, CMCG_E_TIES, i + 1);
        return;
    }
    for (int64_t a = i + 1; a <= e; a++) {
        const uint64_t ka = keys[a];
        const uint32_t v = vals[a];
        int64_t idv = 0;
        bool have_id = false;            // ids are gathered only for equal keys
        int64_t b = a - 1;
        while (b >= i) {
            const uint64_t kb = keys[b];
            const uint32_t vb = vals[b];
            bool less;
            if (ka != kb) {
                less = ka < kb;
            } else {
                if (!have_id) { idv = id[v]; have_id = true; }
                less = idv < id[vb];
            }
            if (!less) break;
            keys[b + 1] = kb;
            vals[b + 1] = vb;
            b--;
        }
        keys[b + 1] = ka;
        vals[b + 1] = v;
    }
}

// Long runs: one block per run (grid-stride over the list), bitonic sort by (key, id)
// in LDS, padded to a power of two with all-ones entries.
constexpr int LR_T = 256;
__global__ __launch_bounds__(LR_T) void rs_long_runs_kernel(uint64_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                            const int64_t *__restrict__ id, DevScalars *s,
                                                            const int64_t *__restrict__ long_runs,
                                                            const int64_t *__restrict__ n_long)
{
    __shared__ uint64_t sk[RUN_LONG];
    __shared__ int64_t sid[RUN_LONG];
    __shared__ uint32_t sv[RUN_LONG];
    __shared__ int64_t len_s;
    const int64_t nl = *n_long < SORT_LONG_RUNS ? *n_long : SORT_LONG_RUNS;
    const int64_t n = s->n_max_new;
    const int t = threadIdx.x;
    for (int64_t r = blockIdx.x; r < nl; r += gridDim.x) {
        const int64_t i = long_runs[r];
        if (t == 0) {
            // keys are ordered by prefix: the last index with the run's prefix by bisection
            const uint64_t p = keys[i] >> RS_SHIFT0;
            int64_t lo = i, hi = i + RUN_LONG < n ? i + RUN_LONG : n - 1;
            if ((keys[hi] >> RS_SHIFT0) == p) {
                len_s = hi == i + RUN_LONG ? -1 : hi - i + 1;     // RUN_LONG + 1 keys do not fit
            } else {
                while (hi - lo > 1) {
                    const int64_t m = lo + ((hi - lo) >> 1);
                    if ((keys[m] >> RS_SHIFT0) == p) lo = m;
                    else hi = m;
                }
                len_s = lo - i + 1;
            }
            if (len_s < 0) dev_set_error(s, CMCG_E_TIES, i + 1);
        }
        __syncthreads();
        const int64_t len = len_s;
        __syncthreads();
        if (len < 0) continue;
        int P = 1;
        while (P < len) P <<= 1;
        for (int q = t; q < P; q += LR_T) {
            if (q < len) {
                const uint32_t v = vals[i + q];
                sk[q] = keys[i + q];
                sv[q] = v;
                sid[q] = id[v];
            } else {
                sk[q] = ~0ull;
                sid[q] = INT64_MAX;
                sv[q] = 0u;
            }
        }
        __syncthreads();
        for (int kk = 2; kk <= P; kk <<= 1)
            for (int j = kk >> 1; j > 0; j >>= 1) {
                for (int q = t; q < P; q += LR_T) {
                    const int x = q ^ j;
                    if (x > q) {
                        const bool up = (q & kk) == 0;
                        const bool gt = key_id_less(sk[x], sid[x], sk[q], sid[q]);
                        if (gt == up) {
                            uint64_t tk = sk[q]; sk[q] = sk[x]; sk[x] = tk;
                            int64_t ti = sid[q]; sid[q] = sid[x]; sid[x] = ti;
                            uint32_t tv = sv[q]; sv[q] = sv[x]; sv[x] = tv;
                        }
                    }
                }
                __syncthreads();
            }
        for (int q = t; q < len; q += LR_T) {
            keys[i + q] = sk[q];
            vals[i + q] = sv[q];
        }
        __syncthreads();
    }
}

// number of live stars = keys below the SF_INFINITY key
__global__ void rs_count_kernel(const uint64_t *__restrict__ keys, DevScalars *s, uint64_t kinf)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = s->n_max_new;
    if (i > n) return;
    bool below_prev = (i == 0) ? true : keys[i - 1] < kinf;
    bool below_here = (i == n) ? false : keys[i] < kinf;
    if (below_prev && !below_here) s->sort_n = i;
}

__global__ __launch_bounds__(256) void rs_gather_kernel(const uint64_t *__restrict__ keys,
                                                        const uint32_t *__restrict__ vals, DevScalars *s,
                                                        double *__restrict__ r_out, GatherList g)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = s->sort_n;
    if (i > n) return;
    if (i == n) {   // the outer sentinel (potential_calculate sets it again)
        r_out[n + 1] = CMCG_SF_INFINITY;
        return;
    }
    uint32_t src = vals[i];
    r_out[i + 1] = unsortable(keys[i]);
    for (int f = 0; f < g.nf; f++) g.dst[f][i + 1] = g.src[f][src];
}

__global__ __launch_bounds__(256) void rs_gather_aos_kernel(const uint64_t *__restrict__ keys,
                                                            const uint32_t *__restrict__ vals, DevScalars *s,
                                                            double *__restrict__ r_out,
                                                            const uint64_t *__restrict__ aos, GatherList g)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = s->sort_n;
    if (i > n) return;
    if (i == n) {
        r_out[n + 1] = CMCG_SF_INFINITY;
        return;
    }
    uint32_t src = vals[i];
    r_out[i + 1] = unsortable(keys[i]);
    const ulonglong2 *rec = (const ulonglong2 *)(aos + (int64_t)AOS_W * (src - 1));
    uint64_t w[AOS_W];
#pragma unroll
    for (int q = 0; q < AOS_W / 2; q++) {
        ulonglong2 v = 2 * q < g.nf ? rec[q] : ulonglong2{0ull, 0ull};
        w[2 * q] = v.x;
        w[2 * q + 1] = v.y;
    }
#pragma unroll
    for (int f = 0; f < AOS_W; f++)
        if (f < g.nf) g.dst[f][i + 1] = w[f];
}

__global__ void rs_finish_kernel(DevScalars *s)
{
    s->n_max = s->sort_n;
    s->n_max_new = s->sort_n;
}

__global__ void rs_begin_kernel(DevScalars *s, int64_t *n_long)
{
    s->sort_n = 0;
    s->sort_in = s->n_max_new;
    *n_long = 0;
}
} // namespace

// Sort slots 1..n_max_new by (key field, id).  src/dst are parallel lists of payload
// fields; key_src is the field holding the radii to sort by; r_dst receives them sorted.
hipError_t launch_sort_fields(const double *key_src, double *r_dst, const int64_t *id_src, int nf,
                              const uint64_t *const *src, uint64_t *const *dst, DevScalars *s, const Scratch &sc,
                              int64_t n_ub, hipStream_t st)
{
    if (nf > RS_MAX_FIELDS) return hipErrorInvalidValue;
    GatherList g;
    g.nf = nf;
    for (int f = 0; f < nf; f++) { g.src[f] = src[f]; g.dst[f] = dst[f]; }
    for (int f = nf; f < RS_MAX_FIELDS; f++) { g.src[f] = nullptr; g.dst[f] = nullptr; }
    uint64_t *aos = (nf <= AOS_W && sc.sort_aos && n_ub <= sc.sort_aos_cap) ? sc.sort_aos : nullptr;
    int64_t ntiles = (n_ub + RS_TILE - 1) / RS_TILE;
    if (ntiles < 1) ntiles = 1;
    if (ntiles > sc.max_sort_tiles) return hipErrorInvalidValue;
    const uint64_t kinf = 0x8000000000000000ull | 0x4202a05f20000000ull;   // sortable(1.0e10)
    hipLaunchKernelGGL(rs_begin_kernel, dim3(1), dim3(1), 0, st, s, sc.sort_nlong);
    hipLaunchKernelGGL(rs_zero_kernel, dim3(8), dim3(256), 0, st, sc.digit_totals, (int64_t)(8 * 256));
    // the passes alternate between the two buffer pairs; start so that the sorted keys
    // and the permutation end in keys_a / vals_a (set_velocities3 reads vals_a)
    constexpr bool odd = RS_PASSES % 2 == 1;
    uint64_t *kin = odd ? sc.keys_b : sc.keys_a, *kout = odd ? sc.keys_a : sc.keys_b;
    uint32_t *vin = odd ? sc.vals_b : sc.vals_a, *vout = odd ? sc.vals_a : sc.vals_b;
    hipLaunchKernelGGL(rs_prep_kernel, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, st, key_src, s, kin, vin,
                       sc.digit_totals, aos, g, kinf);
    for (int pass = 8 - RS_PASSES; pass < 8; pass++) {
        int shift = 8 * pass;
        hipLaunchKernelGGL(rs_hist_kernel, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, st, kin, s, shift,
                           sc.block_hist, ntiles);
        hipLaunchKernelGGL(rs_scan_kernel, dim3(256), dim3(RS_THREADS), 0, st, sc.block_hist, ntiles,
                           (const uint32_t *)(sc.digit_totals + 256 * pass));
        hipLaunchKernelGGL(rs_scatter_kernel, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, st, kin, vin, kout, vout,
                           s, shift, (const uint32_t *)sc.block_hist, ntiles);
        uint64_t *tk = kin; kin = kout; kout = tk;
        uint32_t *tv = vin; vin = vout; vout = tv;
    }
    unsigned g1 = (unsigned)((n_ub + 256) / 256);
    hipLaunchKernelGGL(rs_runs_kernel, dim3(g1), dim3(256), 0, st, kin, vin, id_src, s, sc.sort_long, sc.sort_nlong);
    hipLaunchKernelGGL(rs_long_runs_kernel, dim3(64), dim3(LR_T), 0, st, kin, vin, id_src, s,
                       (const int64_t *)sc.sort_long, (const int64_t *)sc.sort_nlong);
    hipLaunchKernelGGL(rs_count_kernel, dim3(g1), dim3(256), 0, st, (const uint64_t *)kin, s, kinf);
    if (aos)
        hipLaunchKernelGGL(rs_gather_aos_kernel, dim3(g1), dim3(256), 0, st, (const uint64_t *)kin,
                           (const uint32_t *)vin, s, r_dst, (const uint64_t *)aos, g);
    else
        hipLaunchKernelGGL(rs_gather_kernel, dim3(g1), dim3(256), 0, st, (const uint64_t *)kin, (const uint32_t *)vin,
                           s, r_dst, g);
    hipLaunchKernelGGL(rs_finish_kernel, dim3(1), dim3(1), 0, st, s);
    return hipGetLastError();
}
