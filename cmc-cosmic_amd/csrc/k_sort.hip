// a1: qsorts_new -> sample_sort at one rank (src/cmc/cmc_sort.c:375-629): order the
// stars by r ascending (compare_type, :245-250), ties by id (the reference's libc
// qsort leaves them unordered), drop the r == SF_INFINITY tail
// (remove_stripped_stars, :343-354), then refresh r and m (post_sort_comm,
// src/cmc/cmc_utils.c:2259-2288).
//
// Device algorithm: LSD radix sort of one 64-bit word per star: a 32-bit key z on
// top (the order-preserving bits of r, shifted by the live minimum and scaled so the
// live range fills 31 bits; the SF_INFINITY tail gets z = 2^32 - 1) and the star's
// slot below.  Four passes of 8-bit digits over z, each three kernels (tile
// histogram, per-digit tile scan, stable scatter).  The scatter ranks each 4096-word
// tile with 64-lane ballot multisplit (stable in tile order), stages the sorted tile
// in LDS and writes it out digit-run by digit-run, so global writes are coalesced.
// Runs of equal z are then put in (r, id) order (in place up to 64 words, by a bitonic
// sort in LDS up to 4096, by a merge sort in global memory beyond: any input sorts),
// and one gather moves the payload.
#include "kernels.hpp"

namespace {
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = SORT_TILE / RS_THREADS;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_MAX_FIELDS = 24;
// LSD passes over z, the top 32 bits of each word; words equal in z are then put in
// (r, id) order by the run pass (one thread per run up to RUN_SHORT words, one block
// per longer run: sorting in LDS up to RUN_LONG words, a merge sort beyond).
constexpr int RS_PASSES = 4;
constexpr int RS_SHIFT0 = 64 - 8 * RS_PASSES;     // z = word >> RS_SHIFT0
constexpr uint32_t Z_TAIL = 0xffffffffu;          // z of the SF_INFINITY tail; live z < 2^31
static_assert(RS_PASSES % 2 == 0, "the sorted words must end in keys_a");
constexpr int RUN_SHORT = SORT_RUN_SHORT;
constexpr int RUN_LONG = SORT_RUN_LONG;
// Packed payload records: W = 8 words (64 B) when the fields and r fit, else 16 (one
// 128-B line); an odd LDS stride per record keeps the packing free of bank conflicts.
constexpr int AOS_MAX = SORT_AOS_WORDS;

__device__ __forceinline__ uint64_t sortable(double x)
{
    uint64_t b = (uint64_t)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

struct GatherList {
    int nf;
    const uint64_t *src[RS_MAX_FIELDS];
    uint64_t *dst[RS_MAX_FIELDS];
};

__device__ __forceinline__ uint32_t word_slot(uint64_t w) { return (uint32_t)w; }
__device__ __forceinline__ uint32_t word_z(uint64_t w) { return (uint32_t)(w >> RS_SHIFT0); }

// Minimum and maximum order-preserving key of the live stars (r below SF_INFINITY), one
// pair per block: mm[2b] = min, mm[2b + 1] = max (~0 and 0 for a block that saw none).
// rs_prep_kernel combines the pairs; partials per block replace thousands of device-scope
// atomics on one address arriving from every XCD.  The kernel also zeroes the digit
// totals (zn words at z) and resets the sort's scalars and long-run counters.
struct rs_minmax_kernel {
    static constexpr int LB = 256, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ keysrc, DevScalars *s,
                                                        uint64_t kinf, unsigned long long *__restrict__ mm,
                                                        uint32_t *__restrict__ z, int64_t zn, int64_t *__restrict__ n_long)
{
    __shared__ unsigned long long rmin[4], rmax[4];
    const int64_t n = s->n_max_new;
    const int64_t stride = (int64_t)kg_.nb * blockDim.x;
    for (int64_t q = (int64_t)kg_.bid * blockDim.x + threadIdx.x; q < zn; q += stride) z[q] = 0u;
    if (kg_.bid == 0 && threadIdx.x == 0) {
        s->sort_n = 0;
        s->sort_in = n;
        n_long[0] = 0;
    }
    unsigned long long lo = ~0ull, hi = 0ull;
    auto take = [&](uint64_t k) {
        if (k < kinf) {
            lo = k < lo ? k : lo;
            hi = k > hi ? k : hi;
        }
    };
    int64_t i = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {       // four independent loads in flight
        const uint64_t k0 = sortable(keysrc[i + 1]), k1 = sortable(keysrc[i + stride + 1]);
        const uint64_t k2 = sortable(keysrc[i + 2 * stride + 1]), k3 = sortable(keysrc[i + 3 * stride + 1]);
        take(k0); take(k1); take(k2); take(k3);
    }
    for (; i < n; i += stride) take(sortable(keysrc[i + 1]));
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long a = __shfl_down(lo, off, 64), b = __shfl_down(hi, off, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { rmin[w] = lo; rmax[w] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 4; q++) {
            lo = rmin[q] < lo ? rmin[q] : lo;
            hi = rmax[q] > hi ? rmax[q] : hi;
        }
        mm[2 * kg_.bid] = lo;
        mm[2 * kg_.bid + 1] = hi;
    }
}
};

// Blocks of rs_minmax_kernel for a sort over `ntiles` tiles (the same count on both sides).
__host__ __device__ __forceinline__ int minmax_parts(int64_t ntiles)
{
    return ntiles * 4 < SORT_MM_PARTS ? (int)(ntiles * 4) : SORT_MM_PARTS;
}

// Words from slots 1..n plus the four global digit histograms, and the first pass's
// per-tile histogram (a block's words are that pass's tile); with `aos` the payload
// fields of each slot, and r itself after them, are packed into one 128-byte record,
// so the final gather reads one record per star instead of one cache line per field.
template <int W>
struct rs_prep_kernel {
    static constexpr int LB = RS_THREADS, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ keysrc, DevScalars *s,
                                                             uint64_t *__restrict__ words,
                                                             uint32_t *__restrict__ digit_totals,
                                                             uint64_t *__restrict__ aos, GatherList g, uint64_t kinf,
                                                             const unsigned long long *__restrict__ mm,
                                                             uint32_t *__restrict__ block_hist, int64_t ntiles)
{
    constexpr int P0 = 8 - RS_PASSES;      // first byte sorted on
    __shared__ uint32_t h[RS_PASSES][256];
    constexpr int LD = W + 1;
    __shared__ uint64_t recs[RS_THREADS * LD];         // one chunk of packed records
    const int t = threadIdx.x;
    for (int p = 0; p < RS_PASSES; p++) h[p][t] = 0;
    // combine the per-block (min, max) pairs of rs_minmax_kernel
    __shared__ unsigned long long red[2][RS_THREADS / 64];
    {
        unsigned long long lo = ~0ull, hi = 0ull;
        for (int q = t; q < minmax_parts(kg_.nb); q += RS_THREADS) {
            const unsigned long long a = mm[2 * q], b = mm[2 * q + 1];
            lo = a < lo ? a : lo;
            hi = b > hi ? b : hi;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long a = __shfl_down(lo, off, 64), b = __shfl_down(hi, off, 64);
            lo = a < lo ? a : lo;
            hi = b > hi ? b : hi;
        }
        if ((t & 63) == 0) { red[0][t >> 6] = lo; red[1][t >> 6] = hi; }
    }
    __syncthreads();
    // z = (key - min) >> S with S chosen so the live range fits 31 bits
    uint64_t kmin = red[0][0], kmax = red[1][0];
    for (int q = 1; q < RS_THREADS / 64; q++) {
        kmin = red[0][q] < kmin ? red[0][q] : kmin;
        kmax = red[1][q] > kmax ? red[1][q] : kmax;
    }
    int S = 0;
    if (kmax >= kmin) {
        const uint64_t range = kmax - kmin;
        const int bits = range ? 64 - __clzll((long long)range) : 0;
        S = bits > 31 ? bits - 31 : 0;
    } else {
        kmin = 0;      // no live star
    }
    __syncthreads();
    const int64_t n = s->n_max_new;
    const int64_t base = (int64_t)kg_.bid * RS_TILE;
    auto key_word = [&](double r, int64_t i) {
        const uint64_t k = sortable(r);
        const uint32_t z = k >= kinf ? Z_TAIL : (uint32_t)((k - kmin) >> S);
        const uint64_t w = ((uint64_t)z << RS_SHIFT0) | (uint64_t)(uint32_t)(i + 1);
        words[i] = w;
        for (int p = P0; p < 8; p++) atomicAdd(&h[p - P0][(w >> (8 * p)) & 255], 1u);
    };
    if (aos && g.nf == W - 1) {
        // the usual full record, software-pipelined: the next chunk's field loads are in
        // flight while this chunk goes through LDS
        uint64_t v[W - 1];
        double kr = 0.0;
        auto load = [&](int64_t i) {
            if (i < n) {
#pragma unroll
                for (int f = 0; f < W - 1; f++) v[f] = g.src[f][i + 1];
                kr = keysrc[i + 1];
            }
        };
        load(base + t);
        for (int u = 0; u < RS_ITEMS; u++) {
            const int64_t i0 = base + u * RS_THREADS;
            if (i0 >= n) break;                          // block-uniform
            const int64_t i = i0 + t;
            if (i < n) {
                key_word(kr, i);
#pragma unroll
                for (int f = 0; f < W - 1; f++) recs[LD * t + f] = v[f];
                recs[LD * t + W - 1] = (uint64_t)__double_as_longlong(kr);
            }
            if (u + 1 < RS_ITEMS) load(i + RS_THREADS);
            __syncthreads();
            const int64_t cnt = n - i0 < RS_THREADS ? n - i0 : RS_THREADS;
            uint64_t *dst = aos + (int64_t)W * i0;
            for (int w = t; w < cnt * W; w += RS_THREADS) dst[w] = recs[LD * (w / W) + w % W];
            __syncthreads();
        }
    } else {
        for (int u = 0; u < RS_ITEMS; u++) {
            const int64_t i0 = base + u * RS_THREADS;
            const int64_t i = i0 + t;
            if (i < n) key_word(keysrc[i + 1], i);
            if (aos && i0 < n) {        // block-uniform: pack the chunk through LDS, write it contiguously
                if (i < n) {
                    for (int f = 0; f < g.nf; f++) recs[LD * t + f] = g.src[f][i + 1];
                    recs[LD * t + g.nf] = (uint64_t)__double_as_longlong(keysrc[i + 1]);
                }
                __syncthreads();
                const int64_t cnt = n - i0 < RS_THREADS ? n - i0 : RS_THREADS;
                uint64_t *dst = aos + (int64_t)W * i0;
                // whole records (unused words included): partial writes cost more
                for (int w = t; w < cnt * W; w += RS_THREADS) dst[w] = recs[LD * (w / W) + w % W];
                __syncthreads();
            }
        }
    }
    __syncthreads();
    block_hist[(int64_t)t * ntiles + kg_.bid] = h[0][t];
    for (int p = P0; p < 8; p++)
        if (h[p - P0][t]) atomicAdd(&digit_totals[p * 256 + t], h[p - P0][t]);
}
};

struct rs_hist_kernel {
    static constexpr int LB = RS_THREADS, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const uint64_t *__restrict__ keys, DevScalars *s,
                                                             int shift, uint32_t *__restrict__ block_hist,
                                                             int64_t ntiles)
{
    __shared__ uint32_t h[256];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const int64_t n = s->n_max_new;
    const int64_t base = (int64_t)kg_.bid * RS_TILE;
    uint64_t k[RS_ITEMS];                   // every load of the tile in flight before the counting
#pragma unroll
    for (int u = 0; u < RS_ITEMS; u++) {
        const int64_t i = base + u * RS_THREADS + t;
        k[u] = i < n ? keys[i] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < RS_ITEMS; u++)
        if (base + u * RS_THREADS + t < n) atomicAdd(&h[(k[u] >> shift) & 255], 1u);
    __syncthreads();
    block_hist[(int64_t)t * ntiles + kg_.bid] = h[t];
}
};

// One block per digit: exclusive scan of that digit's tile counts, plus the digit's
// global base (sum of the lower digits' totals for this pass).
struct rs_scan_kernel {
    static constexpr int LB = RS_THREADS, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, uint32_t *__restrict__ block_hist, int64_t ntiles,
                                                             const uint32_t *__restrict__ totals)
{
    __shared__ uint32_t red[RS_THREADS];
    const int t = threadIdx.x;
    const int d = kg_.bid;
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
};

// Stable scatter of one 4096-word tile.  Wave w ranks the contiguous sub-tile
// [w*1024, (w+1)*1024) on its own (ballot matching of the 8 digit bits, a per-wave
// digit counter in LDS updated by the highest lane of each match group), so the
// item loop needs no block barrier; the (digit, wave) counts then give each wave's
// offset inside the tile's digit runs.  Index order (wave, item, lane) is the tile's
// index order, so equal digits keep their order.
struct rs_scatter_kernel {
    static constexpr int LB = RS_THREADS, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const uint64_t *__restrict__ words_in,
                                                                uint64_t *__restrict__ words_out, DevScalars *s,
                                                                int shift, const uint32_t *__restrict__ tile_off,
                                                                int64_t ntiles)
{
    constexpr int NW = RS_THREADS / 64;
    __shared__ uint64_t swords[RS_TILE];
    __shared__ uint32_t wcnt[NW][256];
    __shared__ uint32_t dstart[256];       // 38 KB in all: four blocks per CU
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t n = s->n_max_new;
    const int64_t base = (int64_t)kg_.bid * RS_TILE;
    const int64_t wbase = base + (int64_t)w * (64 * RS_ITEMS);
    for (int d = lane; d < 256; d += 64) wcnt[w][d] = 0;
    uint64_t k[RS_ITEMS];
    uint32_t rank[RS_ITEMS];
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    // every load of the wave's sub-tile in flight before the ranking
#pragma unroll
    for (int u = 0; u < RS_ITEMS; u++) {
        const int64_t i = wbase + u * 64 + lane;
        k[u] = i < n ? words_in[i] : 0ull;
    }
    // this tile's output base of every digit
    __shared__ uint32_t dout[256];
    dout[t] = tile_off[(int64_t)t * ntiles + kg_.bid];
#pragma unroll
    for (int u = 0; u < RS_ITEMS; u++) {
        int64_t i = wbase + u * 64 + lane;
        bool valid = i < n;
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
            swords[p] = k[u];
        }
    }
    __syncthreads();
    const int64_t cnt_tile = n - base < RS_TILE ? n - base : RS_TILE;
    for (int u = 0; u < RS_ITEMS; u++) {
        int p = u * RS_THREADS + t;
        if (p < cnt_tile) {
            uint64_t word = swords[p];
            uint32_t d = (uint32_t)((word >> shift) & 255);
            int64_t g = (int64_t)dout[d] + (p - (int64_t)dstart[d]);
            words_out[g] = word;
        }
    }
}
};

// Runs of words equal in z (live stars only: the tail has z = Z_TAIL) are put in
// (r, id) order, r read back through the slot.  Run starts sort runs of up to
// RUN_SHORT words themselves (insertion sort) and list longer ones for
// rs_long_runs_kernel.
__device__ __forceinline__ bool key_id_less(uint64_t ka, int64_t ia, uint64_t kb, int64_t ib)
{
    return ka < kb || (ka == kb && ia < ib);
}

struct rs_runs_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, uint64_t *__restrict__ words, const double *__restrict__ keysrc,
                               const int64_t *__restrict__ id, DevScalars *s, int64_t *__restrict__ long_runs,
                               int64_t *__restrict__ n_long, int64_t long_cap)
{
    const int64_t i = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    const int64_t n = s->n_max_new;
    if (i > n || n <= 0) return;       // n <= 0: sort_n stays 0
    // the words at i - 1, i and i + 1, loaded together (indices clamped into [0, n),
    // the clamped ones are not compared)
    const uint64_t wp = words[i > 0 ? i - 1 : 0];
    const uint64_t wi = words[i < n ? i : n - 1];
    const uint64_t wn = words[i + 1 < n ? i + 1 : n - 1];
    // the number of live stars: the position where the words reach the tail's z (runs
    // reorder words of one z only, so the z sequence read here is final)
    const bool below_prev = i == 0 || word_z(wp) != Z_TAIL;
    const bool below_here = i < n && word_z(wi) != Z_TAIL;
    if (below_prev && !below_here) s->sort_n = i;
    if (i + 1 >= n) return;
    const uint32_t z = word_z(wi);
    if (z == Z_TAIL) return;
    if (word_z(wn) != z) return;
    if (i > 0 && word_z(wp) == z) return;
    int64_t e = i + 1;
    while (e + 1 < n && word_z(words[e + 1]) == z && e - i < RUN_SHORT) e++;
    if (e + 1 < n && word_z(words[e + 1]) == z) {
        unsigned long long slot = atomicAdd((unsigned long long *)n_long, 1ull);
        if ((int64_t)slot < long_cap) long_runs[slot] = i;          // the list holds n / 65 runs
        else dev_set_error(s, CMCG_E_STATE, i + 1);
        return;
    }
    for (int64_t a = i + 1; a <= e; a++) {
        const uint64_t wa = words[a];
        const uint32_t va = word_slot(wa);
        const uint64_t ka = sortable(keysrc[va]);
        const int64_t ida = id[va];
        int64_t b = a - 1;
        while (b >= i) {
            const uint64_t wb = words[b];
            const uint32_t vb = word_slot(wb);
            if (!key_id_less(ka, ida, sortable(keysrc[vb]), id[vb])) break;
            words[b + 1] = wb;
            b--;
        }
        words[b + 1] = wa;
    }
}
};

// Long runs: one block per run (grid-stride over the list), bitonic sort by (r, id)
// in LDS, padded to a power of two with all-ones entries.  Every word of a run has
// the run's z, so only the slots travel.
constexpr int LR_T = 256;
// Bitonic sort by (r, id) of words[i .. i + len) (len <= RUN_LONG, one z) in LDS.
__device__ void lds_sort_run(uint64_t *__restrict__ words, const double *__restrict__ keysrc,
                             const int64_t *__restrict__ id, int64_t i, int64_t len, uint32_t z,
                             uint64_t *sk, int64_t *sid, uint32_t *sv)
{
    const int t = threadIdx.x;
    int P = 1;
    while (P < len) P <<= 1;
    for (int q = t; q < P; q += LR_T) {
        if (q < len) {
            const uint32_t v = word_slot(words[i + q]);
            sk[q] = sortable(keysrc[v]);
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
    for (int q = t; q < len; q += LR_T) words[i + q] = ((uint64_t)z << RS_SHIFT0) | sv[q];
    __syncthreads();
}

// Runs longer than RUN_LONG: chunks of RUN_LONG sorted in LDS, then bottom-up merges
// between words and tmp (the same positions of the second key buffer), each merge
// split among the threads along its merge path.
__device__ __forceinline__ bool word_less(uint64_t wa, uint64_t wb, const double *__restrict__ keysrc,
                                          const int64_t *__restrict__ id)
{
    const uint32_t va = word_slot(wa), vb = word_slot(wb);
    return key_id_less(sortable(keysrc[va]), id[va], sortable(keysrc[vb]), id[vb]);
}

__device__ void merge_sort_run(uint64_t *__restrict__ words, uint64_t *__restrict__ tmp,
                               const double *__restrict__ keysrc, const int64_t *__restrict__ id, int64_t i,
                               int64_t len, uint32_t z, uint64_t *sk, int64_t *sid, uint32_t *sv)
{
    const int t = threadIdx.x;
    for (int64_t c = 0; c < len; c += RUN_LONG)
        lds_sort_run(words, keysrc, id, i + c, len - c < RUN_LONG ? len - c : RUN_LONG, z, sk, sid, sv);
    uint64_t *src = words + i, *dst = tmp + i;
    for (int64_t w = RUN_LONG; w < len; w <<= 1) {
        for (int64_t p = 0; p < len; p += 2 * w) {
            const int64_t la = len - p < w ? len - p : w;
            const int64_t lb = len - p - la < w ? len - p - la : w;
            const uint64_t *A = src + p, *B = src + p + la;
            const int64_t L = la + lb;
            const int64_t d0 = L * t / LR_T, d1 = L * (t + 1) / LR_T;
            // merge path: the first d0 outputs hold A[0, ia) and B[0, d0 - ia); A
            // goes first on equal keys.  A[m] is among them iff A[m] <= B[d0 - 1 - m].
            int64_t a_lo = d0 > lb ? d0 - lb : 0, a_hi = d0 < la ? d0 : la;
            while (a_lo < a_hi) {
                const int64_t m = (a_lo + a_hi) >> 1;
                if (!word_less(B[d0 - 1 - m], A[m], keysrc, id)) a_lo = m + 1;
                else a_hi = m;
            }
            int64_t ia = a_lo, ib = d0 - a_lo;
            for (int64_t d = d0; d < d1; d++) {
                const bool takeA = ib >= lb || (ia < la && !word_less(B[ib], A[ia], keysrc, id));
                dst[p + d] = takeA ? A[ia++] : B[ib++];
            }
        }
        __syncthreads();
        uint64_t *x = src; src = dst; dst = x;
    }
    if (src != words + i)
        for (int64_t q = t; q < len; q += LR_T) words[i + q] = src[q];
    __syncthreads();
}

// One block per listed run (grid-stride over the list): its length by bisection (words
// are ordered by z), then the LDS bitonic sort, or the merge sort beyond RUN_LONG words.
struct rs_long_runs_kernel {
    static constexpr int LB = LR_T, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, uint64_t *__restrict__ words, uint64_t *__restrict__ tmp,
                                                            const double *__restrict__ keysrc,
                                                            const int64_t *__restrict__ id, DevScalars *s,
                                                            const int64_t *__restrict__ long_runs,
                                                            const int64_t *__restrict__ n_long, int64_t long_cap)
{
    __shared__ uint64_t sk[RUN_LONG];
    __shared__ int64_t sid[RUN_LONG];
    __shared__ uint32_t sv[RUN_LONG];
    __shared__ int64_t len_s;
    const int64_t nl = n_long[0] < long_cap ? n_long[0] : long_cap;
    const int64_t n = s->n_max_new;
    const int t = threadIdx.x;
    for (int64_t r = kg_.bid; r < nl; r += kg_.nb) {
        const int64_t i = long_runs[r];
        const uint32_t z = word_z(words[i]);
        if (t == 0) {
            // the last index with the run's z: words[lo] has it, words[hi] not (or lo == hi)
            int64_t lo = i, hi = i + RUN_LONG < n - 1 ? i + RUN_LONG : n - 1;
            if (word_z(words[hi]) == z) { lo = hi; hi = n - 1; }     // more than RUN_LONG words
            if (word_z(words[hi]) == z) lo = hi;
            while (hi - lo > 1) {
                const int64_t m = lo + ((hi - lo) >> 1);
                if (word_z(words[m]) == z) lo = m;
                else hi = m;
            }
            len_s = lo - i + 1;
        }
        __syncthreads();
        const int64_t len = len_s;
        __syncthreads();
        if (len <= RUN_LONG) lds_sort_run(words, keysrc, id, i, len, z, sk, sid, sv);
        else merge_sort_run(words, tmp, keysrc, id, i, len, z, sk, sid, sv);
    }
}
};

// The sorted permutation goes to perm (slot of the star now at position i), r and the
// payload fields to their destinations.
struct rs_gather_kernel {
    static constexpr int LB = 256, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const uint64_t *__restrict__ words, DevScalars *s,
                                                        const double *__restrict__ keysrc, double *__restrict__ r_out,
                                                        uint32_t *__restrict__ perm, GatherList g)
{
    int64_t i = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    const int64_t n = s->sort_n;
    if (i == 0) { s->n_max = n; s->n_max_new = n; }     // no block reads them
    if (i > n) return;
    if (i == n) {   // the outer sentinel (potential_calculate sets it again)
        r_out[n + 1] = CMCG_SF_INFINITY;
        return;
    }
    const uint32_t src = word_slot(words[i]);
    perm[i] = src;
    r_out[i + 1] = keysrc[src];
    for (int f = 0; f < g.nf; f++) g.dst[f][i + 1] = g.src[f][src];
}
};

template <int W>
struct rs_gather_aos_kernel {
    static constexpr int LB = 256, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const uint64_t *__restrict__ words, DevScalars *s,
                                               double *__restrict__ r_out, uint32_t *__restrict__ perm,
                                               const uint64_t *__restrict__ aos, GatherList g)
{
    // The block's 256 records are read cooperatively, W/2 lanes per record (one
    // 16-byte piece each, so each load instruction covers whole records), through LDS;
    // then each thread writes its star's fields, coalesced across the block.
    constexpr int PC = W / 2;                          // 16-byte pieces per record
    __shared__ ulonglong2 recs[256 * PC + 256 * PC / 32];   // one pad slot per 32 pieces
    __shared__ uint32_t ssrc[256];
    const int t = threadIdx.x;
    const int64_t i = (int64_t)kg_.bid * 256 + t;
    const int64_t n = s->sort_n;
    if (i == 0) { s->n_max = n; s->n_max_new = n; }     // no block reads them
    uint32_t src = 0;
    if (i < n) {
        src = word_slot(words[i]);
        perm[i] = src;
    } else if (i == n) {
        r_out[n + 1] = CMCG_SF_INFINITY;
    }
    ssrc[t] = src;
    __syncthreads();
    const int64_t live = n - (int64_t)kg_.bid * 256;   // stars of this block (may be <= 0)
#pragma unroll
    for (int u = 0; u < PC; u++) {
        const int task = u * 256 + t, sl = task / PC, pc = task % PC;
        if (sl < live) {
            const ulonglong2 *rec = (const ulonglong2 *)(aos + (int64_t)W * (ssrc[sl] - 1));
            recs[task + task / 32] = 2 * pc <= g.nf ? rec[pc] : ulonglong2{0ull, 0ull};    // fields, then r in word nf
        }
    }
    __syncthreads();
    if (i >= n) return;
    uint64_t w[W];
#pragma unroll
    for (int q = 0; q < PC; q++) {
        const int task = t * PC + q;
        const ulonglong2 v = recs[task + task / 32];
        w[2 * q] = v.x;
        w[2 * q + 1] = v.y;
    }
#pragma unroll
    for (int f = 0; f < W; f++) {
        if (f < g.nf) g.dst[f][i + 1] = w[f];
        if (f == g.nf) r_out[i + 1] = __longlong_as_double((long long)w[f]);
    }
}
};

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
    // packed records hold the fields and r after them
    uint64_t *aos = (nf < AOS_MAX && sc.sort_aos && n_ub <= sc.sort_aos_cap) ? sc.sort_aos : nullptr;
    const bool narrow = nf < 8;          // fields and r in 8 words
    int64_t ntiles = (n_ub + RS_TILE - 1) / RS_TILE;
    if (ntiles < 1) ntiles = 1;
    if (ntiles > sc.max_sort_tiles) return hipErrorInvalidValue;
    const uint64_t kinf = 0x8000000000000000ull | 0x4202a05f20000000ull;   // sortable(1.0e10)
    unsigned long long *mm = (unsigned long long *)sc.sort_mm;
    kl<rs_minmax_kernel>(dim3((unsigned)minmax_parts(ntiles)), dim3(256), 0, st, key_src, s, kinf,
                       mm, sc.digit_totals, (int64_t)(8 * 256), sc.sort_nlong);
    // an even number of passes: the sorted words end in keys_a
    uint64_t *kin = sc.keys_a, *kout = sc.keys_b;
    if (narrow)
        kl<rs_prep_kernel<8>>(dim3((unsigned)ntiles), dim3(RS_THREADS), 0, st, key_src, s, kin,
                           sc.digit_totals, aos, g, kinf, (const unsigned long long *)mm, sc.block_hist, ntiles);
    else
        kl<rs_prep_kernel<AOS_MAX>>(dim3((unsigned)ntiles), dim3(RS_THREADS), 0, st, key_src, s, kin,
                           sc.digit_totals, aos, g, kinf, (const unsigned long long *)mm, sc.block_hist, ntiles);
    for (int pass = 8 - RS_PASSES; pass < 8; pass++) {
        int shift = 8 * pass;
        if (pass > 8 - RS_PASSES)          // the first pass's tile counts come from rs_prep_kernel
            kl<rs_hist_kernel>(dim3((unsigned)ntiles), dim3(RS_THREADS), 0, st, kin, s, shift,
                           sc.block_hist, ntiles);
        kl<rs_scan_kernel>(dim3(256), dim3(RS_THREADS), 0, st, sc.block_hist, ntiles,
                           (const uint32_t *)(sc.digit_totals + 256 * pass));
        kl<rs_scatter_kernel>(dim3((unsigned)ntiles), dim3(RS_THREADS), 0, st, kin, kout, s, shift,
                           (const uint32_t *)sc.block_hist, ntiles);
        uint64_t *tk = kin; kin = kout; kout = tk;
    }
    unsigned g1 = (unsigned)((n_ub + 256) / 256);
    kl<rs_runs_kernel>(dim3(g1), dim3(256), 0, st, kin, key_src, id_src, s, sc.sort_long,
                       sc.sort_nlong, sc.sort_long_cap);
    kl<rs_long_runs_kernel>(dim3(64), dim3(LR_T), 0, st, kin, kout, key_src, id_src, s,
                       (const int64_t *)sc.sort_long, (const int64_t *)sc.sort_nlong, sc.sort_long_cap);
    if (aos && narrow)
        kl<rs_gather_aos_kernel<8>>(dim3(g1), dim3(256), 0, st, (const uint64_t *)kin, s, r_dst,
                           sc.vals_a, (const uint64_t *)aos, g);
    else if (aos)
        kl<rs_gather_aos_kernel<AOS_MAX>>(dim3(g1), dim3(256), 0, st, (const uint64_t *)kin, s, r_dst,
                           sc.vals_a, (const uint64_t *)aos, g);
    else
        kl<rs_gather_kernel>(dim3(g1), dim3(256), 0, st, (const uint64_t *)kin, s, key_src, r_dst,
                           sc.vals_a, g);
    return hipGetLastError();
}
