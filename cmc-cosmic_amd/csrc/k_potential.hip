// a2: potential_calculate (src/cmc/cmc_utils.c:675-736), reproducing the reference's
// sequential FP64 rounding bit for bit with parallel scans.
//
// The reference evaluates three floating-point chains one element at a time:
//   A  mprev += m_k                  k = 1..N          (total mass, :685-696)
//   B  mprev -= m_k / N_STAR         k = N..2          (enclosed mass M_k, :723)
//   C  phi_k = phi_{k+1} - M_k (1/r_k - 1/r_{k+1})     (:722)
// Each chain is v_j = fl(v_{j-1} + s*x_j) and is monotone.  Repeated rounding makes
// the result drift from the exact sums (6e-11 relative at N = 1e5), so a correctly
// rounded parallel scan does not reproduce it.  Exact reproduction uses:
//
// * If v_{j-1} and v_j lie in one binade with ulp u, then fl(v_{j-1} + s x_j) =
//   v_{j-1} + RN_u(s x_j) unless s x_j / u is a tie: the increment does not depend
//   on v_{j-1}.  Such "regular" elements are recognised from the exact prefix sum
//   of the x (2^-K fixed point, int128) and a rigorous bound on the accumulated
//   rounding error; the rest ("sensitive": binade crossings, ties, tiny values) are
//   few, about one per binade.
// * Segment and walk: an exact prefix sum of the regular increments, then one
//   thread walks the sensitive elements in order, each getting
//   fl(v_prev + s x) from the exact value before it (the run of regular elements
//   in between stays in one binade, so adding the run's sum is exact), then every
//   element is rebuilt in parallel from the last sensitive value before it.
// * Every element is verified against fl(v_{j-1} + s x_j).  A violation (only when
//   the error bound or the sensitive-list capacity is exceeded) is repaired by
//   Jacobi rounds in which each bracket fl(g_{i-1} + s x_i) - g_{i-1} of the
//   previous iterate g is an exact int128 and each round restarts at the first
//   violation, and finally by a sequential pass.  The result equals the sequential
//   chain exactly.
#include "kernels.hpp"

namespace {
constexpr int CT = 256;             // threads per tile
constexpr int CI = 8;               // items per thread
constexpr int CTILE = CT * CI;
constexpr int SCAN_T = 1024;
constexpr int REPAIR_ROUNDS = 1;      // Jacobi rounds after the segment-and-walk pass (then sequential)
constexpr int RAW_CAP = 2048;         // sensitive elements sorted in LDS by the walk

typedef __int128 i128;

struct ChainArgs {
    const double *x;      // increments in processing order, x[0..L-1]
    double *g;            // iterate, updated in place, g[0..L-1]
    double sign;          // +1 or -1
    const double *v0p;    // start value v_{-1} (device scalar)
    int64_t *first_bad;   // [0] restart index of this round, [1] first violation found
    i128 *tsum;           // per-tile increment sums for the next round
    i128 *tpre;           // per-tile exclusive prefixes
    double *bnd;          // per-tile last value of the previous iterate
    const int64_t *Lp;    // chain length (device scalar)
    int K;                // fixed-point scale 2^-K
};

__device__ __forceinline__ i128 fxk(double d, int K)
{
    // d * 2^K as an integer, exact when d is a multiple of 2^-K (truncated otherwise)
    if (d == 0.0) return 0;
    bool neg = d < 0.0;
    double a = neg ? -d : d;
    int e;
    double mfrac = frexp(a, &e);
    uint64_t M = (uint64_t)ldexp(mfrac, 53);
    int sh = e - 53 + K;
    i128 v;
    if (sh >= 0) v = sh >= 74 ? ((i128)1 << 126) : ((i128)M << sh);
    else v = sh <= -64 ? (i128)0 : (i128)(M >> (-sh));
    return neg ? -v : v;
}

__device__ __forceinline__ double unfxk(i128 v, int K)
{
    bool neg = v < 0;
    unsigned __int128 u = neg ? (unsigned __int128)(-v) : (unsigned __int128)v;
    uint64_t hi = (uint64_t)(u >> 64), lo = (uint64_t)u;
    double d = ldexp((double)hi, 64) + (double)lo;    // exact when the value is a double
    d = ldexp(d, -K);
    return neg ? -d : d;
}

__device__ __forceinline__ i128 incr(double prev, double x, double sign, int K)
{
    double s = prev + sign * x;
    return fxk(s, K) - fxk(prev, K);
}

__device__ __forceinline__ i128 shfl_up_i128(i128 v, int d)
{
    const unsigned __int128 u = (unsigned __int128)v;
    unsigned long long lo = (unsigned long long)u, hi = (unsigned long long)(u >> 64);
    lo = __shfl_up(lo, d, 64);
    hi = __shfl_up(hi, d, 64);
    return (i128)(((unsigned __int128)hi << 64) | lo);
}

// block-wide exclusive scan of int128: inclusive scans inside each 64-lane wave by
// shuffles, then the wave totals through LDS (lds needs NT / 64 entries); every
// thread of the block must call it
template <int NT>
__device__ i128 block_excl_scan(i128 v, i128 *lds, i128 *total)
{
    constexpr int NW = NT / 64;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    i128 x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const i128 y = shfl_up_i128(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[w] = x;
    __syncthreads();
    i128 off = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < NW; q++) {
        const i128 s = lds[q];
        if (q < w) off += s;
        tot += s;
    }
    *total = tot;
    __syncthreads();
    return off + x - v;
}

// One block: takes the restart index found by the previous round, snapshots the
// tile-boundary values of the current iterate and scans the tile sums from the
// restart tile on.
__global__ __launch_bounds__(SCAN_T) void chain_scan_kernel(ChainArgs a)
{
    __shared__ i128 lds[SCAN_T];
    __shared__ int64_t fbs;
    const int64_t L = *a.Lp;
    if (threadIdx.x == 0) {
        int64_t fb = a.first_bad[1];
        if (a.first_bad[0] >= L) fb = L;          // already converged
        a.first_bad[0] = fb;
        a.first_bad[1] = L;
        fbs = fb;
    }
    __syncthreads();
    const int64_t fb = fbs;
    if (fb >= L) return;
    const int64_t t_lo = fb / CTILE;
    const int64_t ntiles = (L + CTILE - 1) / CTILE;
    for (int64_t b = (t_lo > 0 ? t_lo - 1 : 0) + threadIdx.x; b < ntiles; b += SCAN_T) {
        int64_t last = (b + 1) * CTILE - 1;
        a.bnd[b] = a.g[last < L ? last : L - 1];
    }
    const int64_t nt = ntiles - t_lo;
    const int64_t chunk = (nt + SCAN_T - 1) / SCAN_T;
    const int64_t b0 = t_lo + (int64_t)threadIdx.x * chunk;
    i128 acc = 0;
    for (int64_t b = b0; b < b0 + chunk && b < ntiles; b++) acc += a.tsum[b];
    i128 tot;
    i128 run = block_excl_scan<SCAN_T>(acc, lds, &tot);
    for (int64_t b = b0; b < b0 + chunk && b < ntiles; b++) {
        a.tpre[b] = run;
        run += a.tsum[b];
    }
}

// One round, in place: new iterate from exact prefix sums of the previous
// iterate's increments, verification of the recurrence, next increments.
__global__ __launch_bounds__(CT) void chain_round_kernel(ChainArgs a)
{
    __shared__ i128 lds[CT];
    __shared__ double lastv[CT];
    const int64_t L = *a.Lp;
    const int64_t fb = a.first_bad[0];
    if (fb >= L) return;
    const int64_t t_lo = fb / CTILE;
    const int64_t tile = t_lo + blockIdx.x;
    const int64_t base = tile * CTILE;
    if (base >= L) return;
    const double v0 = *a.v0p;
    const int t = threadIdx.x;
    const int64_t j0 = base + (int64_t)t * CI;
    // value before the restart tile is final
    const double vstart = t_lo == 0 ? v0 : a.bnd[t_lo - 1];
    const i128 Vstart = fxk(vstart, a.K);
    double xv[CI], gv[CI];
    for (int u = 0; u < CI; u++) {
        int64_t j = j0 + u;
        xv[u] = j < L ? a.x[j] : 0.0;
        gv[u] = j < L ? a.g[j] : 0.0;
    }
    lastv[t] = gv[CI - 1];
    __syncthreads();
    // predecessor of the first item under the previous iterate
    double prev = t > 0 ? lastv[t - 1] : (tile == 0 ? v0 : a.bnd[tile - 1]);
    i128 inc[CI];
    i128 acc = 0;
    for (int u = 0; u < CI; u++) {
        int64_t j = j0 + u;
        inc[u] = 0;
        if (j < L) {
            inc[u] = incr(j == 0 ? v0 : prev, xv[u], a.sign, a.K);
            acc += inc[u];
        }
        prev = gv[u];
    }
    i128 tot;
    i128 run = Vstart + a.tpre[tile] + block_excl_scan<CT>(acc, lds, &tot);
    double nv[CI];
    for (int u = 0; u < CI; u++) {
        run += inc[u];
        nv[u] = unfxk(run, a.K);
    }
    __syncthreads();
    lastv[t] = nv[CI - 1];
    __syncthreads();
    double before = t > 0 ? lastv[t - 1] : (tile == t_lo ? vstart : unfxk(Vstart + a.tpre[tile], a.K));
    i128 nacc = 0;
    int64_t bad = L;
    for (int u = 0; u < CI; u++) {
        int64_t j = j0 + u;
        if (j < L) {
            double p = j == 0 ? v0 : before;
            double want = p + a.sign * xv[u];
            if (want != nv[u] && j < bad) bad = j;
            nacc += incr(p, xv[u], a.sign, a.K);
            a.g[j] = nv[u];
        }
        before = nv[u];
    }
    if (bad < L) atomicMin((unsigned long long *)&a.first_bad[1], (unsigned long long)bad);
    block_excl_scan<CT>(nacc, lds, &tot);
    if (t == 0) a.tsum[tile] = tot;
}

// Bounded sequential completion from the first element that still violates the
// recurrence after the last round (normally none).
__global__ void chain_finish_kernel(ChainArgs a)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t L = *a.Lp;
    int64_t j = a.first_bad[0] >= L ? L : a.first_bad[1];
    if (j >= L) return;
    double v = j == 0 ? *a.v0p : a.g[j - 1];
    for (; j < L; j++) {
        v = v + a.sign * a.x[j];
        a.g[j] = v;
    }
    a.first_bad[0] = L;
}

// ---------------------------------------------------------------- segment and walk
struct EsaArgs {
    const double *x;
    double *g;
    double sign;
    const double *v0p;
    int64_t *first_bad;         // as in ChainArgs, initialised here
    i128 *tsumP, *tpreP;        // per tile: fixed-point sums of s*x and exclusive prefixes
    i128 *tsumR, *tpreR;        // per tile: regular-increment sums and prefixes
    i128 *tcnt, *toff;          // per tile: sensitive counts and prefixes
    i128 *tsum;                 // per tile: next Jacobi increments (for the repair rounds)
    int64_t *s_idx;             // sensitive list: element index
    i128 *s_racc;               //   regular-increment prefix before the element
    double *s_v;                //   exact chain value at the element
    int64_t *s_n;               // number of sensitive elements (uncapped)
    int64_t smax;               // list capacity
    int64_t *raw_n;             // sensitive elements appended by the classification, unordered:
    int64_t *raw_idx;           //   index
    i128 *raw_part;             //   regular-increment prefix inside its tile
    const int64_t *Lp;
    int K;
};

__device__ __forceinline__ int binade_of(double v)
{
    // exponent of a normal double, or INT_MIN for zero, subnormal, inf, nan
    uint64_t b = (uint64_t)__double_as_longlong(v);
    int e = (int)((b >> 52) & 0x7ff);
    return (e == 0 || e == 0x7ff) ? INT32_MIN : e - 1023;
}

// binade of every value within e of the estimate v, or INT32_MIN if they differ
__device__ __forceinline__ int interval_binade(double v, double e)
{
    double lo = v - e, hi = v + e;
    if (!(lo > 0.0) && !(hi < 0.0)) return INT32_MIN;
    int b1 = binade_of(lo), b2 = binade_of(hi);
    return b1 == b2 ? b1 : INT32_MIN;
}

// Bound on |v - estimate| after n additions: n roundings of at most half an ulp of
// values below 2 vmax (monotone chain), n truncations to 2^-K, and the conversion of
// the fixed-point estimate to double.
__device__ __forceinline__ double err_bound(int64_t n, double vmax, double a, int K)
{
    return ((double)n * (ldexp(vmax, -51) + ldexp(1.0, 1 - K)) + ldexp(a, -50)) * 1.001;
}

// Binade of the chain value before element j (after elements 0..j-1, fixed-point
// prefix Pp), or INT32_MIN when its error interval straddles binades; before
// element 0 the value is v0 exactly.
__device__ __forceinline__ int start_binade(int64_t j, double v0, i128 Pp, int K)
{
    if (j == 0) return interval_binade(v0, 0.0);
    const double p = unfxk(Pp, K);
    return interval_binade(p, err_bound(j, fmax(fabs(v0), fabs(p)), fabs(p), K));
}

// regular increment of element j in 2^-K units, or sensitive: bp, bq are the binades
// before and after the element (bq = start_binade(j + 1, ...), so a thread carries
// it on to its next element)
__device__ __forceinline__ bool classify(double x, double sign, int bp, int bq, int K, i128 *R)
{
    if (bp == INT32_MIN || bp != bq || bp < 52 - K || bp > 52 + 60 - K) return true;
    double y = ldexp(sign * x, 52 - bp);         // exact: power-of-two scaling
    double fy = floor(y);
    if (y - fy == 0.5 || !(fabs(y) <= 9007199254740992.0)) return true;
    double ry = rint(y);
    *R = (i128)(int64_t)ry * ((i128)1 << (bp - 52 + K));
    return false;
}

// Classification of one tile's items (CI consecutive per thread).  Needs tpreP.
struct TileItems {
    double xv[CI];
    i128 R[CI];
    uint32_t sens;      // bit u: item u sensitive (or past the end)
};

__device__ void tile_classify(const EsaArgs &a, int64_t tile, int64_t L, double v0, TileItems &it, i128 *lds)
{
    const int t = threadIdx.x;
    const int64_t j0 = tile * CTILE + (int64_t)t * CI;
    i128 fxs[CI];
    i128 acc = 0;
    for (int u = 0; u < CI; u++) {
        int64_t j = j0 + u;
        it.xv[u] = j < L ? a.x[j] : 0.0;
        fxs[u] = j < L ? fxk(a.sign * it.xv[u], a.K) : (i128)0;
        acc += fxs[u];
    }
    i128 tot;
    i128 P = fxk(v0, a.K) + a.tpreP[tile] + block_excl_scan<CT>(acc, lds, &tot);
    it.sens = 0;
    int bp = start_binade(j0, v0, P, a.K);
    for (int u = 0; u < CI; u++) {
        int64_t j = j0 + u;
        i128 Pq = P + fxs[u];
        it.R[u] = 0;
        bool sens = j >= L;
        if (!sens) {
            const int bq = start_binade(j + 1, v0, Pq, a.K);
            sens = classify(it.xv[u], a.sign, bp, bq, a.K, &it.R[u]);
            bp = bq;
        }
        if (sens) {
            it.sens |= 1u << u;
            it.R[u] = 0;
        }
        P = Pq;
    }
}

__global__ __launch_bounds__(CT) void esa_sum_kernel(EsaArgs a)
{
    __shared__ i128 lds[CT];
    const int64_t L = *a.Lp;
    if (blockIdx.x == 0 && threadIdx.x == 0) { a.first_bad[0] = 0; a.first_bad[1] = L; *a.raw_n = 0; }
    const int64_t base = (int64_t)blockIdx.x * CTILE;
    if (base >= L) return;
    i128 acc = 0;
    for (int u = 0; u < CI; u++) {
        int64_t j = base + (int64_t)threadIdx.x * CI + u;
        if (j < L) acc += fxk(a.sign * a.x[j], a.K);
    }
    i128 tot;
    block_excl_scan<CT>(acc, lds, &tot);
    if (threadIdx.x == 0) a.tsumP[blockIdx.x] = tot;
}

// exclusive scan of up to two per-tile int128 arrays (one block); the total of the
// second goes to *total
__global__ __launch_bounds__(SCAN_T) void esa_tile_scan_kernel(const i128 *in1, i128 *out1, const i128 *in2,
                                                               i128 *out2, int64_t *total, const int64_t *Lp)
{
    __shared__ i128 lds[SCAN_T];
    const int64_t L = *Lp;
    const int64_t ntiles = (L + CTILE - 1) / CTILE;
    const int64_t chunk = (ntiles + SCAN_T - 1) / SCAN_T;
    const int64_t b0 = (int64_t)threadIdx.x * chunk;
    for (int pass = 0; pass < (in2 ? 2 : 1); pass++) {
        const i128 *in = pass ? in2 : in1;
        i128 *out = pass ? out2 : out1;
        i128 acc = 0;
        for (int64_t b = b0; b < b0 + chunk && b < ntiles; b++) acc += in[b];
        i128 tot;
        i128 run = block_excl_scan<SCAN_T>(acc, lds, &tot);
        for (int64_t b = b0; b < b0 + chunk && b < ntiles; b++) {
            out[b] = run;
            run += in[b];
        }
        if (pass == 1 && threadIdx.x == 0) *total = (int64_t)tot;
    }
}

__global__ __launch_bounds__(CT) void esa_classify_kernel(EsaArgs a)
{
    __shared__ i128 lds[CT];
    const int64_t L = *a.Lp;
    const int64_t tile = blockIdx.x;
    if (tile * CTILE >= L) return;
    const double v0 = *a.v0p;
    TileItems it;
    tile_classify(a, tile, L, v0, it, lds);
    i128 racc = 0, cnt = 0;
    for (int u = 0; u < CI; u++) {
        racc += it.R[u];
        int64_t j = tile * CTILE + (int64_t)threadIdx.x * CI + u;
        if (j < L && ((it.sens >> u) & 1)) cnt += 1;
    }
    i128 tr, tc;
    i128 R = block_excl_scan<CT>(racc, lds, &tr);
    block_excl_scan<CT>(cnt, lds, &tc);
    if (threadIdx.x == 0) { a.tsumR[tile] = tr; a.tcnt[tile] = tc; }
    // the few sensitive elements, unordered, with their in-tile regular prefix
    for (int u = 0; u < CI; u++) {
        int64_t j = tile * CTILE + (int64_t)threadIdx.x * CI + u;
        if (j >= L) break;
        if ((it.sens >> u) & 1) {
            unsigned long long slot = atomicAdd((unsigned long long *)a.raw_n, 1ull);
            if (slot < (unsigned long long)RAW_CAP) { a.raw_idx[slot] = j; a.raw_part[slot] = R; }
        }
        R += it.R[u];
    }
}

// in-tile exclusive prefixes of the regular increments and sensitive counts
__device__ void tile_prefixes(const EsaArgs &a, int64_t tile, int64_t L, const TileItems &it, i128 *lds,
                              i128 *Racc, int64_t *kpos)
{
    i128 racc = 0, cnt = 0;
    const int64_t j0 = tile * CTILE + (int64_t)threadIdx.x * CI;
    for (int u = 0; u < CI; u++) {
        racc += it.R[u];
        if (j0 + u < L && ((it.sens >> u) & 1)) cnt += 1;
    }
    i128 tr, tc;
    *Racc = a.tpreR[tile] + block_excl_scan<CT>(racc, lds, &tr);
    *kpos = (int64_t)(a.toff[tile] + block_excl_scan<CT>(cnt, lds, &tc));
}

// Ordered compaction, used only when more than RAW_CAP elements are sensitive.
__global__ __launch_bounds__(CT) void esa_compact_kernel(EsaArgs a)
{
    __shared__ i128 lds[CT];
    if (*a.s_n <= RAW_CAP) return;
    const int64_t L = *a.Lp;
    const int64_t tile = blockIdx.x;
    if (tile * CTILE >= L) return;
    const double v0 = *a.v0p;
    TileItems it;
    tile_classify(a, tile, L, v0, it, lds);
    i128 R;
    int64_t k;
    tile_prefixes(a, tile, L, it, lds, &R, &k);
    const int64_t j0 = tile * CTILE + (int64_t)threadIdx.x * CI;
    for (int u = 0; u < CI; u++) {
        int64_t j = j0 + u;
        if (j >= L) break;
        if ((it.sens >> u) & 1) {
            if (k < a.smax) { a.s_idx[k] = j; a.s_racc[k] = R; }
            k++;
        }
        R += it.R[u];
    }
}

// One thread walks the sensitive elements in order; the block stages the list
// through LDS.  Between two sensitive elements the regular run stays in one binade,
// so v_prev + (run sum) is exact in double.
__global__ __launch_bounds__(256) void esa_walk_kernel(EsaArgs a)
{
    __shared__ i128 sr[256];
    __shared__ double sx[256], sv[256];
    __shared__ double Vd;
    __shared__ i128 Rl;
    const int t = threadIdx.x;
    int64_t ns = *a.s_n;
    if (ns <= RAW_CAP) {
        // rank sort of the classification's list by index (indices are distinct);
        // the regular prefix gains its tile's offset
        __shared__ int64_t ridx[RAW_CAP];
        for (int64_t e = t; e < ns; e += blockDim.x) ridx[e] = a.raw_idx[e];
        __syncthreads();
        for (int64_t e = t; e < ns; e += blockDim.x) {
            const int64_t j = ridx[e];
            int64_t rank = 0;
            for (int64_t q = 0; q < ns; q++) rank += ridx[q] < j;
            a.s_idx[rank] = j;
            a.s_racc[rank] = a.raw_part[e] + a.tpreR[j / CTILE];
        }
        __threadfence();
        __syncthreads();
    }
    if (ns > a.smax) ns = a.smax;
    if (t == 0) { Vd = *a.v0p; Rl = 0; }
    __syncthreads();
    for (int64_t c0 = 0; c0 < ns; c0 += 256) {
        int64_t k = c0 + t;
        if (k < ns) { sr[t] = a.s_racc[k]; sx[t] = a.x[a.s_idx[k]]; }
        __syncthreads();
        if (t == 0) {
            double vd = Vd;
            i128 rl = Rl;
            int cnt = (int)(ns - c0 < 256 ? ns - c0 : 256);
            for (int u = 0; u < cnt; u++) {
                i128 ra = sr[u];
                double pred = ra == rl ? vd : vd + unfxk(ra - rl, a.K);
                vd = pred + a.sign * sx[u];
                rl = ra;
                sv[u] = vd;
            }
            Vd = vd;
            Rl = rl;
        }
        __syncthreads();
        if (k < ns) a.s_v[k] = sv[t];
        __syncthreads();
    }
}

// value of an element from the last sensitive element kk at or before it
__device__ __forceinline__ double esa_value(const EsaArgs &a, int64_t kk, int64_t ns_eff, double v0, i128 Rincl)
{
    if (kk >= ns_eff) kk = ns_eff - 1;
    if (kk < 0) return v0 + unfxk(Rincl, a.K);
    return a.s_v[kk] + unfxk(Rincl - a.s_racc[kk], a.K);
}

// Parallel rebuild of every element, verification of the recurrence, and the
// tile sums of the next Jacobi increments.
__global__ __launch_bounds__(CT) void esa_rebuild_kernel(EsaArgs a)
{
    __shared__ i128 lds[CT];
    __shared__ double lastv[CT];
    const int64_t L = *a.Lp;
    const int64_t tile = blockIdx.x;
    const int64_t base = tile * CTILE;
    if (base >= L) return;
    const double v0 = *a.v0p;
    const int t = threadIdx.x;
    int64_t ns_eff = *a.s_n;
    if (ns_eff > a.smax) ns_eff = a.smax;
    TileItems it;
    tile_classify(a, tile, L, v0, it, lds);
    i128 R;
    int64_t k;
    tile_prefixes(a, tile, L, it, lds, &R, &k);
    const int64_t j0 = base + (int64_t)t * CI;
    double nv[CI];
    for (int u = 0; u < CI; u++) {
        int64_t j = j0 + u;
        nv[u] = 0.0;
        if (j >= L) continue;
        bool sens = (it.sens >> u) & 1;
        R += it.R[u];
        if (sens && k < ns_eff) nv[u] = a.s_v[k];
        else nv[u] = esa_value(a, sens ? k : k - 1, ns_eff, v0, R);
        if (sens) k++;
    }
    lastv[t] = nv[CI - 1];
    __syncthreads();
    double before;
    if (t > 0) before = lastv[t - 1];
    else if (tile == 0) before = v0;
    else {
        int64_t kk = (int64_t)a.toff[tile] - 1;
        if (kk >= 0 && kk < ns_eff && a.s_idx[kk] == base - 1) before = a.s_v[kk];
        else before = esa_value(a, kk, ns_eff, v0, a.tpreR[tile]);
    }
    int64_t bad = L;
    for (int u = 0; u < CI; u++) {
        int64_t j = j0 + u;
        if (j < L) {
            double want = before + a.sign * it.xv[u];
            if (want != nv[u] && j < bad) bad = j;
            a.g[j] = nv[u];
        }
        before = nv[u];
    }
    if (bad < L) atomicMin((unsigned long long *)&a.first_bad[1], (unsigned long long)bad);
}

// Only when the rebuild found a violation: per-tile sums of the Jacobi increments
// fl(g_{j-1} + s x_j) - g_{j-1} of the rebuilt chain, which the repair rounds scan.
__global__ __launch_bounds__(CT) void chain_seed_from_g_kernel(ChainArgs a)
{
    __shared__ i128 lds[CT];
    const int64_t L = *a.Lp;
    if (a.first_bad[1] >= L) return;
    const int64_t base = (int64_t)blockIdx.x * CTILE;
    if (base >= L) return;
    const double v0 = *a.v0p;
    i128 acc = 0;
    for (int u = 0; u < CI; u++) {
        int64_t j = base + (int64_t)threadIdx.x * CI + u;
        if (j < L) acc += incr(j == 0 ? v0 : a.g[j - 1], a.x[j], a.sign, a.K);
    }
    i128 tot;
    block_excl_scan<CT>(acc, lds, &tot);
    if (threadIdx.x == 0) a.tsum[blockIdx.x] = tot;
}

// ---------------------------------------------------------------- the three chains
__global__ void pot_init_kernel(DevScalars *s) { s->first_inf = s->n_max + 1; }

// cmc_utils.c:685-696: the count stops at the first k with !(r[k] < SF_INFINITY).
__global__ void pot_count_kernel(const double *__restrict__ r, DevScalars *s)
{
    int64_t k = 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > s->n_max) return;
    if (!(r[k] < CMCG_SF_INFINITY)) atomicMin((unsigned long long *)&s->first_inf, (unsigned long long)k);
}

struct PotWork {
    double *xa, *ga, *gb;      // chain buffers (x, iterate, scratch iterate)
    double *Mk;                // chain B result: M_{N-1-j}
    double *scal;              // [0] 0.0  [1] Mtotal  [2] chain lengths as doubles (unused)
    int64_t *len;              // [0] L_A  [1] L_B  [2] L_C
    int64_t *fbad;             // 2 per chain
    i128 *tsum, *tpre;
    double *bnd;
};

// N_MAX from the count, the chain lengths, and the sentinels before the recurrence
// reads r[N+1] (cmc_utils.c:713-714)
__global__ void pot_setup_kernel(DevScalars *s, PotWork w, double *__restrict__ r, double *__restrict__ phi)
{
    int64_t n = s->first_inf - 1;
    s->n_max = n;
    w.len[0] = n;                       // A: k = 1..N
    w.len[1] = n > 1 ? n - 1 : 0;       // B: k = N..2 gives M_{N-1}..M_1
    w.len[2] = n;                       // C: k = N..1
    w.scal[0] = 0.0;
    r[n + 1] = CMCG_SF_INFINITY;
    phi[n + 1] = 0.0;
}

// after chain A: Mtotal, Rtidal (cmc_utils.c:703-707), NaN check (:688-691)
__device__ __forceinline__ void pot_total(DevScalars *s, const KParams &P, const double *__restrict__ ga, PotWork &w)
{
    int64_t n = s->n_max;
    double mprev = n > 0 ? ga[n - 1] : 0.0;
    if (isnan(mprev)) dev_set_error(s, CMCG_E_NAN, n);
    double Mtotal = mprev * P.madhoc + P.cenma_m * P.madhoc;
    s->Mtotal = Mtotal;
    s->Rtidal = P.orbit_r * pow(Mtotal, 1.0 / 3.0);
    w.scal[1] = Mtotal;
}

// chain B increments (star_m[k] / clus.N_STAR, k = N - j); its first thread also
// finishes chain A (Mtotal, the start value of chain B), which the increments do not use
__global__ void prep_b_kernel(const double *__restrict__ m, DevScalars *s, double nstar, double *__restrict__ x,
                              KParams P, const double *__restrict__ ga, PotWork w)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) pot_total(s, P, ga, w);
    int64_t n = s->n_max;
    if (j < n - 1) x[j] = m[n - j] / nstar;
}

// p_k = M_k * (1/r_k - 1/r_{k+1}) with the reference's roundings; chain C runs on
// w = -phi, w_k = fl(w_{k+1} + p_k)
__global__ void prep_c_kernel(const double *__restrict__ r, DevScalars *s, const double *__restrict__ Mk,
                              double *__restrict__ x)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n = s->n_max;
    if (j >= n) return;
    int64_t k = n - j;
    double M = k == n ? s->Mtotal : Mk[n - 1 - k];
    x[j] = M * (1.0 / r[k] - 1.0 / r[k + 1]);
}

__global__ void phi_out_kernel(const double *__restrict__ r, double *__restrict__ phi, DevScalars *s, KParams P,
                               const double *__restrict__ g)
{
    int64_t k = 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n = s->n_max;
    if (n == 0 && k == 1) phi[0] = phi[1] + P.cenma_m * P.madhoc / r[1];    // no stars: phi[1] is the sentinel
    if (k > n) return;
    double p = -g[n - k];
    phi[k] = p;
    if (isnan(p)) dev_set_error(s, CMCG_E_NAN, k);
    if (k == 1) phi[0] = p + P.cenma_m * P.madhoc / r[1];    // cmc_utils.c:729
}

__global__ void pot_at_kernel(const double *__restrict__ r, const double *__restrict__ phi, DevScalars *s,
                              const double *__restrict__ x, double *__restrict__ out, int64_t count)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int64_t n = s->n_max;
    bool bad = false;
    double v = potential_at(r, phi, n, x[i], 1, n + 1, &bad);
    if (bad) { dev_set_error(s, CMCG_E_BRACKET, i); v = 0.0; }
    out[i] = v;
}

hipError_t run_chain(ChainArgs a, EsaArgs e, int rounds, int64_t n_ub, hipStream_t st)
{
    int64_t ntiles = (n_ub + CTILE - 1) / CTILE;
    if (ntiles < 1) ntiles = 1;
    dim3 gt((unsigned)ntiles), bt(CT);
    hipLaunchKernelGGL(esa_sum_kernel, gt, bt, 0, st, e);
    hipLaunchKernelGGL(esa_tile_scan_kernel, dim3(1), dim3(SCAN_T), 0, st, (const i128 *)e.tsumP, e.tpreP,
                       (const i128 *)nullptr, (i128 *)nullptr, (int64_t *)nullptr, e.Lp);
    hipLaunchKernelGGL(esa_classify_kernel, gt, bt, 0, st, e);
    hipLaunchKernelGGL(esa_tile_scan_kernel, dim3(1), dim3(SCAN_T), 0, st, (const i128 *)e.tsumR, e.tpreR,
                       (const i128 *)e.tcnt, e.toff, e.s_n, e.Lp);
    hipLaunchKernelGGL(esa_compact_kernel, gt, bt, 0, st, e);
    hipLaunchKernelGGL(esa_walk_kernel, dim3(1), dim3(256), 0, st, e);
    hipLaunchKernelGGL(esa_rebuild_kernel, gt, bt, 0, st, e);
    hipLaunchKernelGGL(chain_seed_from_g_kernel, gt, bt, 0, st, a);
    for (int it = 0; it < rounds; it++) {
        hipLaunchKernelGGL(chain_scan_kernel, dim3(1), dim3(SCAN_T), 0, st, a);
        hipLaunchKernelGGL(chain_round_kernel, gt, bt, 0, st, a);
    }
    hipLaunchKernelGGL(chain_finish_kernel, dim3(1), dim3(1), 0, st, a);
    return hipGetLastError();
}
} // namespace

// ---------------------------------------------------------------- comp_mass_percent
// cmc_utils.c:872-982 for one rank.  Chain M: mprev_k = mprev_{k-1} + m_k / N_STAR
// (k = 1..N_MAX, from cenma.m madhoc or 0); chains T_r, T_t, V_r, V_t: the running
// sums of 0.5 m madhoc vr vr, 0.5 m madhoc vt vt, vr vr, vt vt (k = 1..N_MAX_NEW).
// Each is the reference's sequential recurrence, solved by the same exact emulation
// as the potential's chains.
constexpr int MPC_MAX = 256;
struct MassPcOut {
    double v[8][MPC_MAX];        // mass_r, ave_mass_r, densities_r, ke_rad, ke_tan, v2_rad, v2_tan, (unused)
    int64_t k[MPC_MAX];          // no_star_r (the index k)
    int64_t reached;             // fractions assigned
};

__global__ void mpc_prep_m_kernel(const double *__restrict__ m, const DevScalars *s, double nstar,
                                  double *__restrict__ x, int64_t *__restrict__ len)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = s->n_max;
    if (j == 0) len[3] = n;
    if (j < n) x[j] = m[j + 1] / nstar;                   // star_m[k] / clus.N_STAR, k = j + 1
}

// term 0: 0.5 m madhoc vr vr, 1: 0.5 m madhoc vt vt, 2: vr vr, 3: vt vt (left to right)
__global__ void mpc_prep_v_kernel(const DevFields f, const DevScalars *s, double madhoc, int term,
                                  double *__restrict__ x, int64_t *__restrict__ len)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = s->n_max_new;
    if (j == 0) len[3] = n;
    if (j >= n) return;
    const int64_t k = j + 1;
    double v;
    switch (term) {
    case 0: v = 0.5 * f.m[k] * madhoc * f.vr[k] * f.vr[k]; break;
    case 1: v = 0.5 * f.m[k] * madhoc * f.vt[k] * f.vt[k]; break;
    case 2: v = f.vr[k] * f.vr[k]; break;
    default: v = f.vt[k] * f.vt[k]; break;
    }
    x[j] = v;
}

// x^3 from x*x exactly as a double-double, times x: within 2^-50 ulp of the correctly
// rounded cube (pow(x, 3) of a correctly rounding libm)
__device__ __forceinline__ double cube_rn(double x)
{
    const double p = x * x, e = __builtin_fma(x, x, -p);
    const double q = p * x, f1 = __builtin_fma(p, x, -q);
    return q + (f1 + e * x);
}

// the fractions' indices (one fraction per star at most, as the reference's loop),
// from the enclosed-mass chain g[k-1] = mprev after star k
__global__ void mpc_assign_kernel(const double *__restrict__ r, const DevScalars *s, const double *__restrict__ g,
                                  const double *__restrict__ pc, int count, int bh_include, double mprev0,
                                  double itm, double minimum_r, MassPcOut *o)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double PI = 3.14159265358979323;                  // include/cmc/cmc.h:62
    const double Mtotal = s->Mtotal;
    const int64_t n = s->n_max;
    int mc = 0;
    if (bh_include)
        for (; mc < count && mprev0 / Mtotal > pc[mc]; mc++) {
            o->v[0][mc] = minimum_r;
            for (int q = 1; q < 7; q++) o->v[q][mc] = 0.0;
            o->k[mc] = 0;
        }
    int64_t kprev = 0;
    for (; mc < count; mc++) {
        const double thr = pc[mc] * Mtotal;
        // smallest k in (kprev, n] with g[k-1] > thr; g is non-decreasing
        int64_t lo = kprev + 1, hi = n + 1;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (g[mid - 1] > thr) hi = mid;
            else lo = mid + 1;
        }
        if (lo > n) break;
        const int64_t k = lo;
        const double mprev = g[k - 1];
        o->v[0][mc] = r[k];
        o->v[1][mc] = mprev / Mtotal / k * itm;
        o->v[2][mc] = mprev / (4.0 / 3.0 * PI * cube_rn(r[k]));
        o->k[mc] = k;
        kprev = k;
    }
    o->reached = mc;
}

__global__ void mpc_init_kernel(double *v0, double mprev0)
{
    v0[0] = mprev0;
    v0[1] = 0.0;
}

__global__ void mpc_take_kernel(const double *__restrict__ g, MassPcOut *o, int term, int count)
{
    const int i = threadIdx.x;
    if (i >= count || i >= o->reached) return;
    const int64_t k = o->k[i];
    o->v[3 + term][i] = k > 0 ? g[k - 1] : 0.0;
}

hipError_t launch_potential(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub,
                            const Scratch &sc, hipStream_t st)
{
    int64_t ntiles = (n_ub + CTILE - 1) / CTILE + 1;
    if (ntiles > sc.max_tiles) return hipErrorInvalidValue;
    PotWork w;
    w.xa = sc.chain_x;
    w.ga = sc.chain_g;
    w.gb = nullptr;
    w.Mk = sc.chain_mk;
    w.scal = sc.chain_scal;
    w.len = sc.chain_len;
    w.fbad = sc.chain_fbad;
    // per tile: tsum tpre tsumP tpreP tsumR tpreR tcnt toff (int128) and bnd (double)
    i128 *tb = (i128 *)sc.tile_sums;
    const int64_t T = sc.max_tiles;
    w.tsum = tb;
    w.tpre = tb + T;
    w.bnd = (double *)(tb + 8 * T);
    EsaArgs e{};
    e.tsumP = tb + 2 * T; e.tpreP = tb + 3 * T; e.tsumR = tb + 4 * T; e.tpreR = tb + 5 * T;
    e.tcnt = tb + 6 * T; e.toff = tb + 7 * T;
    e.tsum = w.tsum;
    e.s_idx = sc.esa_idx;
    e.s_racc = (i128 *)sc.esa_racc;
    e.s_v = sc.esa_v;
    e.s_n = sc.esa_n;
    e.smax = sc.esa_max;
    e.raw_n = sc.esa_n + 1;
    e.raw_idx = sc.esa_raw_idx;
    e.raw_part = (i128 *)sc.esa_raw_part;
    unsigned g1 = (unsigned)((n_ub + 256) / 256);
    hipLaunchKernelGGL(pot_init_kernel, dim3(1), dim3(1), 0, st, s);
    if (n_ub > 0) hipLaunchKernelGGL(pot_count_kernel, dim3(g1), dim3(256), 0, st, f.r, s);
    hipLaunchKernelGGL(pot_setup_kernel, dim3(1), dim3(1), 0, st, s, w, f.r, f.phi);
    // chain A: total mass, straight over star_m[1..N]
    // fixed-point scales 2^-70 (stored masses are O(1)) and 2^-100; a test hook may
    // coarsen them, which sends most elements through the fallback paths
    const int KA = 70 - P.chain_k_drop < 8 ? 8 : 70 - P.chain_k_drop;
    const int KB = 100 - P.chain_k_drop < 8 ? 8 : 100 - P.chain_k_drop;
    ChainArgs a{f.m + 1, w.ga, 1.0, w.scal + 0, w.fbad + 0, w.tsum, w.tpre, w.bnd, w.len + 0, KA};
    EsaArgs ae = e;
    ae.x = f.m + 1; ae.g = w.ga; ae.sign = 1.0; ae.v0p = w.scal + 0; ae.first_bad = w.fbad + 0; ae.Lp = w.len + 0; ae.K = KA;
    if (hipError_t err = run_chain(a, ae, REPAIR_ROUNDS, n_ub, st)) return err;
    // chain B: enclosed mass, iterated directly in Mk (chain C's input)
    hipLaunchKernelGGL(prep_b_kernel, dim3(g1), dim3(256), 0, st, f.m, s, (double)P.n_star, w.xa, P,
                       (const double *)w.ga, w);
    ChainArgs b{w.xa, w.Mk, -1.0, w.scal + 1, w.fbad + 2, w.tsum, w.tpre, w.bnd, w.len + 1, KB};
    EsaArgs be = e;
    be.x = w.xa; be.g = w.Mk; be.sign = -1.0; be.v0p = w.scal + 1; be.first_bad = w.fbad + 2; be.Lp = w.len + 1; be.K = KB;
    if (hipError_t err = run_chain(b, be, REPAIR_ROUNDS, n_ub, st)) return err;
    // chain C: w = -phi
    hipLaunchKernelGGL(prep_c_kernel, dim3(g1), dim3(256), 0, st, f.r, s, (const double *)w.Mk, w.xa);
    ChainArgs c{w.xa, w.ga, 1.0, w.scal + 0, w.fbad + 4, w.tsum, w.tpre, w.bnd, w.len + 2, KB};
    EsaArgs ce = e;
    ce.x = w.xa; ce.g = w.ga; ce.sign = 1.0; ce.v0p = w.scal + 0; ce.first_bad = w.fbad + 4; ce.Lp = w.len + 2; ce.K = KB;
    if (hipError_t err = run_chain(c, ce, REPAIR_ROUNDS, n_ub, st)) return err;
    hipLaunchKernelGGL(phi_out_kernel, dim3(g1), dim3(256), 0, st, f.r, f.phi, s, P, (const double *)w.ga);
    return hipGetLastError();
}

hipError_t launch_potential_at(const DevFields &f, DevScalars *s, const KParams &P, const double *x,
                               double *out, int64_t count, hipStream_t st)
{
    (void)P;
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(pot_at_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, f.r, f.phi, s, x,
                       out, count);
    return hipGetLastError();
}

// comp_mass_percent: pc (device, count <= MPC_MAX, ascending), results into out
hipError_t launch_mass_percent(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub, const Scratch &sc,
                               const double *pc, int count, int bh_include, double itm, double minimum_r,
                               void *out, hipStream_t st)
{
    if (count < 0 || count > MPC_MAX) return hipErrorInvalidValue;
    int64_t ntiles = (n_ub + CTILE - 1) / CTILE + 1;
    if (ntiles > sc.max_tiles) return hipErrorInvalidValue;
    i128 *tb = (i128 *)sc.tile_sums;
    const int64_t T = sc.max_tiles;
    EsaArgs e{};
    e.tsumP = tb + 2 * T; e.tpreP = tb + 3 * T; e.tsumR = tb + 4 * T; e.tpreR = tb + 5 * T;
    e.tcnt = tb + 6 * T; e.toff = tb + 7 * T;
    e.tsum = tb;
    e.s_idx = sc.esa_idx;
    e.s_racc = (i128 *)sc.esa_racc;
    e.s_v = sc.esa_v;
    e.s_n = sc.esa_n;
    e.smax = sc.esa_max;
    e.raw_n = sc.esa_n + 1;
    e.raw_idx = sc.esa_raw_idx;
    e.raw_part = (i128 *)sc.esa_raw_part;
    e.x = sc.chain_x; e.g = sc.chain_g; e.sign = 1.0; e.first_bad = sc.chain_fbad + 6;
    e.Lp = sc.chain_len + 3;
    double *v0 = sc.chain_scal + 2;              // start values: [2] mprev0, [3] 0
    const double mprev0 = bh_include ? P.cenma_m * P.madhoc : 0.0;   // cmc_utils.c:878
    hipLaunchKernelGGL(mpc_init_kernel, dim3(1), dim3(1), 0, st, v0, mprev0);
    MassPcOut *o = (MassPcOut *)out;
    unsigned g1 = (unsigned)((n_ub + 256) / 256);
    const int KD = P.chain_k_drop;
    auto Kof = [&](int k) { return k - KD < 8 ? 8 : k - KD; };
    // chain M
    hipLaunchKernelGGL(mpc_prep_m_kernel, dim3(g1), dim3(256), 0, st, f.m, (const DevScalars *)s, (double)P.n_star,
                       sc.chain_x, sc.chain_len);
    EsaArgs em = e;
    em.v0p = v0; em.K = Kof(100);
    ChainArgs cm{sc.chain_x, sc.chain_g, 1.0, v0, e.first_bad, tb, tb + T, (double *)(tb + 8 * T), e.Lp, em.K};
    if (hipError_t err = run_chain(cm, em, REPAIR_ROUNDS, n_ub, st)) return err;
    hipLaunchKernelGGL(mpc_assign_kernel, dim3(1), dim3(1), 0, st, (const double *)f.r, (const DevScalars *)s,
                       (const double *)sc.chain_g, pc, count, bh_include, mprev0, itm, minimum_r, o);
    // the four running sums
    for (int term = 0; term < 4; term++) {
        hipLaunchKernelGGL(mpc_prep_v_kernel, dim3(g1), dim3(256), 0, st, f, (const DevScalars *)s, P.madhoc, term,
                           sc.chain_x, sc.chain_len);
        EsaArgs ev = e;
        ev.v0p = v0 + 1; ev.K = Kof(term < 2 ? 100 : 80);
        ChainArgs cv{sc.chain_x, sc.chain_g, 1.0, v0 + 1, e.first_bad, tb, tb + T, (double *)(tb + 8 * T), e.Lp, ev.K};
        if (hipError_t err = run_chain(cv, ev, REPAIR_ROUNDS, n_ub, st)) return err;
        hipLaunchKernelGGL(mpc_take_kernel, dim3(1), dim3(MPC_MAX), 0, st, (const double *)sc.chain_g, o, term, count);
    }
    return hipGetLastError();
}

size_t mass_percent_out_bytes() { return sizeof(MassPcOut); }

// ---------------------------------------------------------------- central_calculate
// cmc_utils.c:1239-1345 (J = 6) and clusdyn_calculate cmc_utils.c:1614-1624.  The
// half-mass index nave comes from the enclosed-mass chain of comp_mass_percent
// (star_m[k] / N_STAR, from 0); the six density-weighted sums over i = 1..nave are
// sequential recurrences solved by the exact chain emulation.  Each sum's terms are
// first scaled by a power of two that brings nave * max(term) to at most 1: binary
// rounding is invariant under such a scaling, so the scaled chain rounds exactly as
// the unscaled one and fits the chain's fixed-point range.
constexpr int CEN_TERMS = 6;     // rhoj, rhoj^2, rhoj v^2, rhoj r, (rhoj r)^2, rhoj m madhoc
enum { CEN_OK = 0, CEN_E_STATE = 1, CEN_E_RANGE = 2 };

struct CentralOut {
    int64_t nave;
    int32_t error, pad;
    double rh, rhojsum, rhoj2sum, rho, v_rms, rc, rc_nb, m_ave, n, rc_spitzer, N_core, N_core_nb, Trc;
    double S[CEN_TERMS];
    unsigned long long mx[CEN_TERMS];     // bit patterns of the terms' maxima (non-negative doubles)
    double scale[CEN_TERMS], unscale[CEN_TERMS];
};

// rhoj of cmc_utils.c:1263-1274 for star i: the 5 masses between jmin and jmax = jmin + 6
// over the shell volume between r[jmin] and r[jmax]
__device__ __forceinline__ double cen_rhoj(const double *__restrict__ r, const double *__restrict__ m,
                                           double madhoc, int64_t i)
{
    const double PI = 3.14159265358979323;                  // include/cmc/cmc.h:62
    const int64_t jmin = i - 3 > 1 ? i - 3 : 1, jmax = jmin + 6;
    double mrho = 0.0;
#pragma unroll
    for (int q = 1; q <= 5; q++) mrho += m[jmin + q] * madhoc;
    const double a = r[jmax], b = r[jmin];
    const double Vrj = 4.0 / 3.0 * PI * (a * a * a - b * b * b);
    return mrho / Vrj;
}

__device__ __forceinline__ double cen_term(const DevFields &f, double madhoc, int64_t i, int t)
{
    const double rj = cen_rhoj(f.r, f.m, madhoc, i);
    switch (t) {
    case 0: return rj;
    case 1: return rj * rj;
    case 2: return rj * (f.vr[i] * f.vr[i] + f.vt[i] * f.vt[i]);
    case 3: return rj * f.r[i];
    case 4: { const double q = rj * f.r[i]; return q * q; }
    default: return rj * f.m[i] * madhoc;
    }
}

__global__ void cen_init_kernel(CentralOut *o, double *v0)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    o->nave = 0;
    o->error = CEN_OK;
    for (int t = 0; t < CEN_TERMS; t++) {
        o->mx[t] = 0ull;
        o->S[t] = 0.0;
        o->scale[t] = o->unscale[t] = 1.0;
    }
    v0[0] = 0.0;
    v0[1] = 0.0;
}

// nave and rh from the chain g[k-1] = m after star k (cmc_utils.c:1247-1257, 1620-1623);
// the chain length for the sums becomes nave
__global__ void cen_nave_kernel(const double *__restrict__ r, const DevScalars *s, const double *__restrict__ g,
                                int64_t n_star, CentralOut *o, int64_t *len)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t n = s->n_max;
    const double thr = 0.5 * s->Mtotal;
    int64_t lo = 1, hi = n + 1;                 // smallest k with !(g[k-1] < thr); g non-decreasing
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (!(g[mid - 1] < thr)) hi = mid;
        else lo = mid + 1;
    }
    const int64_t nave = lo + 1;
    const int64_t jmax_last = (nave - 3 > 1 ? nave - 3 : 1) + 6;
    if (lo > n || n_star <= 12 || nave >= n_star - 6 || jmax_last > n + 1) {
        o->error = CEN_E_STATE;
        o->nave = 0;
        len[3] = 1;                             // the sums then run over one zero term
        return;
    }
    o->nave = nave;
    o->rh = r[nave];
    len[3] = nave;
}

__global__ __launch_bounds__(256) void cen_max_kernel(const DevFields f, double madhoc, CentralOut *o)
{
    __shared__ unsigned long long red[CEN_TERMS][256 / 64];
    const int64_t nave = o->nave;
    unsigned long long mx[CEN_TERMS] = {0, 0, 0, 0, 0, 0};
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nave; j += (int64_t)gridDim.x * blockDim.x)
        for (int t = 0; t < CEN_TERMS; t++) {
            const unsigned long long b = (unsigned long long)__double_as_longlong(cen_term(f, madhoc, j + 1, t));
            mx[t] = b > mx[t] ? b : mx[t];      // a negative or NaN term has the largest patterns
        }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int t = 0; t < CEN_TERMS; t++) {
        unsigned long long v = mx[t];
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long u = __shfl_down(v, off, 64);
            v = u > v ? u : v;
        }
        if (lane == 0) red[t][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < CEN_TERMS) {
        unsigned long long v = 0;
        for (int q = 0; q < 256 / 64; q++) v = red[threadIdx.x][q] > v ? red[threadIdx.x][q] : v;
        if (v) atomicMax(&o->mx[threadIdx.x], v);
    }
}

__global__ void cen_scale_kernel(CentralOut *o)
{
    if (threadIdx.x != 0 || blockIdx.x != 0 || o->error) return;
    for (int t = 0; t < CEN_TERMS; t++) {
        if (o->mx[t] >= 0x7ff0000000000000ull) {   // infinite, NaN or negative
            o->error = CEN_E_RANGE;
            return;
        }
        const double bound = (double)o->nave * __longlong_as_double((long long)o->mx[t]);
        if (bound > 0.0) {
            int e;
            frexp(bound, &e);                      // bound <= 2^e
            o->scale[t] = ldexp(1.0, -e);
            o->unscale[t] = ldexp(1.0, e);
        }
    }
}

__global__ void cen_prep_kernel(const DevFields f, double madhoc, int t, CentralOut *o, double *__restrict__ x)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nave = o->nave;
    if (o->error) {
        if (j == 0) x[0] = 0.0;
        return;
    }
    if (j >= nave) return;
    const double v = cen_term(f, madhoc, j + 1, t) * o->scale[t];
    if (v != 0.0 && v < 2.2250738585072014e-308) o->error = CEN_E_RANGE;   // scaled into subnormals
    x[j] = v;
}

__global__ void cen_take_kernel(const double *__restrict__ g, CentralOut *o, int t)
{
    if (threadIdx.x != 0 || blockIdx.x != 0 || o->error) return;
    o->S[t] = g[o->nave - 1] * o->unscale[t];
}

// the averages and derived quantities, cmc_utils.c:1323-1345
__global__ void cen_final_kernel(CentralOut *o)
{
    if (threadIdx.x != 0 || blockIdx.x != 0 || o->error) return;
    const double PI = 3.14159265358979323;
    const double rhojsum = o->S[0], rhoj2sum = o->S[1];
    double rho = o->S[1], v_rms = o->S[2], rc = o->S[3], rc_nb = o->S[4], m_ave = o->S[5];
    rho /= rhojsum;
    rho *= 4.0 / 5.0;
    rho *= 2.0;
    v_rms /= rhojsum;
    v_rms = sqrt(v_rms);
    rc /= rhojsum;
    rc_nb = sqrt(rc_nb / rhoj2sum);
    m_ave /= rhojsum;
    const double n = rho / m_ave;
    o->rc_spitzer = sqrt(3.0 * (v_rms * v_rms) / (4.0 * PI * rho));
    o->N_core = 4.0 / 3.0 * PI * (rc * rc * rc) * (n / 2.0);
    o->N_core_nb = 4.0 / 3.0 * PI * (rc_nb * rc_nb * rc_nb) * (n / 2.0);
    o->Trc = 0.065 * (v_rms * v_rms * v_rms) / (rho * m_ave);
    o->rhojsum = rhojsum;
    o->rhoj2sum = rhoj2sum;
    o->rho = rho;
    o->v_rms = v_rms;
    o->rc = rc;
    o->rc_nb = rc_nb;
    o->m_ave = m_ave;
    o->n = n;
}

hipError_t launch_central(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub, const Scratch &sc,
                          void *out, hipStream_t st)
{
    int64_t ntiles = (n_ub + CTILE - 1) / CTILE + 1;
    if (ntiles > sc.max_tiles) return hipErrorInvalidValue;
    i128 *tb = (i128 *)sc.tile_sums;
    const int64_t T = sc.max_tiles;
    EsaArgs e{};
    e.tsumP = tb + 2 * T; e.tpreP = tb + 3 * T; e.tsumR = tb + 4 * T; e.tpreR = tb + 5 * T;
    e.tcnt = tb + 6 * T; e.toff = tb + 7 * T;
    e.tsum = tb;
    e.s_idx = sc.esa_idx;
    e.s_racc = (i128 *)sc.esa_racc;
    e.s_v = sc.esa_v;
    e.s_n = sc.esa_n;
    e.smax = sc.esa_max;
    e.raw_n = sc.esa_n + 1;
    e.raw_idx = sc.esa_raw_idx;
    e.raw_part = (i128 *)sc.esa_raw_part;
    e.x = sc.chain_x; e.g = sc.chain_g; e.sign = 1.0; e.first_bad = sc.chain_fbad + 6;
    e.Lp = sc.chain_len + 3;
    double *v0 = sc.chain_scal + 2;              // both start values 0
    CentralOut *o = (CentralOut *)out;
    const unsigned g1 = (unsigned)((n_ub + 256) / 256);
    const int KD = P.chain_k_drop;
    const int K = 100 - KD < 8 ? 8 : 100 - KD;
    hipLaunchKernelGGL(cen_init_kernel, dim3(1), dim3(1), 0, st, o, v0);
    // enclosed mass, star_m[k] / N_STAR from 0
    hipLaunchKernelGGL(mpc_prep_m_kernel, dim3(g1), dim3(256), 0, st, f.m, (const DevScalars *)s, (double)P.n_star,
                       sc.chain_x, sc.chain_len);
    EsaArgs em = e;
    em.v0p = v0; em.K = K;
    ChainArgs cmass{sc.chain_x, sc.chain_g, 1.0, v0, e.first_bad, tb, tb + T, (double *)(tb + 8 * T), e.Lp, K};
    if (hipError_t err = run_chain(cmass, em, REPAIR_ROUNDS, n_ub, st)) return err;
    hipLaunchKernelGGL(cen_nave_kernel, dim3(1), dim3(1), 0, st, (const double *)f.r, (const DevScalars *)s,
                       (const double *)sc.chain_g, (int64_t)P.n_star, o, sc.chain_len);
    const unsigned gmax = g1 < 1024u ? g1 : 1024u;
    hipLaunchKernelGGL(cen_max_kernel, dim3(gmax), dim3(256), 0, st, f, P.madhoc, o);
    hipLaunchKernelGGL(cen_scale_kernel, dim3(1), dim3(1), 0, st, o);
    for (int t = 0; t < CEN_TERMS; t++) {
        hipLaunchKernelGGL(cen_prep_kernel, dim3(g1), dim3(256), 0, st, f, P.madhoc, t, o, sc.chain_x);
        EsaArgs es = e;
        es.v0p = v0 + 1; es.K = K;
        ChainArgs cs{sc.chain_x, sc.chain_g, 1.0, v0 + 1, e.first_bad, tb, tb + T, (double *)(tb + 8 * T), e.Lp, K};
        if (hipError_t err = run_chain(cs, es, REPAIR_ROUNDS, n_ub, st)) return err;
        hipLaunchKernelGGL(cen_take_kernel, dim3(1), dim3(1), 0, st, (const double *)sc.chain_g, o, t);
    }
    hipLaunchKernelGGL(cen_final_kernel, dim3(1), dim3(1), 0, st, o);
    return hipGetLastError();
}

size_t central_out_bytes() { return sizeof(CentralOut); }
