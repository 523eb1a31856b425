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
//   on v_{j-1}.  Such "regular" elements are recognised from a double estimate of
//   the prefix sums of the x and a rigorous bound on its error and on the chain's
//   own accumulated rounding; the rest ("sensitive": binade crossings, ties, tiny
//   values) are few, about one per binade.
// * Segment and walk: a regular increment is an integer R in units of its binade's
//   ulp, and a run of them between two sensitive elements stays in one binade, so
//   an int64 scan that restarts at every sensitive element gives each run's exact
//   sum.  One thread walks the sensitive elements in order, each getting
//   fl(v_prev + s x) from the exact value before it (the previous one plus the run
//   in between), then every element is rebuilt in parallel from the last sensitive
//   value before it.  Four passes: tile sums (fused into the kernel that produces
//   the increments), classification, walk, rebuild.
// * Every element is verified against fl(v_{j-1} + s x_j), so the estimate's bound
//   decides only the speed, never the result.  A violation (only when the bound or
//   the sensitive-list capacity is exceeded) is repaired by running the recurrence
//   itself from the first violating element.  The result equals the sequential
//   chain exactly.
#include "kernels.hpp"

namespace {
constexpr int CT = 256;             // threads per tile
constexpr int CI = 8;               // items per thread
constexpr int CTILE = CT * CI;
constexpr int RAW_CAP = 2048;         // sensitive elements sorted in LDS by the walk
constexpr int WALK_T = 1024;          // threads of the one-block kernels
constexpr int WALK_CHUNK = 256;       // sensitive elements staged per round of the walk
constexpr int WALK_REG = 16;          // tiles per walk thread held in registers (N <= 8.4e6 at CT*CI = 2048)

// ---------------------------------------------------------------- segment and walk
// The classification needs only the binade of each chain value, so it runs on a
// double estimate of the prefix sums with a rigorous error bound.  A regular element
// adds RN_u(s x) = R u, u = 2^(b-52) the ulp of its binade b, and all regular
// elements between two sensitive ones share one binade: a sum of the integers R that
// restarts at every sensitive element (a segmented scan) is exact in int64, and times
// u exact in double.
struct EsaArgs {
    const double *x;
    double *g;
    double sign;
    const double *v0p;
    int64_t *first_bad;         // [1] first element that violates the recurrence (L: none)
    const int64_t *Lp;
    int slack;                  // error bound widened by 2^slack (0; a test hook raises it)
    // per tile
    double *tsumP, *tpreP;      // estimated sums of s*x and their exclusive prefixes
    int64_t *tcnt, *ttail;      // sensitive count; regular sum after the tile's last sensitive element
    int64_t *toff, *tcarry;     // sensitive elements before the tile; regular sum of the segment open at its start
    // the sensitive elements in chain order
    int64_t *s_idx;
    double *s_pre;              // regular sum between the previous sensitive element and this one
    double *s_v;                // chain value after the element
    int64_t *s_n;               // count (uncapped)
    int64_t smax;               // capacity
    // appended by the classification in any order (RAW_CAP entries): index, regular sum
    // since the previous sensitive element of its tile (or the tile start), binade
    // before it
    int64_t *raw_n, *raw_idx, *raw_part;
    int32_t *raw_bp;
    double *raw_sx;                         // s x of the element
};

// A run of elements: regular sum after its last sensitive element (all of it when it
// has none) and its number of sensitive elements.
struct Seg {
    int64_t sum;
    int64_t cnt;
};
// a followed by b
__device__ __forceinline__ Seg seg_then(Seg a, Seg b) { return Seg{b.cnt ? b.sum : a.sum + b.sum, a.cnt + b.cnt}; }
__device__ __forceinline__ Seg shfl_up_seg(Seg s, int d)
{
    return Seg{__shfl_up(s.sum, d, 64), __shfl_up(s.cnt, d, 64)};
}

// Block-wide exclusive scans (every thread of the block calls; lds holds blockDim/64
// entries).  The wave totals are folded in wave order, so the results do not depend
// on timing.
__device__ double block_excl_sum(double v, double *lds, double *total)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, nw = blockDim.x >> 6;
    double x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    double ex = __shfl_up(x, 1, 64);
    if (lane == 0) ex = 0.0;
    if (lane == 63) lds[w] = x;
    __syncthreads();
    double off = 0.0, tot = 0.0;
    for (int q = 0; q < nw; q++) {
        const double s = lds[q];
        if (q < w) off += s;
        tot += s;
    }
    *total = tot;
    __syncthreads();
    return off + ex;
}

__device__ Seg block_excl_seg(Seg v, Seg *lds, Seg *total)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, nw = blockDim.x >> 6;
    Seg x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const Seg y = shfl_up_seg(x, d);
        if (lane >= d) x = seg_then(y, x);
    }
    Seg ex = shfl_up_seg(x, 1);
    if (lane == 0) ex = Seg{0, 0};
    if (lane == 63) lds[w] = x;
    __syncthreads();
    Seg off{0, 0}, tot{0, 0};
    for (int q = 0; q < nw; q++) {
        const Seg s = lds[q];
        if (q < w) off = seg_then(off, s);
        tot = seg_then(tot, s);
    }
    *total = tot;
    __syncthreads();
    return seg_then(off, ex);
}

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

// Bound on |v - P| for the estimate P of the chain value v after j elements: the
// chain's own j roundings (at most 2^-53 |v| each, |v| <= vmax for a monotone chain)
// and the roundings of the estimate's sums in any order (at most (j + 2) 2^-53
// (|v0| + sum |s x|) <= (j + 2) 2^-52 vmax), with a factor 2 to spare.  A chain for
// which it fails only takes the repair path: every element is verified.
__device__ __forceinline__ double err_bound(int64_t j, double vmax, int slack)
{
    return ldexp(((double)j + 2.0) * vmax, slack - 50) * 1.001;
}

// binade of the chain value before element j, or INT32_MIN when its error interval
// straddles binades; before element 0 the value is v0 exactly
__device__ __forceinline__ int start_binade(int64_t j, double v0, double P, int slack)
{
    if (j == 0) return interval_binade(v0, 0.0);
    return interval_binade(P, err_bound(j, fmax(fabs(v0), fabs(P)), slack));
}

// regular increment of an element in units of 2^(bp-52), or sensitive: bp, bq are the
// binades before and after it
__device__ __forceinline__ bool classify(double sx, int bp, int bq, int64_t *R)
{
    if (bp == INT32_MIN || bp != bq) return true;
    const double y = ldexp(sx, 52 - bp);      // exact unless it underflows, far below 1/2
    const double fy = floor(y);
    if (y - fy == 0.5 || !(fabs(y) <= 9007199254740992.0)) return true;
    *R = (int64_t)rint(y);
    return false;
}

// One tile's items, CI consecutive per thread.
struct TileItems {
    double sx[CI];      // s x_j (exact: s = +-1)
    double pa[CI];      // estimate of the chain value after the item
    int64_t R[CI];      // regular increment in units of 2^(bp - 52), 0 if sensitive
    int bp[CI];         // binade before the item
    uint32_t sens;      // bit u: item u sensitive
};

// A tile staged through LDS so that global memory is read and written in coalesced
// rows (element i of the tile by thread i mod CT) while each thread works on CI
// consecutive elements; one pad double per 8 keeps the per-thread reads free of bank
// conflicts.
constexpr int TILE_LDS = CTILE + CTILE / 8;
__device__ __forceinline__ int tile_slot(int i) { return i + (i >> 3); }

__device__ __forceinline__ void load_tile(const double *__restrict__ x, double sign, int64_t base, int64_t L,
                                          double *tbuf, double (&sx)[CI])
{
    const int t = threadIdx.x;
    double v[CI];                           // every load in flight before the LDS writes
#pragma unroll
    for (int u = 0; u < CI; u++) {
        const int64_t j = base + u * CT + t;
        v[u] = j < L ? x[j] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < CI; u++) tbuf[tile_slot(u * CT + t)] = base + u * CT + t < L ? sign * v[u] : 0.0;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CI; u++) sx[u] = tbuf[tile_slot(t * CI + u)];
}

// the reverse; tbuf must be free (every thread past its last read of it)
__device__ __forceinline__ void store_tile(double *__restrict__ g, int64_t base, int64_t L, double *tbuf,
                                           const double (&v)[CI])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int u = 0; u < CI; u++) tbuf[tile_slot(t * CI + u)] = v[u];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CI; u++) {
        const int i = u * CT + t;
        const int64_t j = base + i;
        if (j < L) g[j] = tbuf[tile_slot(i)];
    }
}

// Classification of one tile (items past the end are regular with R = 0), from the
// estimated prefixes before the tile and before the next one.  Every kernel
// classifies through this function on the same inputs, so they agree element for
// element; the binade at a tile start is taken from the same expression in both tiles
// that meet there.  Nearly always a thread's whole range of
// values lies in one binade, checked once with the bound at its last element: then
// only the ties remain to be found.
__device__ void tile_classify(const EsaArgs &a, int64_t tile, int64_t L, double v0, TileItems &it, double *lds,
                              double *tbuf, double pre_t, double pre_t1)
{
    const int64_t j0 = tile * CTILE + (int64_t)threadIdx.x * CI;
    load_tile(a.x, a.sign, tile * CTILE, L, tbuf, it.sx);
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < CI; u++) acc += it.sx[u];
    double tot;
    const double E = block_excl_sum(acc, lds, &tot);
    const double P0 = v0 + (pre_t + E);
    double P = P0, Pmin = P0, Pmax = P0;
#pragma unroll
    for (int u = 0; u < CI; u++) {
        P = P + it.sx[u];
        it.pa[u] = P;
        Pmin = fmin(Pmin, P);
        Pmax = fmax(Pmax, P);
    }
    it.sens = 0;
    const int64_t jl = j0 + CI < L ? j0 + CI : L;           // one past the thread's last element
    const bool edge = jl < L && jl % CTILE == 0;             // the next element opens the next tile
    const double Pn = edge ? v0 + (pre_t1 + 0.0) : 0.0;
    // fast path: every value before and after the thread's elements in one binade
    int bt = INT32_MIN;
    if (j0 < L) {
        const double vmax = fmax(fabs(v0), fmax(fabs(Pmin), fabs(Pmax)));
        const double err = err_bound(jl, vmax, a.slack);
        const double lo = fmin(Pmin, j0 == 0 ? v0 : Pmin) - err, hi = fmax(Pmax, j0 == 0 ? v0 : Pmax) + err;
        if (lo > 0.0 || hi < 0.0) {
            const int b1 = binade_of(lo), b2 = binade_of(hi);
            if (b1 == b2) bt = b1;
        }
        if (j0 == 0 && bt != interval_binade(v0, 0.0)) bt = INT32_MIN;
        if (edge && bt != start_binade(jl, v0, Pn, a.slack)) bt = INT32_MIN;
    }
    if (bt != INT32_MIN) {
#pragma unroll
        for (int u = 0; u < CI; u++) {
            it.R[u] = 0;
            it.bp[u] = bt;
            if (j0 + u < L && classify(it.sx[u], bt, bt, &it.R[u])) {
                it.sens |= 1u << u;
                it.R[u] = 0;
            }
        }
        return;
    }
    P = P0;
    int bp = start_binade(j0, v0, P, a.slack);
#pragma unroll
    for (int u = 0; u < CI; u++) {
        const int64_t j = j0 + u;
        const double Pq = it.pa[u];
        it.R[u] = 0;
        it.bp[u] = bp;
        if (j < L) {
            const double Pe = edge && j + 1 == jl ? Pn : Pq;
            const int bq = start_binade(j + 1, v0, Pe, a.slack);
            if (classify(it.sx[u], bp, bq, &it.R[u])) {
                it.sens |= 1u << u;
                it.R[u] = 0;
            }
            bp = bq;
        }
        P = Pq;
    }
}

__device__ __forceinline__ Seg items_seg(const TileItems &it)
{
    Seg s{0, 0};
#pragma unroll
    for (int u = 0; u < CI; u++) {
        if ((it.sens >> u) & 1) {
            s.sum = 0;
            s.cnt++;
        } else {
            s.sum += it.R[u];
        }
    }
    return s;
}

__device__ __forceinline__ double seg_value(int64_t sum, int bp)
{
    return sum == 0 || bp == INT32_MIN ? 0.0 : ldexp((double)sum, bp - 52);
}

// Tile sums of s*x (acc: this thread's share); block 0 also resets the chain's flags.
__device__ void esa_tile_sums(const EsaArgs &a, int64_t L, double acc, double *lds, uint32_t tile)
{
    double tot;
    block_excl_sum(acc, lds, &tot);
    if (threadIdx.x == 0) {
        a.tsumP[tile] = tot;
        if (tile == 0) {
            a.first_bad[0] = 0;
            a.first_bad[1] = L;
            *a.raw_n = 0;
        }
    }
}

// one block: exclusive prefixes of the nt tile sums, in tile order
struct esa_tile_scan_kernel {
    static constexpr int LB = WALK_T, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, EsaArgs a, int64_t nt)
{
    __shared__ double lds[WALK_T / 64];
    const int64_t chunk = (nt + WALK_T - 1) / WALK_T;
    const int64_t b0 = (int64_t)threadIdx.x * chunk;
    double s = 0.0;
    for (int64_t b = b0; b < b0 + chunk && b < nt; b++) s += a.tsumP[b];
    double tot;
    double run = block_excl_sum(s, lds, &tot);
    for (int64_t b = b0; b < b0 + chunk && b < nt; b++) {
        a.tpreP[b] = run;
        run += a.tsumP[b];
    }
}
};

struct esa_sum_kernel {
    static constexpr int LB = CT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, EsaArgs a)
{
    __shared__ double lds[CT / 64];
    const int64_t L = *a.Lp;
    const int64_t base = (int64_t)kg_.bid * CTILE;
    double v[CI];
#pragma unroll
    for (int u = 0; u < CI; u++) {
        const int64_t j = base + u * CT + threadIdx.x;
        v[u] = j < L ? a.x[j] : 0.0;
    }
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < CI; u++)
        if (base + u * CT + threadIdx.x < L) acc += a.sign * v[u];
    esa_tile_sums(a, L, acc, lds, kg_.bid);
}
};

// The estimated prefixes of tile t and t + 1 (the sums of tsumP before them), by one
// block: thread i sums the tiles i, i + CT, ... in order, and the threads' sums are
// reduced by a fixed tree.  Tile t + 1's own block forms the same sums term for term
// (its thread t mod CT ends with tile t), so both blocks get the same bits.
constexpr int64_t SELF_PREFIX_MAX = 16384;    // tiles; beyond, a one-block scan forms them
__device__ __forceinline__ double tree_sum(double v, double *lds)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    double r = lds[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) r += lds[q];
    __syncthreads();
    return r;
}
__device__ void tile_prefixes(const double *__restrict__ tsumP, int64_t t, double *lds, double *pre_t, double *pre_t1)
{
    double p = 0.0;
    for (int64_t q = threadIdx.x; q < t; q += CT) p += tsumP[q];
    const double p1 = (int64_t)threadIdx.x == t % CT ? p + tsumP[t] : p;
    *pre_t = tree_sum(p, lds);
    *pre_t1 = tree_sum(p1, lds);
}

// self_prefix: form the tile prefixes here (and store this tile's in tpreP for the
// later passes); otherwise read them from tpreP (esa_tile_scan_kernel).
struct esa_classify_kernel {
    static constexpr int LB = CT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, EsaArgs a, int self_prefix)
{
    __shared__ double ldsd[CT / 64];
    __shared__ Seg ldss[CT / 64];
    __shared__ double tbuf[TILE_LDS];
    const int64_t L = *a.Lp;
    const int64_t tile = kg_.bid;
    if (tile * CTILE >= L) return;
    double pre_t, pre_t1;
    if (self_prefix) {
        tile_prefixes(a.tsumP, tile, ldsd, &pre_t, &pre_t1);
        if (threadIdx.x == 0) a.tpreP[tile] = pre_t;
    } else {
        pre_t = a.tpreP[tile];
        pre_t1 = a.tpreP[tile + 1];
    }
    TileItems it;
    tile_classify(a, tile, L, *a.v0p, it, ldsd, tbuf, pre_t, pre_t1);
    Seg tot;
    const Seg pre = block_excl_seg(items_seg(it), ldss, &tot);
    if (threadIdx.x == 0) {
        a.tcnt[tile] = tot.cnt;
        a.ttail[tile] = tot.sum;
    }
    if (!it.sens) return;
    int64_t cur = pre.sum;
    const int64_t j0 = tile * CTILE + (int64_t)threadIdx.x * CI;
    for (int u = 0; u < CI; u++) {
        if ((it.sens >> u) & 1) {
            const unsigned long long slot = atomicAdd((unsigned long long *)a.raw_n, 1ull);
            if (slot < (unsigned long long)RAW_CAP) {
                a.raw_idx[slot] = j0 + u;
                a.raw_part[slot] = cur;
                a.raw_bp[slot] = it.bp[u];
                a.raw_sx[slot] = it.sx[u];
            }
            cur = 0;
        } else {
            cur += it.R[u];
        }
    }
}
};

// The sequential part: one thread walks the sensitive elements in order from the
// chain's start value, each getting fl(value before it + s x), the value before it
// being the previous one plus the regular sum in between (exact: the run stays in one
// binade).  The list is in global memory (staged through LDS in rounds) or, from the
// one-block kernel, already in LDS (lpre, lsx non-null).  A list beyond its capacity
// leaves every element after the last one kept to the repair.
__device__ void esa_walk_block(const EsaArgs &a, int64_t ns, const double *lpre, const double *lsx)
{
    __shared__ double sp[WALK_CHUNK], sx[WALK_CHUNK], sv[WALK_CHUNK];
    __shared__ double Vd;
    const int t = threadIdx.x;
    if (ns > a.smax) {
        if (t == 0)
            atomicMin((unsigned long long *)&a.first_bad[1], (unsigned long long)(a.s_idx[a.smax - 1] + 1));
        ns = a.smax;
    }
    if (lpre) {
        if (t == 0) {
            double vd = *a.v0p;
            for (int64_t k = 0; k < ns; k++) {
                const double pred = lpre[k] == 0.0 ? vd : vd + lpre[k];
                vd = pred + lsx[k];
                a.s_v[k] = vd;
            }
        }
        return;
    }
    if (t == 0) Vd = *a.v0p;
    __syncthreads();
    for (int64_t c0 = 0; c0 < ns; c0 += WALK_CHUNK) {
        const int64_t k = c0 + t;
        if (t < WALK_CHUNK && k < ns) {
            sp[t] = a.s_pre[k];
            sx[t] = a.sign * a.x[a.s_idx[k]];
        }
        __syncthreads();
        if (t == 0) {
            double vd = Vd;
            const int cnt = (int)(ns - c0 < WALK_CHUNK ? ns - c0 : WALK_CHUNK);
            for (int u = 0; u < cnt; u++) {
                const double pred = sp[u] == 0.0 ? vd : vd + sp[u];
                vd = pred + sx[u];
                sv[u] = vd;
            }
            Vd = vd;
        }
        __syncthreads();
        if (t < WALK_CHUNK && k < ns) a.s_v[k] = sv[t];
        __syncthreads();
    }
}

// One block: the tiles' sensitive offsets and open-segment carries (a segmented scan
// of the tile aggregates), then the ordered list of the sensitive elements and the
// walk.  The list is the classification's appended one put in order by index when it
// holds every sensitive element; otherwise (more than RAW_CAP, never for an ordinary
// cluster) this block classifies the tiles again one after another and lists them in
// order, slowly but exactly.
struct esa_walk_kernel {
    static constexpr int LB = CT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, EsaArgs a)
{
    __shared__ Seg lds[CT / 64];
    __shared__ int64_t ridx[RAW_CAP];
    __shared__ double lpre[RAW_CAP], lsx[RAW_CAP];
    const int t = threadIdx.x;
    const int64_t L = *a.Lp;
    const int64_t nt = (L + CTILE - 1) / CTILE;
    const int64_t chunk = (nt + CT - 1) / CT;
    const int64_t b0 = (int64_t)t * chunk;
    Seg tot;
    if (chunk <= WALK_REG) {
        // the thread's tile entries in registers, every load issued before the first use
        Seg v[WALK_REG];
#pragma unroll
        for (int u = 0; u < WALK_REG; u++) {
            const int64_t b = b0 + u;
            v[u] = u < chunk && b < nt ? Seg{a.ttail[b], a.tcnt[b]} : Seg{0, 0};
        }
        Seg acc{0, 0};
#pragma unroll
        for (int u = 0; u < WALK_REG; u++) acc = seg_then(acc, v[u]);
        Seg run = block_excl_seg(acc, lds, &tot);
#pragma unroll
        for (int u = 0; u < WALK_REG; u++) {
            const int64_t b = b0 + u;
            if (u < chunk && b < nt) {
                a.toff[b] = run.cnt;
                a.tcarry[b] = run.sum;
            }
            run = seg_then(run, v[u]);
        }
    } else {
        Seg acc{0, 0};
        for (int64_t b = b0; b < b0 + chunk && b < nt; b++) acc = seg_then(acc, Seg{a.ttail[b], a.tcnt[b]});
        Seg run = block_excl_seg(acc, lds, &tot);
        for (int64_t b = b0; b < b0 + chunk && b < nt; b++) {
            a.toff[b] = run.cnt;
            a.tcarry[b] = run.sum;
            run = seg_then(run, Seg{a.ttail[b], a.tcnt[b]});
        }
    }
    const int64_t ns = tot.cnt;
    if (t == 0) *a.s_n = ns;
    __syncthreads();                                // (also makes the block's global writes visible to it)
    if (ns > RAW_CAP) {
        __shared__ double ldsd[CT / 64];
        __shared__ double tbuf[TILE_LDS];
        for (int64_t tile = 0; tile < nt; tile++) {
            TileItems it;
            tile_classify(a, tile, L, *a.v0p, it, ldsd, tbuf, a.tpreP[tile], a.tpreP[tile + 1]);
            Seg ttot;
            const Seg pre = block_excl_seg(items_seg(it), lds, &ttot);
            int64_t k = a.toff[tile] + pre.cnt;
            int64_t cur = pre.sum + (pre.cnt ? 0 : a.tcarry[tile]);
            const int64_t j0 = tile * CTILE + (int64_t)t * CI;
            for (int u = 0; u < CI; u++) {
                if ((it.sens >> u) & 1) {
                    if (k < a.smax) {
                        a.s_idx[k] = j0 + u;
                        a.s_pre[k] = seg_value(cur, it.bp[u]);
                    }
                    k++;
                    cur = 0;
                } else {
                    cur += it.R[u];
                }
            }
            __syncthreads();
        }
        esa_walk_block(a, ns, nullptr, nullptr);
        return;
    }
    for (int64_t e = t; e < ns; e += CT) ridx[e] = a.raw_idx[e];
    __syncthreads();
    for (int64_t e = t; e < ns; e += CT) {
        const int64_t j = ridx[e];
        int64_t rank = 0;
        for (int64_t q = 0; q < ns; q++) rank += ridx[q] < j;      // the indices are distinct
        const int64_t tile = j / CTILE;
        const int64_t pre = a.raw_part[e] + (rank == a.toff[tile] ? a.tcarry[tile] : 0);
        a.s_idx[rank] = j;
        lpre[rank] = seg_value(pre, a.raw_bp[e]);
        lsx[rank] = a.raw_sx[e];
    }
    __syncthreads();
    esa_walk_block(a, ns, lpre, lsx);
}
};

// Parallel rebuild of every element from the last sensitive value before it,
// verification of the recurrence against the values actually written (the value
// before a tile is rebuilt exactly as the previous tile wrote it), and the first
// violation.
struct esa_rebuild_kernel {
    static constexpr int LB = CT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, EsaArgs a)
{
    __shared__ double ldsd[CT / 64];
    __shared__ Seg ldss[CT / 64];
    __shared__ double lastv[CT];
    __shared__ double tbuf[TILE_LDS];
    const int64_t L = *a.Lp;
    const int64_t tile = kg_.bid;
    const int64_t base = tile * CTILE;
    if (base >= L) return;
    const double v0 = *a.v0p;
    const int t = threadIdx.x;
    int64_t ns_eff = *a.s_n;
    if (ns_eff > a.smax) ns_eff = a.smax;
    TileItems it;
    tile_classify(a, tile, L, v0, it, ldsd, tbuf, a.tpreP[tile], a.tpreP[tile + 1]);
    Seg tot;
    const Seg pre = block_excl_seg(items_seg(it), ldss, &tot);
    const int64_t toff = a.toff[tile], tcarry = a.tcarry[tile];
    int64_t k = toff + pre.cnt;                       // rank of the next sensitive element
    int64_t cur = pre.sum + (pre.cnt ? 0 : tcarry);
    const int64_t j0 = base + (int64_t)t * CI;
    double nv[CI];
#pragma unroll
    for (int u = 0; u < CI; u++) {
        if ((it.sens >> u) & 1) {
            nv[u] = k < ns_eff ? a.s_v[k] : it.pa[u];
            k++;
            cur = 0;
        } else {
            cur += it.R[u];
            const int64_t kk = k - 1;                 // the last sensitive element before
            if (kk >= ns_eff) {
                nv[u] = it.pa[u];
            } else {
                const double b = kk >= 0 ? a.s_v[kk] : v0;
                nv[u] = cur == 0 ? b : b + seg_value(cur, it.bp[u]);
            }
        }
    }
    lastv[t] = nv[CI - 1];
    __syncthreads();
    double before;
    if (t > 0) {
        before = lastv[t - 1];
    } else if (tile == 0) {
        before = v0;
    } else {
        const int64_t kk = toff - 1;
        if (kk >= ns_eff) before = it.pa[0] - it.sx[0];        // past the list: already left to the repair
        else if (kk >= 0 && a.s_idx[kk] == base - 1) before = a.s_v[kk];
        else {
            const double b = kk >= 0 ? a.s_v[kk] : v0;
            before = tcarry == 0 ? b : b + seg_value(tcarry, it.bp[0]);
        }
    }
    int64_t bad = L;
#pragma unroll
    for (int u = 0; u < CI; u++) {
        const int64_t j = j0 + u;
        if (j < L) {
            const double want = before + it.sx[u];
            if (want != nv[u] && j < bad) bad = j;
        }
        before = nv[u];
    }
    if (bad < L) atomicMin((unsigned long long *)&a.first_bad[1], (unsigned long long)bad);
    store_tile(a.g, base, L, tbuf, nv);
}
};

// The repair, only after a violation (the estimate's bound or the list's capacity
// exceeded, which a monotone chain of ordinary size never does): the recurrence
// itself from the first violating element on, one thread over chunks that the block
// stages through LDS in coalesced rows.
constexpr int FIN_CHUNK = 4 * WALK_T;
struct chain_finish_kernel {
    static constexpr int LB = WALK_T, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, EsaArgs a)
{
    __shared__ double buf[FIN_CHUNK];
    __shared__ double vlast;
    const int64_t L = *a.Lp;
    const int64_t j0 = a.first_bad[1];
    if (j0 >= L) return;
    const int t = threadIdx.x;
    if (t == 0) vlast = j0 == 0 ? *a.v0p : a.g[j0 - 1];
    for (int64_t c0 = j0; c0 < L; c0 += FIN_CHUNK) {
        const int cnt = (int)(L - c0 < FIN_CHUNK ? L - c0 : FIN_CHUNK);
        for (int i = t; i < cnt; i += WALK_T) buf[i] = a.sign * a.x[c0 + i];
        __syncthreads();
        if (t == 0) {
            double v = vlast;
            for (int i = 0; i < cnt; i++) {
                v = v + buf[i];
                buf[i] = v;
            }
            vlast = v;
        }
        __syncthreads();
        for (int i = t; i < cnt; i += WALK_T) a.g[c0 + i] = buf[i];
        __syncthreads();
    }
}
};

// ---------------------------------------------------------------- the three chains
struct pot_init_kernel { static constexpr int LB = 1024, MB = 0; static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s) { s->first_inf = s->n_max + 1; } };

// cmc_utils.c:685-696: the count stops at the first k with !(r[k] < SF_INFINITY).
struct pot_count_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ r, DevScalars *s)
{
    int64_t k = 1 + (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (k > s->n_max) return;
    if (!(r[k] < CMCG_SF_INFINITY)) atomicMin((unsigned long long *)&s->first_inf, (unsigned long long)k);
}
};

struct PotWork {
    double *xa, *ga;           // chain buffers (increments, values)
    double *Mk;                // chain B result: M_{N-1-j}
    double *scal;              // [0] 0.0  [1] Mtotal
    int64_t *len;              // [0] L_A  [1] L_B  [2] L_C
    int64_t *fbad;             // 2 per chain
};

// N_MAX from the count, the chain lengths, and the sentinels before the recurrence
// reads r[N+1] (cmc_utils.c:713-714)
struct pot_setup_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s, PotWork w, double *__restrict__ r, double *__restrict__ phi)
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
};

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

// chain B increments (star_m[k] / clus.N_STAR, k = N - j) and their tile sums; the
// first thread also finishes chain A (Mtotal, the start value of chain B), which the
// increments do not use
struct prep_b_kernel {
    static constexpr int LB = CT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ m, DevScalars *s, double nstar,
                                                    double *__restrict__ x, KParams P, const double *__restrict__ ga,
                                                    PotWork w, EsaArgs e)
{
    __shared__ double lds[CT / 64];
    if (kg_.bid == 0 && threadIdx.x == 0) pot_total(s, P, ga, w);
    const int64_t n = s->n_max;
    const int64_t L = n > 1 ? n - 1 : 0;
    const int64_t base = (int64_t)kg_.bid * CTILE;
    double mv[CI];
#pragma unroll
    for (int u = 0; u < CI; u++) {
        const int64_t j = base + u * CT + threadIdx.x;
        mv[u] = j < L ? m[n - j] : 0.0;
    }
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < CI; u++) {
        const int64_t j = base + u * CT + threadIdx.x;
        if (j < L) {
            const double v = mv[u] / nstar;
            x[j] = v;
            acc += e.sign * v;
        }
    }
    esa_tile_sums(e, L, acc, lds, kg_.bid);
}
};

// p_k = M_k * (1/r_k - 1/r_{k+1}) with the reference's roundings, and their tile sums;
// chain C runs on w = -phi, w_k = fl(w_{k+1} + p_k)
struct prep_c_kernel {
    static constexpr int LB = CT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ r, DevScalars *s,
                                                    const double *__restrict__ Mk, double *__restrict__ x, EsaArgs e)
{
    __shared__ double lds[CT / 64];
    const int64_t n = s->n_max;
    const int64_t base = (int64_t)kg_.bid * CTILE;
    double Mv[CI], rk[CI], rk1[CI];
    const double Mtot = s->Mtotal;
#pragma unroll
    for (int u = 0; u < CI; u++) {
        const int64_t j = base + u * CT + threadIdx.x;
        const int64_t k = n - j;
        Mv[u] = rk[u] = rk1[u] = 1.0;
        if (j < n) {
            Mv[u] = k == n ? Mtot : Mk[n - 1 - k];
            rk[u] = r[k];
            rk1[u] = r[k + 1];
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < CI; u++) {
        const int64_t j = base + u * CT + threadIdx.x;
        if (j < n) {
            const double v = Mv[u] * (1.0 / rk[u] - 1.0 / rk1[u]);
            x[j] = v;
            acc += e.sign * v;
        }
    }
    esa_tile_sums(e, n, acc, lds, kg_.bid);
}
};

// phi[k] = -g[n-k] (k = 1..n) and phi[0] (cmc_utils.c:729); thread t handles node
// k = t for k = 0..n+1 and also writes the packed (r, phi) node table and its every
// MID_STEP-th entry, which the searches of set_velocities3 and the orbit phase read.
struct phi_out_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ r, double *__restrict__ phi, DevScalars *s, KParams P,
                               const double *__restrict__ g, double2 *__restrict__ nodes, double2 *__restrict__ mid)
{
    const int64_t k = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    const int64_t n = s->n_max;
    if (k > n + 1) return;
    double p;
    if (k == 0) {
        // no stars: phi[1] is the sentinel 0
        p = (n == 0 ? 0.0 : -g[n - 1]) + P.cenma_m * P.madhoc / r[1];
        phi[0] = p;
    } else if (k <= n) {
        p = -g[n - k];
        phi[k] = p;
        if (isnan(p)) dev_set_error(s, CMCG_E_NAN, k);
    } else {
        p = 0.0;                                  // phi[n + 1], set with r[n + 1] by pot_setup
    }
    const double2 v = make_double2(r[k], p);
    nodes[k] = v;
    if (k % MID_STEP == 0) mid[k / MID_STEP] = v;
}
};

struct pot_at_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ r, const double *__restrict__ phi, DevScalars *s,
                              const double *__restrict__ x, double *__restrict__ out, int64_t count)
{
    int64_t i = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int64_t n = s->n_max;
    bool bad = false;
    double v = potential_at(r, phi, n, x[i], 1, n + 1, &bad);
    if (bad) { dev_set_error(s, CMCG_E_BRACKET, i); v = 0.0; }
    out[i] = v;
}
};

// The chain's increments must be in e.x; `summed`: their tile sums are already formed
// (by a producer that calls esa_tile_sums over the same grid).
hipError_t run_chain(const EsaArgs &e, int64_t n_ub, hipStream_t st, bool summed = false)
{
    int64_t ntiles = (n_ub + CTILE - 1) / CTILE;
    if (ntiles < 1) ntiles = 1;
    dim3 gt((unsigned)ntiles), bt(CT);
    if (!summed) kl<esa_sum_kernel>(gt, bt, 0, st, e);
    const bool self_prefix = ntiles <= SELF_PREFIX_MAX;
    if (!self_prefix) kl<esa_tile_scan_kernel>(dim3(1), dim3(WALK_T), 0, st, e, ntiles);
    kl<esa_classify_kernel>(gt, bt, 0, st, e, (int)self_prefix);
    kl<esa_walk_kernel>(dim3(1), dim3(CT), 0, st, e);
    kl<esa_rebuild_kernel>(gt, bt, 0, st, e);
    kl<chain_finish_kernel>(dim3(1), dim3(WALK_T), 0, st, e);
    return hipGetLastError();
}

// The chains' per-tile and list buffers in the context's scratch (tile_sums: six
// 8-byte words per tile).
EsaArgs esa_args(const Scratch &sc, int slack)
{
    double *tb = sc.tile_sums;
    const int64_t T = sc.max_tiles;
    EsaArgs e{};
    e.slack = slack;
    e.tsumP = tb;
    e.tpreP = tb + T;
    e.tcnt = (int64_t *)(tb + 2 * T);
    e.ttail = (int64_t *)(tb + 3 * T);
    e.toff = (int64_t *)(tb + 4 * T);
    e.tcarry = (int64_t *)(tb + 5 * T);
    e.s_idx = sc.esa_idx;
    e.s_pre = sc.esa_racc;
    e.s_v = sc.esa_v;
    e.s_n = sc.esa_n;
    e.smax = sc.esa_max;
    e.raw_n = sc.esa_n + 1;
    e.raw_idx = sc.esa_raw_idx;
    e.raw_part = (int64_t *)sc.esa_raw_part;
    e.raw_bp = (int32_t *)(sc.esa_raw_part + RAW_CAP);
    e.raw_sx = sc.esa_raw_part + 2 * RAW_CAP;
    return e;
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

struct mpc_prep_m_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ m, const DevScalars *s, double nstar,
                                  double *__restrict__ x, int64_t *__restrict__ len)
{
    const int64_t j = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    const int64_t n = s->n_max;
    if (j == 0) len[3] = n;
    if (j < n) x[j] = m[j + 1] / nstar;                   // star_m[k] / clus.N_STAR, k = j + 1
}
};

// term 0: 0.5 m madhoc vr vr, 1: 0.5 m madhoc vt vt, 2: vr vr, 3: vt vt (left to right)
struct mpc_prep_v_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const DevFields f, const DevScalars *s, double madhoc, int term,
                                  double *__restrict__ x, int64_t *__restrict__ len)
{
    const int64_t j = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
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
};

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
struct mpc_assign_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ r, const DevScalars *s, const double *__restrict__ g,
                                  const double *__restrict__ pc, int count, int bh_include, double mprev0,
                                  double itm, double minimum_r, MassPcOut *o)
{
    if (threadIdx.x != 0 || kg_.bid != 0) return;
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
};

struct mpc_init_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, double *v0, double mprev0)
{
    v0[0] = mprev0;
    v0[1] = 0.0;
}
};

struct mpc_take_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ g, MassPcOut *o, int term, int count)
{
    const int i = threadIdx.x;
    if (i >= count || i >= o->reached) return;
    const int64_t k = o->k[i];
    o->v[3 + term][i] = k > 0 ? g[k - 1] : 0.0;
}
};

hipError_t launch_potential(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub,
                            const Scratch &sc, hipStream_t st)
{
    int64_t ntiles = (n_ub + CTILE - 1) / CTILE + 1;
    if (ntiles > sc.max_tiles) return hipErrorInvalidValue;
    PotWork w;
    w.xa = sc.chain_x;
    w.ga = sc.chain_g;
    w.Mk = sc.chain_mk;
    w.scal = sc.chain_scal;
    w.len = sc.chain_len;
    w.fbad = sc.chain_fbad;
    // a test hook widens the classification's error bound, which sends most elements
    // through the fallback paths
    const EsaArgs e = esa_args(sc, P.chain_k_drop);
    const unsigned g1 = (unsigned)((n_ub + 256) / 256);
    const dim3 gt((unsigned)(n_ub > 0 ? (n_ub + CTILE - 1) / CTILE : 1));
    kl<pot_init_kernel>(dim3(1), dim3(1), 0, st, s);
    if (n_ub > 0) kl<pot_count_kernel>(dim3(g1), dim3(256), 0, st, f.r, s);
    kl<pot_setup_kernel>(dim3(1), dim3(1), 0, st, s, w, f.r, f.phi);
    // chain A: total mass, straight over star_m[1..N]
    EsaArgs ae = e;
    ae.x = f.m + 1; ae.g = w.ga; ae.sign = 1.0; ae.v0p = w.scal + 0; ae.first_bad = w.fbad + 0; ae.Lp = w.len + 0;
    if (hipError_t err = run_chain(ae, n_ub, st)) return err;
    // chain B: enclosed mass, iterated directly in Mk (chain C's input)
    EsaArgs be = e;
    be.x = w.xa; be.g = w.Mk; be.sign = -1.0; be.v0p = w.scal + 1; be.first_bad = w.fbad + 2; be.Lp = w.len + 1;
    kl<prep_b_kernel>(gt, dim3(CT), 0, st, f.m, s, (double)P.n_star, w.xa, P, (const double *)w.ga, w,
                       be);
    if (hipError_t err = run_chain(be, n_ub, st, true)) return err;
    // chain C: w = -phi
    EsaArgs ce = e;
    ce.x = w.xa; ce.g = w.ga; ce.sign = 1.0; ce.v0p = w.scal + 0; ce.first_bad = w.fbad + 4; ce.Lp = w.len + 2;
    kl<prep_c_kernel>(gt, dim3(CT), 0, st, f.r, s, (const double *)w.Mk, w.xa, ce);
    if (hipError_t err = run_chain(ce, n_ub, st, true)) return err;
    kl<phi_out_kernel>(dim3((unsigned)((n_ub + 2 + 255) / 256)), dim3(256), 0, st, f.r, f.phi, s, P,
                       (const double *)w.ga, (double2 *)sc.nodes, (double2 *)sc.mid);
    return hipGetLastError();
}

hipError_t launch_potential_at(const DevFields &f, DevScalars *s, const KParams &P, const double *x,
                               double *out, int64_t count, hipStream_t st)
{
    (void)P;
    if (count <= 0) return hipSuccess;
    kl<pot_at_kernel>(dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, f.r, f.phi, s, x,
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
    EsaArgs e = esa_args(sc, P.chain_k_drop);
    e.x = sc.chain_x; e.g = sc.chain_g; e.sign = 1.0; e.first_bad = sc.chain_fbad + 6;
    e.Lp = sc.chain_len + 3;
    double *v0 = sc.chain_scal + 2;              // start values: [2] mprev0, [3] 0
    const double mprev0 = bh_include ? P.cenma_m * P.madhoc : 0.0;   // cmc_utils.c:878
    kl<mpc_init_kernel>(dim3(1), dim3(1), 0, st, v0, mprev0);
    MassPcOut *o = (MassPcOut *)out;
    unsigned g1 = (unsigned)((n_ub + 256) / 256);
    // chain M
    kl<mpc_prep_m_kernel>(dim3(g1), dim3(256), 0, st, f.m, (const DevScalars *)s, (double)P.n_star,
                       sc.chain_x, sc.chain_len);
    EsaArgs em = e;
    em.v0p = v0;
    if (hipError_t err = run_chain(em, n_ub, st)) return err;
    kl<mpc_assign_kernel>(dim3(1), dim3(1), 0, st, (const double *)f.r, (const DevScalars *)s,
                       (const double *)sc.chain_g, pc, count, bh_include, mprev0, itm, minimum_r, o);
    // the four running sums
    for (int term = 0; term < 4; term++) {
        kl<mpc_prep_v_kernel>(dim3(g1), dim3(256), 0, st, f, (const DevScalars *)s, P.madhoc, term,
                           sc.chain_x, sc.chain_len);
        EsaArgs ev = e;
        ev.v0p = v0 + 1;
        if (hipError_t err = run_chain(ev, n_ub, st)) return err;
        kl<mpc_take_kernel>(dim3(1), dim3(MPC_MAX), 0, st, (const double *)sc.chain_g, o, term, count);
    }
    return hipGetLastError();
}

size_t mass_percent_out_bytes() { return sizeof(MassPcOut); }

// ---------------------------------------------------------------- central_calculate
// cmc_utils.c:1239-1345 (J = 6) and clusdyn_calculate cmc_utils.c:1614-1624.  The
// half-mass index nave comes from the enclosed-mass chain of comp_mass_percent
// (star_m[k] / N_STAR, from 0); the six density-weighted sums over i = 1..nave are
// sequential recurrences solved by the exact chain emulation.  Each sum's terms are
// first scaled by a power of two that brings nave * max(term) to at most 1, so no
// partial sum can overflow; binary rounding is invariant under such a scaling, so the
// scaled chain rounds exactly as the unscaled one.
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

struct cen_init_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, CentralOut *o, double *v0)
{
    if (threadIdx.x != 0 || kg_.bid != 0) return;
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
};

// nave and rh from the chain g[k-1] = m after star k (cmc_utils.c:1247-1257, 1620-1623);
// the chain length for the sums becomes nave
struct cen_nave_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ r, const DevScalars *s, const double *__restrict__ g,
                                int64_t n_star, CentralOut *o, int64_t *len)
{
    if (threadIdx.x != 0 || kg_.bid != 0) return;
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
};

struct cen_max_kernel {
    static constexpr int LB = 256, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const DevFields f, double madhoc, CentralOut *o)
{
    __shared__ unsigned long long red[CEN_TERMS][256 / 64];
    const int64_t nave = o->nave;
    unsigned long long mx[CEN_TERMS] = {0, 0, 0, 0, 0, 0};
    for (int64_t j = (int64_t)kg_.bid * blockDim.x + threadIdx.x; j < nave; j += (int64_t)kg_.nb * blockDim.x)
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
};

struct cen_scale_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, CentralOut *o)
{
    if (threadIdx.x != 0 || kg_.bid != 0 || o->error) return;
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
};

struct cen_prep_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const DevFields f, double madhoc, int t, CentralOut *o, double *__restrict__ x)
{
    const int64_t j = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
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
};

struct cen_take_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ g, CentralOut *o, int t)
{
    if (threadIdx.x != 0 || kg_.bid != 0 || o->error) return;
    o->S[t] = g[o->nave - 1] * o->unscale[t];
}
};

// the averages and derived quantities, cmc_utils.c:1323-1345
struct cen_final_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, CentralOut *o)
{
    if (threadIdx.x != 0 || kg_.bid != 0 || o->error) return;
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
};

hipError_t launch_central(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub, const Scratch &sc,
                          void *out, hipStream_t st)
{
    int64_t ntiles = (n_ub + CTILE - 1) / CTILE + 1;
    if (ntiles > sc.max_tiles) return hipErrorInvalidValue;
    EsaArgs e = esa_args(sc, P.chain_k_drop);
    e.x = sc.chain_x; e.g = sc.chain_g; e.sign = 1.0; e.first_bad = sc.chain_fbad + 6;
    e.Lp = sc.chain_len + 3;
    double *v0 = sc.chain_scal + 2;              // both start values 0
    CentralOut *o = (CentralOut *)out;
    const unsigned g1 = (unsigned)((n_ub + 256) / 256);
    kl<cen_init_kernel>(dim3(1), dim3(1), 0, st, o, v0);
    // enclosed mass, star_m[k] / N_STAR from 0
    kl<mpc_prep_m_kernel>(dim3(g1), dim3(256), 0, st, f.m, (const DevScalars *)s, (double)P.n_star,
                       sc.chain_x, sc.chain_len);
    EsaArgs em = e;
    em.v0p = v0;
    if (hipError_t err = run_chain(em, n_ub, st)) return err;
    kl<cen_nave_kernel>(dim3(1), dim3(1), 0, st, (const double *)f.r, (const DevScalars *)s,
                       (const double *)sc.chain_g, (int64_t)P.n_star, o, sc.chain_len);
    const unsigned gmax = g1 < 1024u ? g1 : 1024u;
    kl<cen_max_kernel>(dim3(gmax), dim3(256), 0, st, f, P.madhoc, o);
    kl<cen_scale_kernel>(dim3(1), dim3(1), 0, st, o);
    for (int t = 0; t < CEN_TERMS; t++) {
        kl<cen_prep_kernel>(dim3(g1), dim3(256), 0, st, f, P.madhoc, t, o, sc.chain_x);
        EsaArgs es = e;
        es.v0p = v0 + 1;
        if (hipError_t err = run_chain(es, n_ub, st)) return err;
        kl<cen_take_kernel>(dim3(1), dim3(1), 0, st, (const double *)sc.chain_g, o, t);
    }
    kl<cen_final_kernel>(dim3(1), dim3(1), 0, st, o);
    return hipGetLastError();
}

size_t central_out_bytes() { return sizeof(CentralOut); }
