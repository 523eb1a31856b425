// a2: potential_calculate (src/cmc/cmc_utils.c:675-736), reproducing the reference's
// sequential FP64 rounding bit for bit with parallel scans.
//
// The reference evaluates three floating-point chains one element at a time:
//   A  mprev += m_k                  k = 1..N          (total mass, :685-696)
//   B  mprev -= m_k / N_STAR         k = N..2          (enclosed mass M_k, :723)
//   C  phi_k = phi_{k+1} - M_k (1/r_k - 1/r_{k+1})     (:722)
// Each chain is v_j = fl(v_{j-1} + s*x_j).  Repeated rounding makes the result
// drift from the exact sums (6e-11 relative at N = 1e5), so a correctly rounded
// parallel scan does not reproduce it.  Here each chain is solved as the fixed
// point of
//     v_j = v_{-1} + sum_{i<=j} [ fl(g_{i-1} + s*x_i) - g_{i-1} ],   g = previous iterate,
// where every bracket is an exact integer in 2^-100 units (int128), so each
// iteration is an exact parallel prefix sum.  An element's increment depends on
// its predecessor only through the predecessor's binade (and last bit at binade
// crossings and ties), so the iteration converges in a few rounds; every round
// restarts at the first element that does not yet satisfy the recurrence
// (everything before it is final), and a bounded sequential pass finishes any
// remainder after the round limit.  The result equals the sequential chain
// exactly: it is verified element by element against fl(v_{j-1} + s*x_j).
#include "kernels.hpp"

namespace {
constexpr int CT = 256;             // threads per tile
constexpr int CI = 8;               // items per thread
constexpr int CTILE = CT * CI;
constexpr int SCAN_T = 1024;

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

// block-wide exclusive scan of int128 through LDS (Hillis-Steele)
template <int NT>
__device__ i128 block_excl_scan(i128 v, i128 *lds, i128 *total)
{
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {
        i128 add = t >= off ? lds[t - off] : (i128)0;
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    i128 incl = lds[t];
    *total = lds[NT - 1];
    __syncthreads();
    return incl - v;
}

// Round 0: zero iterate (every predecessor 0 except v_{-1} = v0): its increments are
// the exact-arithmetic partial sums.
__global__ __launch_bounds__(CT) void chain_seed_kernel(ChainArgs a)
{
    __shared__ i128 lds[CT];
    const int64_t L = *a.Lp;
    if (blockIdx.x == 0 && threadIdx.x == 0) { a.first_bad[0] = 0; a.first_bad[1] = 0; }
    const int64_t base = (int64_t)blockIdx.x * CTILE;
    if (base >= L) return;
    const double v0 = *a.v0p;
    i128 acc = 0;
    for (int u = 0; u < CI; u++) {
        int64_t j = base + (int64_t)threadIdx.x * CI + u;
        if (j < L) {
            acc += incr(j == 0 ? v0 : 0.0, a.x[j], a.sign, a.K);
            a.g[j] = 0.0;
        }
    }
    i128 tot;
    block_excl_scan<CT>(acc, lds, &tot);
    if (threadIdx.x == 0) a.tsum[blockIdx.x] = tot;
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

__global__ void pot_setup_kernel(DevScalars *s, PotWork w)
{
    int64_t n = s->first_inf - 1;
    s->n_max = n;
    w.len[0] = n;                       // A: k = 1..N
    w.len[1] = n > 1 ? n - 1 : 0;       // B: k = N..2 gives M_{N-1}..M_1
    w.len[2] = n;                       // C: k = N..1
    w.scal[0] = 0.0;
}

// sentinels before the recurrence reads r[N+1] (cmc_utils.c:713-714)
__global__ void sentinel_kernel(double *__restrict__ r, double *__restrict__ phi, DevScalars *s)
{
    int64_t n = s->n_max;
    r[n + 1] = CMCG_SF_INFINITY;
    phi[n + 1] = 0.0;
}

__global__ void prep_a_kernel(const double *__restrict__ m, DevScalars *s, double *__restrict__ x)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n = s->n_max;
    if (j < n) x[j] = m[j + 1];
}

// after chain A: Mtotal, Rtidal (cmc_utils.c:703-707), NaN check (:688-691)
__global__ void total_kernel(DevScalars *s, KParams P, const double *__restrict__ ga, PotWork w)
{
    int64_t n = s->n_max;
    double mprev = n > 0 ? ga[n - 1] : 0.0;
    if (isnan(mprev)) dev_set_error(s, CMCG_E_NAN, n);
    double Mtotal = mprev * P.madhoc + P.cenma_m * P.madhoc;
    s->Mtotal = Mtotal;
    s->Rtidal = P.orbit_r * pow(Mtotal, 1.0 / 3.0);
    w.scal[1] = Mtotal;
}

__global__ void prep_b_kernel(const double *__restrict__ m, DevScalars *s, double nstar, double *__restrict__ x)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n = s->n_max;
    if (j < n - 1) x[j] = m[n - j] / nstar;           // star_m[k] / clus.N_STAR, k = N - j
}

__global__ void save_mk_kernel(DevScalars *s, const double *__restrict__ g, double *__restrict__ Mk)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < s->n_max - 1) Mk[j] = g[j];
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
    if (k > n) return;
    double p = -g[n - k];
    phi[k] = p;
    if (isnan(p)) dev_set_error(s, CMCG_E_NAN, k);
    if (k == 1) phi[0] = p + P.cenma_m * P.madhoc / r[1];    // cmc_utils.c:729
}

__global__ void phi0_empty_kernel(const double *__restrict__ r, double *__restrict__ phi, DevScalars *s, KParams P)
{
    if (s->n_max == 0) phi[0] = phi[1] + P.cenma_m * P.madhoc / r[1];
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

hipError_t run_chain(ChainArgs a, int rounds, int64_t n_ub, hipStream_t st)
{
    int64_t ntiles = (n_ub + CTILE - 1) / CTILE;
    if (ntiles < 1) ntiles = 1;
    hipLaunchKernelGGL(chain_seed_kernel, dim3((unsigned)ntiles), dim3(CT), 0, st, a);
    for (int it = 0; it < rounds; it++) {
        hipLaunchKernelGGL(chain_scan_kernel, dim3(1), dim3(SCAN_T), 0, st, a);
        hipLaunchKernelGGL(chain_round_kernel, dim3((unsigned)ntiles), dim3(CT), 0, st, a);
    }
    hipLaunchKernelGGL(chain_finish_kernel, dim3(1), dim3(1), 0, st, a);
    return hipGetLastError();
}
} // namespace

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
    w.tsum = (i128 *)sc.tile_sums;          // 3 x 16 B per tile
    w.tpre = w.tsum + sc.max_tiles;
    w.bnd = (double *)(w.tpre + sc.max_tiles);
    unsigned g1 = (unsigned)((n_ub + 256) / 256);
    hipLaunchKernelGGL(pot_init_kernel, dim3(1), dim3(1), 0, st, s);
    if (n_ub > 0) hipLaunchKernelGGL(pot_count_kernel, dim3(g1), dim3(256), 0, st, f.r, s);
    hipLaunchKernelGGL(pot_setup_kernel, dim3(1), dim3(1), 0, st, s, w);
    hipLaunchKernelGGL(sentinel_kernel, dim3(1), dim3(1), 0, st, f.r, f.phi, s);
    // chain A: total mass (stored masses are O(1): 2^-70 resolution)
    hipLaunchKernelGGL(prep_a_kernel, dim3(g1), dim3(256), 0, st, f.m, s, w.xa);
    ChainArgs a{w.xa, w.ga, 1.0, w.scal + 0, w.fbad + 0, w.tsum, w.tpre, w.bnd, w.len + 0, 70};
    if (hipError_t e = run_chain(a, 16, n_ub, st)) return e;
    hipLaunchKernelGGL(total_kernel, dim3(1), dim3(1), 0, st, s, P, (const double *)w.ga, w);
    // chain B: enclosed mass
    hipLaunchKernelGGL(prep_b_kernel, dim3(g1), dim3(256), 0, st, f.m, s, (double)P.n_star, w.xa);
    ChainArgs b{w.xa, w.ga, -1.0, w.scal + 1, w.fbad + 2, w.tsum, w.tpre, w.bnd, w.len + 1, 100};
    if (hipError_t e = run_chain(b, 8, n_ub, st)) return e;
    hipLaunchKernelGGL(save_mk_kernel, dim3(g1), dim3(256), 0, st, s, (const double *)w.ga, w.Mk);
    // chain C: w = -phi
    hipLaunchKernelGGL(prep_c_kernel, dim3(g1), dim3(256), 0, st, f.r, s, (const double *)w.Mk, w.xa);
    ChainArgs c{w.xa, w.ga, 1.0, w.scal + 0, w.fbad + 4, w.tsum, w.tpre, w.bnd, w.len + 2, 100};
    if (hipError_t e = run_chain(c, 16, n_ub, st)) return e;
    hipLaunchKernelGGL(phi_out_kernel, dim3(g1), dim3(256), 0, st, f.r, f.phi, s, P, (const double *)w.ga);
    hipLaunchKernelGGL(phi0_empty_kernel, dim3(1), dim3(1), 0, st, f.r, f.phi, s, P);
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
