// a2: potential_calculate (src/cmc/cmc_utils.c:675-736) as an FP64 suffix scan.
//
// The reference's reverse recurrence
//     phi_k = phi_{k+1} - M_k (1/r_k - 1/r_{k+1}),   M_{k-1} = M_k - m_k/N_STAR,
// with phi_{N+1} = 0, r_{N+1} = SF_INFINITY and M_N = Mtotal, telescopes to
//     phi_k = Mtotal/r_{N+1} - S_k - M_k/r_k,
//     T_k = sum_{i>k} a_i,  S_k = sum_{i>k} a_i/r_i,  M_k = Mtotal - T_k,
// where a_i = m_i/N_STAR rounded as the reference rounds each decrement.  The two
// suffix sums are scanned in double-double, so each phi_k is the correctly rounded
// value of that closed form; the reference's sequential rounding differs from it by
// ~1e-13 relative at N = 4e6.
//
// Kernels: count (N_MAX), tile reduce, tile scan (1 block), tile down-sweep.
// HBM traffic per star: r and m read twice (reduce, down-sweep), phi written once:
// 40 B; algorithmic 24 B.
#include "kernels.hpp"
#include "dd.hpp"

namespace {
constexpr int POT_THREADS = 256;
constexpr int POT_ITEMS = 8;
constexpr int POT_TILE = POT_THREADS * POT_ITEMS;
constexpr int SCAN_THREADS = 1024;

__global__ void pot_init_kernel(DevScalars *s)
{
    s->first_inf = s->n_max + 1;
}

// cmc_utils.c:685-696: the count stops at the first k with !(r[k] < SF_INFINITY).
__global__ void pot_count_kernel(const double *__restrict__ r, DevScalars *s)
{
    int64_t k = 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n = s->n_max;
    if (k > n) return;
    if (!(r[k] < CMCG_SF_INFINITY)) atomicMin((unsigned long long *)&s->first_inf, (unsigned long long)k);
}

struct Sums { dd T, S, M; };

__device__ __forceinline__ Sums sums_zero() { return {dd_from(0.0), dd_from(0.0), dd_from(0.0)}; }
__device__ __forceinline__ Sums sums_add(Sums a, Sums b)
{
    return {dd_add(a.T, b.T), dd_add(a.S, b.S), dd_add(a.M, b.M)};
}
__device__ __forceinline__ void sums_store(double *p, Sums a)
{
    p[0] = a.T.hi; p[1] = a.T.lo; p[2] = a.S.hi; p[3] = a.S.lo; p[4] = a.M.hi; p[5] = a.M.lo;
}
__device__ __forceinline__ Sums sums_load(const double *p)
{
    return {{p[0], p[1]}, {p[2], p[3]}, {p[4], p[5]}};
}

__device__ __forceinline__ Sums elem(const double *__restrict__ r, const double *__restrict__ m, double nstar, int64_t k)
{
    double mk = m[k];
    double a = mk / nstar;                  // star_m[k] / clus.N_STAR
    Sums e;
    e.T = dd_from(a);
    e.S = dd_div_d(a, r[k]);
    e.M = dd_from(mk);
    return e;
}

// Block-wide exclusive suffix scan over threads (thread t receives the sum of
// threads t+1..nthreads-1) through LDS.
template <int NT>
__device__ Sums block_excl_suffix(Sums v, double *lds)
{
    const int t = threadIdx.x;
    sums_store(lds + 6 * t, v);
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {
        Sums cur = sums_load(lds + 6 * t);
        Sums add = sums_zero();
        bool has = t + off < NT;
        if (has) add = sums_load(lds + 6 * (t + off));
        __syncthreads();
        if (has) sums_store(lds + 6 * t, sums_add(cur, add));
        __syncthreads();
    }
    Sums ex = sums_zero();
    if (t + 1 < NT) ex = sums_load(lds + 6 * (t + 1));
    __syncthreads();
    return ex;
}

__global__ __launch_bounds__(POT_THREADS) void pot_reduce_kernel(const double *__restrict__ r,
                                                                  const double *__restrict__ m,
                                                                  DevScalars *s, double nstar,
                                                                  double *__restrict__ tile_sums)
{
    __shared__ double lds[6 * POT_THREADS];
    int64_t n = s->first_inf - 1;
    int64_t base = 1 + (int64_t)blockIdx.x * POT_TILE;
    if (base > n) return;                      // whole block exits together
    Sums acc = sums_zero();
    int64_t k0 = base + (int64_t)threadIdx.x * POT_ITEMS;
    for (int u = 0; u < POT_ITEMS; u++) {
        int64_t k = k0 + u;
        if (k <= n) acc = sums_add(acc, elem(r, m, nstar, k));
    }
    Sums ex = block_excl_suffix<POT_THREADS>(acc, lds);
    if (threadIdx.x == 0) {
        Sums tot = sums_add(acc, ex);
        sums_store(tile_sums + 6 * blockIdx.x, tot);
        if (isnan(tot.M.hi)) dev_set_error(s, CMCG_E_NAN, base);
    }
}

// One block: exclusive suffix scan over tile totals, Mtotal, Rtidal, sentinels.
__global__ __launch_bounds__(SCAN_THREADS) void pot_scan_kernel(double *__restrict__ tile_sums, DevScalars *s,
                                                                 KParams P, double *__restrict__ r,
                                                                 double *__restrict__ phi)
{
    __shared__ double lds[6 * SCAN_THREADS];
    int64_t n = s->first_inf - 1;
    int64_t ntiles = (n + POT_TILE - 1) / POT_TILE;
    int64_t chunk = (ntiles + SCAN_THREADS - 1) / SCAN_THREADS;
    int64_t t0 = (int64_t)threadIdx.x * chunk;
    Sums acc = sums_zero();
    for (int64_t b = t0; b < t0 + chunk && b < ntiles; b++) acc = sums_add(acc, sums_load(tile_sums + 6 * b));
    Sums ex = block_excl_suffix<SCAN_THREADS>(acc, lds);
    // write exclusive suffix prefixes, walking each chunk from its high end
    Sums run = ex;
    int64_t hiend = t0 + chunk < ntiles ? t0 + chunk : ntiles;
    for (int64_t b = hiend - 1; b >= t0; b--) {
        Sums tot = sums_load(tile_sums + 6 * b);
        sums_store(tile_sums + 6 * b, run);
        run = sums_add(run, tot);
    }
    if (threadIdx.x == 0) {
        Sums total = sums_add(acc, ex);
        double mprev = dd_round(total.M);
        if (isnan(mprev)) dev_set_error(s, CMCG_E_NAN, 0);
        s->n_max = n;
        double Mtotal = mprev * P.madhoc + P.cenma_m * P.madhoc;
        s->Mtotal = Mtotal;
        s->Rtidal = P.orbit_r * pow(Mtotal, 1.0 / 3.0);
        r[n + 1] = CMCG_SF_INFINITY;
        phi[n + 1] = 0.0;
        if (n == 0) phi[0] = phi[1] + P.cenma_m * P.madhoc / r[1];
    }
}

__global__ __launch_bounds__(POT_THREADS) void pot_down_kernel(const double *__restrict__ r,
                                                                const double *__restrict__ m,
                                                                double *__restrict__ phi, DevScalars *s,
                                                                KParams P, const double *__restrict__ tile_sums)
{
    __shared__ double lds[6 * POT_THREADS];
    int64_t n = s->first_inf - 1;
    int64_t base = 1 + (int64_t)blockIdx.x * POT_TILE;
    if (base > n) return;
    double nstar = (double)P.n_star;
    double Mtotal = s->Mtotal;
    Sums e[POT_ITEMS];
    Sums acc = sums_zero();
    int64_t k0 = base + (int64_t)threadIdx.x * POT_ITEMS;
    for (int u = 0; u < POT_ITEMS; u++) {
        int64_t k = k0 + u;
        if (k <= n) e[u] = elem(r, m, nstar, k);
        else e[u] = sums_zero();
        acc = sums_add(acc, e[u]);
    }
    Sums run = sums_add(block_excl_suffix<POT_THREADS>(acc, lds), sums_load(tile_sums + 6 * blockIdx.x));
    dd C = dd_div_d(Mtotal, CMCG_SF_INFINITY);     // M_N / r_{N+1}
    for (int u = POT_ITEMS - 1; u >= 0; u--) {
        int64_t k = k0 + u;
        if (k <= n) {
            double rk = r[k];
            dd Mk = dd_add(dd_from(Mtotal), dd_neg(run.T));
            dd Mr = dd_div_dd_d(Mk, rk);
            dd ph = dd_add(dd_add(C, dd_neg(run.S)), dd_neg(Mr));
            double p = dd_round(ph);
            phi[k] = p;
            if (isnan(p)) dev_set_error(s, CMCG_E_NAN, k);
            if (k == 1) phi[0] = p + P.cenma_m * P.madhoc / rk;     // cmc_utils.c:729
        }
        run = sums_add(run, e[u]);
    }
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
} // namespace

hipError_t launch_potential(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub,
                            const Scratch &sc, hipStream_t st)
{
    int64_t ntiles = (n_ub + POT_TILE - 1) / POT_TILE;
    if (ntiles < 1) ntiles = 1;
    if (ntiles > sc.max_tiles || ntiles > (int64_t)SCAN_THREADS * 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pot_init_kernel, dim3(1), dim3(1), 0, st, s);
    if (n_ub > 0)
        hipLaunchKernelGGL(pot_count_kernel, dim3((unsigned)((n_ub + 255) / 256)), dim3(256), 0, st, f.r, s);
    hipLaunchKernelGGL(pot_reduce_kernel, dim3((unsigned)ntiles), dim3(POT_THREADS), 0, st, f.r, f.m, s,
                       (double)P.n_star, sc.tile_sums);
    hipLaunchKernelGGL(pot_scan_kernel, dim3(1), dim3(SCAN_THREADS), 0, st, sc.tile_sums, s, P, f.r, f.phi);
    hipLaunchKernelGGL(pot_down_kernel, dim3((unsigned)ntiles), dim3(POT_THREADS), 0, st, f.r, f.m, f.phi, s, P,
                       sc.tile_sums);
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
