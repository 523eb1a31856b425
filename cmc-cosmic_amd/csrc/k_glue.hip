// Henon-step glue around the hot path (src/cmc/cmc.c:240-407 order):
// energy_conservation1/2 (src/cmc/cmc_utils.c:2153-2227), tidally_strip_stars
// (src/cmc/cmc_evolution_thr.c:443-646, TIDAL_TREATMENT 0), ComputeIntermediateEnergy
// (cmc_utils.c:368-391), set_velocities3 (:254-354) and ComputeEnergy (:604-653).
// All are per-star streaming maps plus deterministic two-level reductions.
#include "kernels.hpp"

namespace {
constexpr int GT = 256;
__device__ __forceinline__ double sqr(double x) { return x * x; }

template <int NV>
__device__ void block_sum_store(double (&v)[NV], double *__restrict__ out)
{
    __shared__ double lds[NV][GT];
    const int t = threadIdx.x;
    for (int i = 0; i < NV; i++) lds[i][t] = v[i];
    __syncthreads();
    for (int off = GT / 2; off > 0; off >>= 1) {
        if (t < off)
            for (int i = 0; i < NV; i++) lds[i][t] += lds[i][t + off];
        __syncthreads();
    }
    if (t == 0)
        for (int i = 0; i < NV; i++) out[i] = lds[i][0];
}

template <int NV>
__device__ void final_sum(const double *__restrict__ partials, int64_t nblocks, double (&res)[NV])
{
    __shared__ double lds[NV][GT];
    const int t = threadIdx.x;
    double acc[NV];
    for (int i = 0; i < NV; i++) acc[i] = 0.0;
#pragma unroll 8
    for (int64_t b = t; b < nblocks; b += GT)      // unrolled: independent loads in flight
        for (int i = 0; i < NV; i++) acc[i] += partials[8 * b + i];
    for (int i = 0; i < NV; i++) lds[i][t] = acc[i];
    __syncthreads();
    for (int off = GT / 2; off > 0; off >>= 1) {
        if (t < off)
            for (int i = 0; i < NV; i++) lds[i][t] += lds[i][t + off];
        __syncthreads();
    }
    for (int i = 0; i < NV; i++) res[i] = lds[i][0];
}

// cmc.c:246-258, 278: per-step resets
struct step_begin_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s)
{
    s->Eescaped_old = s->Eescaped;
    s->Jescaped_old = s->Jescaped;
    s->TidalMassLoss_old = s->TidalMassLoss;
    s->Etidal_old = s->Etidal;
    s->Eescaped = 0.0; s->Jescaped = 0.0; s->TidalMassLoss = 0.0; s->Etidal = 0.0;
    s->n_max_new = s->n_max;
    s->dtrel = __longlong_as_double(0x7ff0000000000000ll);     // the time step's minimum starts at +inf
}
};

struct after_dynamics_kernel { static constexpr int LB = 1024, MB = 0; static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s) { s->Prev_Dt = s->Dt; } };   // cmc.c:297

struct ec1_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P)
{
    int64_t i = 1 + (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (i > s->n_max) return;
    double ri = f.r[i];
    f.vtold[i] = f.vt[i];
    f.vrold[i] = f.vr[i];
    f.Uoldrold[i] = f.phi[i] + phi_s_own(ri, f.m[i], P.inv_nstar, ri);
}
};

// energy_conservation2: Uoldrnew = potential(rnew) + MPI_PHI_S(rnew, i).  The orbit
// phase already evaluated potential(rnew) for sampled stars (flag 256) on the same
// arrays, so only circular / removed stars search here.
struct ec2_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P)
{
    int64_t i = 1 + (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    int64_t n = s->n_max;
    if (i > n) return;
    double x = f.rnew[i];
    double pot;
    int fl = f.flags[i];
    if (fl & 256) {
        pot = f.pot_rnew[i];
    } else {
        bool bad = false;
        pot = potential_at(f.r, f.phi, n, x, 1, n + 1, &bad);
        if (bad) dev_set_error(s, CMCG_E_BRACKET, i);
        f.pot_rnew[i] = pot;
        f.flags[i] = fl | 256;
    }
    f.Uoldrnew[i] = pot + phi_s_own(f.r[i], f.m[i], P.inv_nstar, x);
}
};

// The step's form of energy_conservation2 followed by ComputeIntermediateEnergy's
// first loop (cmc_utils.c:335-351): the stripping between them leaves every field
// that EI reads unchanged for the stars it keeps and sets EI = 0 for the stars it
// strips, so EI can be formed here, before it, with the same result.
// It also forms D = vold^2 + 2 (1 - q) (Uoldrold - Uoldrnew), the part of
// set_velocities3's vnew^2 (cmc_utils.c:242-247) that the sort then carries in place of
// its four inputs: the same operations in the same order, so the same double.
struct ec2_ie_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P, double *__restrict__ dpre)
{
    int64_t i = 1 + (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    int64_t n = s->n_max;
    if (i > n) return;
    const double x = f.rnew[i];
    double pot;
    const int fl = f.flags[i];
    if (fl & 256) {
        pot = f.pot_rnew[i];
    } else {
        bool bad = false;
        pot = potential_at(f.r, f.phi, n, x, 1, n + 1, &bad);
        if (bad) dev_set_error(s, CMCG_E_BRACKET, i);
        f.pot_rnew[i] = pot;
        f.flags[i] = fl | 256;
    }
    const double Uoldrnew = pot + phi_s_own(f.r[i], f.m[i], P.inv_nstar, x);
    f.Uoldrnew[i] = Uoldrnew;
    if (x < 1.0e6) f.EI[i] = sqr(f.vr[i]) + sqr(f.vt[i]) + f.phi[i] - pot;
    const double q = 0.5;
    const double vtold = f.vtold[i], vrold = f.vrold[i];
    const double vold2 = vtold * vtold + vrold * vrold;
    dpre[i] = vold2 + 2.0 * (1.0 - q) * (f.Uoldrold[i] - Uoldrnew);
}
};

struct strip_begin_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s, int32_t *done)
{
    s->Etidal = 0.0;        // cmc_evolution_thr.c:449
    *done = 0;
}
};

// One pass of the stripping loop (cmc_evolution_thr.c:447-645) is two kernels: a
// fixed grid in which every block derives Rtidal from the scalars left by the
// previous pass, strips its stars (grid-stride) and stores its partial sums, then one
// block that reduces them in block order, updates the scalars and decides whether the
// loop ends.  Passes after the end return at once, so a fixed number of passes can be
// enqueued (or captured in a graph) without a host round trip.
constexpr int STRIP_BLOCKS = 1024;
constexpr int STRIP_TAIL_PASSES = 64;

// Block bi's share of one pass (stars 1 + bi*GT + t, stride nb*GT): strip them and
// store the block's partial sums (Eescaped, Jescaped, DTidalMassLoss, Etidal).
__device__ void strip_block(DevFields &f, const KParams &P, double Rt, int64_t n, double *__restrict__ partials,
                            unsigned nb, unsigned bi)
{
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = 1 + (int64_t)bi * blockDim.x + threadIdx.x; i <= n; i += (int64_t)nb * blockDim.x) {
        if (f.r_apo[i] > Rt && f.rnew[i] < 1000000) {
            double m = f.m[i], E = f.E[i], J = f.J[i];
            f.rnew[i] = CMCG_SF_INFINITY; f.vrnew[i] = 0.0; f.vtnew[i] = 0.0;
            v[0] += E * m / (double)P.n_star;   // Eescaped
            v[1] += J * m / (double)P.n_star;   // Jescaped
            v[2] += m / (double)P.n_star;       // DTidalMassLoss
            v[3] += E * m / (double)P.n_star;   // Etidal
            // destroy_obj
            f.vr[i] = 0.0; f.vt[i] = 0.0; f.E[i] = 0.0; f.J[i] = 0.0; f.id[i] = 0;
            f.X[i] = 0.0; f.Y[i] = 0.0; f.r_peri[i] = 0.0; f.r_apo[i] = 0.0;
            f.rOld[i] = 0.0; f.Uoldrold[i] = 0.0; f.Uoldrnew[i] = 0.0; f.vrold[i] = 0.0; f.vtold[i] = 0.0;
            f.EI[i] = 0.0;
            f.m[i] = 2.2250738585072014e-308;
            f.flags[i] |= CMCG_SF_STRIPPED;
        }
    }
    block_sum_store<4>(v, partials + 8 * bi);
}

__device__ __forceinline__ double strip_rtidal(const DevScalars *s, const KParams &P)
{
    return P.orbit_r * pow(s->Mtotal - (s->TidalMassLoss - s->OldTidalMassLoss), 1.0 / 3.0);
}

// The end of a pass: the block partials summed in block order, the scalars updated,
// the loop's exit decided (all threads of the calling block take part).
__device__ void strip_finish(DevScalars *s, double Rt, int32_t *done, const double *__restrict__ partials,
                             unsigned nb)
{
    double res[4];
    final_sum<4>(partials, nb, res);
    if (threadIdx.x == 0) {
        s->Rtidal = Rt;
        s->Eescaped += res[0];
        s->Jescaped += res[1];
        s->Etidal += res[3];
        s->TidalMassLoss += res[2];
        s->DTidalMassLoss = res[2];
        if (!(res[2] > 0)) *done = 1;
    }
}

struct strip_pass_kernel {
    static constexpr int LB = GT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P, const int32_t *done,
                                               double *__restrict__ partials)
    {
        if (*done) return;
        strip_block(f, P, strip_rtidal(s, P), s->n_max, partials, kg_.nb, kg_.bid);
    }
};

struct strip_end_kernel {
    static constexpr int LB = GT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s, KParams P, int32_t *done,
                                               const double *__restrict__ partials, unsigned nb)
    {
        if (*done) return;
        strip_finish(s, strip_rtidal(s, P), done, partials, nb);
    }
};

// The rest of the loop in one block, for a captured step: after the grid passes the
// loop has nearly always ended and this returns at once; otherwise it runs further
// passes here, at most STRIP_TAIL_PASSES (then the step reports CMCG_E_STRIP).  The
// block plays the nb blocks of a grid pass one after another and writes the same
// partials, so the sums add in the same order as in the eager loop.
struct strip_tail_kernel {
    static constexpr int LB = GT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P, int32_t *done,
                                                        double *__restrict__ partials, unsigned nb)
{
    __shared__ int32_t sdone;
    for (int pass = 0; pass < STRIP_TAIL_PASSES; pass++) {
        if (threadIdx.x == 0) sdone = *(volatile int32_t *)done;
        __syncthreads();
        if (sdone) return;
        const double Rt = strip_rtidal(s, P);
        const int64_t n = s->n_max;
        __syncthreads();
        for (unsigned bi = 0; bi < nb; bi++) strip_block(f, P, Rt, n, partials, nb, bi);
        __threadfence_block();
        __syncthreads();
        strip_finish(s, Rt, done, partials, nb);
        __threadfence();
        __syncthreads();
    }
    if (threadIdx.x == 0 && !*(volatile int32_t *)done) dev_set_error(s, CMCG_E_STRIP, 0);
}
};

// ComputeIntermediateEnergy, first loop (EI).
struct ie_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s)
{
    int64_t j = 1 + (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    int64_t n = s->n_max;
    if (j > n) return;
    double x = f.rnew[j];
    if (x < 1.0e6) {
        double pot;
        if (f.flags[j] & 256) pot = f.pot_rnew[j];
        else {
            bool bad = false;
            pot = potential_at(f.r, f.phi, n, x, 1, n + 1, &bad);
            if (bad) dev_set_error(s, CMCG_E_BRACKET, j);
        }
        f.EI[j] = sqr(f.vr[j]) + sqr(f.vt[j]) + f.phi[j] - pot;
    }
}
};

// ComputeIntermediateEnergy, second loop (transfer of the new orbit variables).
struct ie_transfer_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s)
{
    int64_t j = 1 + (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (j > s->n_max) return;
    f.rOld[j] = f.r[j];
    f.r[j] = f.rnew[j];
    f.vr[j] = f.vrnew[j];
    f.vt[j] = f.vtnew[j];
}
};

// set_velocities3 per-star part.  Stars with vnew2 <= 0 ("events") carry excess
// energy that the following stars absorb (cmc_utils.c:268-310); they are counted
// per block for an ordered compaction.
// potential(r) of the new potential at every old node.  The old radii (the array the
// sort read, left intact in the alternate buffer) are sorted, so a block's MERGE_BLK
// consecutive old radii [x_first, x_last] fall into one contiguous run of new nodes:
// two lanes find its ends by bisection, the block stages the run in LDS (coalesced)
// and every radius is placed by bisection in LDS.  A run longer than MERGE_LDS nodes
// is searched in global memory within the same ends.  The interval is the largest
// k with r_k < x, as in potential_at (the search predicate is monotone, so any bracket
// holding that k gives the same k), and the interpolation is the same expression, so
// the values equal potential_at(x) bit for bit.
constexpr int MERGE_T = 256, MERGE_PER = 2, MERGE_BLK = MERGE_T * MERGE_PER, MERGE_LDS = 1024;

// The node index of both ends of every merge block (slot 2b: the block's first old
// radius, 2b + 1: its last), one search per thread, so the merge blocks start from
// their brackets instead of each waiting on two full-range searches.
struct merge_ends_kernel {
    static constexpr int LB = 256, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ r_old,
                                               const double2 *__restrict__ nodes, const DevScalars *s,
                                               int64_t *__restrict__ ends)
    {
        const int64_t q = (int64_t)kg_.bid * 256 + threadIdx.x;       // 2 b + (0: first, 1: last)
        const int64_t n_old = s->sort_in, n = s->n_max;
        const int64_t first = 1 + (q >> 1) * MERGE_BLK;
        if (first > n_old) return;
        const int64_t last = first + MERGE_BLK - 1 <= n_old ? first + MERGE_BLK - 1 : n_old;
        const PackedNodes nd{nodes};
        const double x = r_old[(q & 1) ? last : first];
        // largest k in [1, n+1] with r_k < x (1 if none); monotone in x
        ends[q] = (x == 0.0 || x < nd.R(1)) ? 1 : find_zero_r_n(nd, 1, n + 1, x);
    }
};

struct old_node_potential_kernel {
    static constexpr int LB = MERGE_T, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ r_old,
                                                                     const double2 *__restrict__ nodes,
                                                                     DevScalars *s, double2 *__restrict__ u_old,
                                                                     const int64_t *__restrict__ ends)
{
    __shared__ double2 L[MERGE_LDS];
    const int64_t n_old = s->sort_in, n = s->n_max;
    const int64_t first = 1 + (int64_t)kg_.bid * MERGE_BLK;
    if (first > n_old) return;
    const int64_t last = first + MERGE_BLK - 1 <= n_old ? first + MERGE_BLK - 1 : n_old;
    const PackedNodes nd{nodes};
    const double r1 = nd.R(1);
    const int t = threadIdx.x;
    const int64_t lo = ends[2 * (int64_t)kg_.bid], hi = ends[2 * (int64_t)kg_.bid + 1];
    const double xf = r_old[first], xl = r_old[last];
    const int64_t span = hi + 2 - lo;             // nodes lo .. hi+1
    const bool staged = span <= MERGE_LDS && hi + 1 <= n + 1;
    if (staged)
        for (int64_t k = t; k < span; k += MERGE_T) L[k] = nodes[lo + k];
    __syncthreads();
    for (int u = 0; u < MERGE_PER; u++) {
        const int64_t sl = first + t + (int64_t)u * MERGE_T;
        if (sl > last) break;
        const double x = r_old[sl];
        double v;
        if (x == 0.0 || x < r1 || !(x >= xf && x <= xl)) {    // the latter only if unsorted
            bool bad = false;
            v = potential_at_n(nd, n, x, 1, n + 1, &bad);
            if (bad) dev_set_error(s, CMCG_E_BRACKET, sl);
        } else {
            int64_t i;
            double ri, pi, ri1, pi1;
            if (staged) {
                int a = 0, b = (int)(hi - lo);
                while (a < b) {
                    const int m = (a + b + 1) >> 1;
                    if (L[m].x < x) a = m;
                    else b = m - 1;
                }
                i = lo + a;
                ri = L[a].x; pi = L[a].y; ri1 = L[a + 1].x; pi1 = L[a + 1].y;
            } else {
                i = lo == hi ? lo : find_zero_r_n(nd, lo, hi, x);
                nd.RP(i, ri, pi);
                if (i + 1 <= n + 1) nd.RP(i + 1, ri1, pi1);
                else { ri1 = ri; pi1 = pi; }
            }
            if (i > n) { dev_set_error(s, CMCG_E_BRACKET, sl); v = 0.0; }
            else v = pi + (pi1 - pi) * (1.0 / ri - 1.0 / x) / (1.0 / ri - 1.0 / ri1);
        }
        u_old[sl] = make_double2(v, x);       // with the radius: set_velocities3 reads both at once
    }
}
};

constexpr int SV3_COARSE = 1024;      // LDS samples of r for potential(rOld)

// With u_old (the step): potential(rOld) and rOld itself through the sort's permutation,
// and D (dpost) in place of vtold, vrold, Uoldrold and Uoldrnew.  Without: the
// reference's fields and a search (the entry point of the reference's name).
struct sv3_kernel {
    static constexpr int LB = GT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P, double *__restrict__ excess_c,
                                                 int64_t *__restrict__ blk_cnt, const double2 *__restrict__ mid,
                                                 const double2 *__restrict__ nodes, const double2 *__restrict__ u_old,
                                                 const uint32_t *__restrict__ perm, const double *__restrict__ dpost)
{
    __shared__ int cnt;
    __shared__ double cr[SV3_COARSE + 1];
    const double q = 0.5;
    int64_t i = 1 + (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    int64_t n = s->n_max;
    Coarse C;
    C.cr = cr; C.cphi = nullptr; C.cs = SV3_COARSE; C.n1 = n + 1; C.hint = false;
    C.mid = mid;
    C.stride = coarse_stride(n, SV3_COARSE);
    if (!u_old)      // the samples serve only the search for potential(rOld)
        for (int sI = threadIdx.x; sI <= SV3_COARSE; sI += blockDim.x) cr[sI] = f.r[coarse_k(C, sI)];
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    if (i <= n) {
        double m = f.m[i], ri = f.r[i];
        double Unewrold, D;
        if (u_old) {
            const double2 ux = u_old[perm[i - 1]];
            Unewrold = ux.x + phi_s_own(ri, m, P.inv_nstar, ux.y);
            D = dpost[i];
        } else {
            const double rO = f.rOld[i];
            bool bad = false;
            Unewrold = potential_at_coarse_n(PackedNodes{nodes}, n, rO, 1, n + 1, C, &bad) +
                       phi_s_own(ri, m, P.inv_nstar, rO);
            if (bad) dev_set_error(s, CMCG_E_BRACKET, i);
            const double vtold = f.vtold[i], vrold = f.vrold[i];
            const double vold2 = vtold * vtold + vrold * vrold;
            D = vold2 + 2.0 * (1.0 - q) * (f.Uoldrold[i] - f.Uoldrnew[i]);
        }
        double Unewrnew = f.phi[i] + phi_s_own(ri, m, P.inv_nstar, ri);
        double vnew2 = D + 2.0 * q * (Unewrold - Unewrnew);
        double vr = f.vr[i], vt = f.vt[i];
        if (vnew2 <= 0.0) {
            excess_c[i] = 0.5 * (sqr(vr) + sqr(vt) - vnew2) * m;
            atomicAdd(&cnt, 1);
        } else {
            excess_c[i] = -1.0;
            double alpha = sqrt(vnew2 / (sqr(vr) + sqr(vt)));
            vr *= alpha;
            vt *= alpha;
            f.vr[i] = vr; f.vt[i] = vt;
            f.vrnew[i] = vr; f.vtnew[i] = vt;      // post-scaling values, read by the chain
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) blk_cnt[kg_.bid] = cnt;
}
};

// Exclusive scan of the per-block event counts, in place, and their total.  Each thread
// takes a contiguous chunk, held in registers when it has at most SV3_CHUNK entries (so
// the write-back waits on no load), the chunk sums are scanned by wave shuffles and the
// 16 wave totals through LDS.
constexpr int SV3_CHUNK = 16;
struct sv3_scan_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s, int64_t *__restrict__ blk_cnt, int64_t nb)
{
    __shared__ int64_t wtot[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t chunk = (nb + 1023) / 1024;
    const int64_t b0 = (int64_t)t * chunk;
    const bool regs = chunk <= SV3_CHUNK;
    int64_t v[SV3_CHUNK];
    int64_t acc = 0;
    if (regs) {
#pragma unroll
        for (int q = 0; q < SV3_CHUNK; q++) {
            v[q] = q < chunk && b0 + q < nb ? blk_cnt[b0 + q] : 0;
            acc += v[q];
        }
    } else {
        for (int64_t b = b0; b < b0 + chunk && b < nb; b++) acc += blk_cnt[b];
    }
    int64_t inc = acc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) wtot[w] = inc;
    __syncthreads();
    int64_t run = inc - acc, tot = 0;
    for (int q = 0; q < 16; q++) {
        if (q < w) run += wtot[q];
        tot += wtot[q];
    }
    if (t == 1023) s->n_excess = tot;
    if (regs) {
#pragma unroll
        for (int q = 0; q < SV3_CHUNK; q++)
            if (q < chunk && b0 + q < nb) {
                blk_cnt[b0 + q] = run;
                run += v[q];
            }
    } else {
        for (int64_t b = b0; b < b0 + chunk && b < nb; b++) {
            const int64_t c = blk_cnt[b];
            blk_cnt[b] = run;
            run += c;
        }
    }
}
};

// ordered list of the event stars
struct sv3_compact_kernel {
    static constexpr int LB = GT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s, const double *__restrict__ excess_c,
                                                         const int64_t *__restrict__ blk_off,
                                                         int64_t *__restrict__ events)
{
    __shared__ int wcnt[GT / 64];
    int64_t i = 1 + (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    bool ev = i <= s->n_max && excess_c[i] >= 0.0;
    uint64_t bal = __ballot(ev);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wcnt[w] = __popcll(bal);
    __syncthreads();
    int64_t pos = blk_off[kg_.bid];
    for (int w2 = 0; w2 < w; w2++) pos += wcnt[w2];
    pos += __popcll(bal & ((1ull << lane) - 1ull));
    if (ev) events[pos] = i;
}
};

constexpr int64_t SPEC_WALK_MAX = 4096;

// One episode of the excess chain starting at event star `i` with no excess
// carried into it: event stars add their excess, the other stars absorb
// min(0.8 KE, excess) each (cmc_utils.c:268-310).  Returns the star at which the
// excess reaches 0, n + 1 when excess is left at the end (dropped at one rank,
// cmc_utils.c:313-353), or -1 after `max_steps` stars.
__device__ int64_t excess_walk(const DevFields &f, const double *__restrict__ excess_c, int64_t i, int64_t n,
                               bool write, int64_t max_steps)
{
    const double E_dump_factor = 0.8;
    double Eexcess = 0.0;
    for (int64_t steps = 0; i <= n; i++, steps++) {
        if (steps >= max_steps) return -1;
        double c = excess_c[i];
        if (c >= 0.0) {
            Eexcess += c;
        } else if (Eexcess > 0) {
            double m = f.m[i], vr = f.vrnew[i], vt = f.vtnew[i];
            double v2 = sqr(vr) + sqr(vt);
            double E_dump_capacity = E_dump_factor * 0.5 * m * v2;
            double E_dump = E_dump_capacity <= Eexcess ? E_dump_capacity : Eexcess;
            if (write) {
                double exc_ratio = sqrt((v2 - 2 * E_dump / m) / v2);
                f.vr[i] = vr * exc_ratio;
                f.vt[i] = vt * exc_ratio;
            }
            Eexcess -= E_dump;
        }
        if (!(Eexcess > 0)) return i;
    }
    return n + 1;
}

// Fresh (no carried excess) bounded walk of every event.  joins[d] = 1 when that
// walk reaches event d + 1 or was cut off, so event d + 1 is presumed to lie inside
// the same episode.  An episode starting at a true head is exactly its fresh walk.
struct sv3_spec_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, const double *__restrict__ excess_c,
                                const int64_t *__restrict__ events, int64_t *__restrict__ wend,
                                int32_t *__restrict__ joins)
{
    const int64_t ne = s->n_excess;
    for (int64_t d = (int64_t)kg_.bid * blockDim.x + threadIdx.x; d < ne; d += (int64_t)kg_.nb * blockDim.x) {
        int64_t e = excess_walk(f, excess_c, events[d], s->n_max, false, SPEC_WALK_MAX);
        wend[d] = e;
        joins[d] = (e < 0) || (d + 1 < ne && e >= events[d + 1]);
    }
}
};

// Presumed heads are d = 0 and every d with joins[d-1] = 0; the presumed group of
// head d runs to the first q >= d with joins[q] = 0.  If every presumed head's walk
// ends inside its group's span [events[q], events[q+1]), then by induction from
// event 0 (which starts with no carried excess) each presumed head is a true head
// and its walk is the true episode.  Otherwise the heads are relinked exactly.
struct sv3_verify_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s, const int64_t *__restrict__ events,
                                  const int32_t *__restrict__ joins, const int64_t *__restrict__ wend,
                                  int64_t *__restrict__ relink)
{
    const int64_t ne = s->n_excess;
    for (int64_t d = (int64_t)kg_.bid * blockDim.x + threadIdx.x; d < ne; d += (int64_t)kg_.nb * blockDim.x) {
        if (d > 0 && joins[d - 1]) continue;
        int64_t q = d;
        while (q + 1 < ne && joins[q]) q++;
        int64_t e = wend[d];
        if (e < 0 || e < events[q] || (q + 1 < ne && e >= events[q + 1]))
            atomicMin((unsigned long long *)relink, (unsigned long long)d);
    }
}
};

// Exact head chain in event order: the next head is the first event after the end
// of the current episode.  Runs only when a presumed head failed the check, and
// starts at the first one that did: every presumed head before it passed, so by
// induction it is a true head and the groups before it are exact.
struct sv3_relink_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, const double *__restrict__ excess_c,
                                  const int64_t *__restrict__ events, const int64_t *__restrict__ wend,
                                  int32_t *__restrict__ joins, const int64_t *__restrict__ relink)
{
    if (threadIdx.x != 0 || kg_.bid != 0) return;
    int64_t n = s->n_max, ne = s->n_excess;
    int64_t d = *relink;
    if (d >= ne) return;
    while (d < ne) {
        int64_t end = wend[d];
        if (end < 0) end = excess_walk(f, excess_c, events[d], n, false, INT64_MAX);
        int64_t d2 = d + 1;
        while (d2 < ne && events[d2] <= end) joins[d2 - 1] = 1, d2++;
        if (d2 < ne) joins[d2 - 1] = 0;
        d = d2;
    }
}
};

// every head walks its episode and writes the dumped velocities
struct sv3_heads_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, const double *__restrict__ excess_c,
                                 const int64_t *__restrict__ events, const int32_t *__restrict__ joins)
{
    const int64_t ne = s->n_excess;
    for (int64_t d = (int64_t)kg_.bid * blockDim.x + threadIdx.x; d < ne; d += (int64_t)kg_.nb * blockDim.x)
        if (d == 0 || !joins[d - 1]) excess_walk(f, excess_c, events[d], s->n_max, true, INT64_MAX);
}
};

struct sv3_begin_kernel { static constexpr int LB = 1024, MB = 0; static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s, int64_t *relink) { s->n_excess = 0; *relink = INT64_MAX; } };

struct energy_kernel {
    static constexpr int LB = GT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P,
                                                    double *__restrict__ partials)
{
    int64_t i = 1 + (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    int64_t n = s->n_max;
    if (i <= n) {
        double vr = f.vr[i], vt = f.vt[i], ph = f.phi[i], m = f.m[i];
        f.E[i] = ph + 0.5 * (sqr(vr) + sqr(vt));
        f.J[i] = f.r[i] * vt;
        v[0] = 0.5 * (sqr(vr) + sqr(vt)) * m * P.madhoc;
        v[1] = ph * m * P.madhoc + f.phi[0] * P.cenma_m * P.madhoc / n;
    }
    block_sum_store<2>(v, partials + 8 * kg_.bid);
}
};

// compute_energy_new (cmc_utils.c:2017-2028) + the totals of ComputeEnergy
struct energy_end_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s, const double *__restrict__ partials, int64_t nblocks)
{
    double res[2];
    final_sum<2>(partials, nblocks, res);
    if (threadIdx.x == 0) {
        s->Eescaped += s->Eescaped_old;
        s->Jescaped += s->Jescaped_old;
        s->Etidal += s->Etidal_old;
        double K = res[0], Pt = res[1] * 0.5;
        s->Ekin = K;
        s->Epot = Pt;
        s->Etot = K + Pt + 0.0 + 0.0;
        s->Etot += 0.0 + s->Eescaped + 0.0 + 0.0;
    }
}
};

unsigned grid_for(int64_t n) { return (unsigned)((n + GT - 1) / GT > 0 ? (n + GT - 1) / GT : 1); }
} // namespace

hipError_t launch_step_begin(DevScalars *s, hipStream_t st)
{
    kl<step_begin_kernel>(dim3(1), dim3(1), 0, st, s);
    return hipGetLastError();
}
hipError_t launch_after_dynamics(DevScalars *s, hipStream_t st)
{
    kl<after_dynamics_kernel>(dim3(1), dim3(1), 0, st, s);
    return hipGetLastError();
}
hipError_t launch_energy_conservation1(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub,
                                       hipStream_t st)
{
    kl<ec1_kernel>(dim3(grid_for(n_ub)), dim3(GT), 0, st, f, s, P);
    return hipGetLastError();
}
hipError_t launch_energy_conservation2(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub,
                                       hipStream_t st, bool with_ie, double *dpre)
{
    if (with_ie && !dpre) return hipErrorInvalidValue;
    if (with_ie) kl<ec2_ie_kernel>(dim3(grid_for(n_ub)), dim3(GT), 0, st, f, s, P, dpre);
    else kl<ec2_kernel>(dim3(grid_for(n_ub)), dim3(GT), 0, st, f, s, P);
    return hipGetLastError();
}
hipError_t launch_strip(const DevFields &f, DevScalars *s, const KParams &P, const Scratch &sc, int32_t *done,
                        int64_t n_ub, int passes, bool begin, hipStream_t st)
{
    (void)n_ub;
    // a fixed grid (grid-stride over the stars) that the per-block partials hold
    const unsigned nb = STRIP_BLOCKS;              // the scratch holds at least STRIP_BLOCKS partials
    if (sc.max_blocks < STRIP_BLOCKS) return hipErrorInvalidValue;
    if (begin) kl<strip_begin_kernel>(dim3(1), dim3(1), 0, st, s, done);
    for (int it = 0; it < passes; it++) {
        kl<strip_pass_kernel>(dim3(nb), dim3(GT), 0, st, f, s, P, (const int32_t *)done, sc.partials);
        kl<strip_end_kernel>(dim3(1), dim3(GT), 0, st, s, P, done, (const double *)sc.partials, nb);
    }
    return hipGetLastError();
}

hipError_t launch_strip_tail(const DevFields &f, DevScalars *s, const KParams &P, const Scratch &sc, hipStream_t st)
{
    if (sc.max_blocks < STRIP_BLOCKS) return hipErrorInvalidValue;
    kl<strip_tail_kernel>(dim3(1), dim3(GT), 0, st, f, s, P, sc.strip_done, sc.partials,
                       (unsigned)STRIP_BLOCKS);
    return hipGetLastError();
}
hipError_t launch_intermediate_energy(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub,
                                      hipStream_t st)
{
    (void)P;
    kl<ie_kernel>(dim3(grid_for(n_ub)), dim3(GT), 0, st, f, s);
    return hipGetLastError();
}
hipError_t launch_ie_transfer(const DevFields &f, DevScalars *s, int64_t n_ub, hipStream_t st)
{
    kl<ie_transfer_kernel>(dim3(grid_for(n_ub)), dim3(GT), 0, st, f, s);
    return hipGetLastError();
}
hipError_t launch_set_velocities3(const DevFields &f, DevScalars *s, const KParams &P, const Scratch &sc,
                                  int64_t n_ub, hipStream_t st, const double *r_old_sorted, const double *dpost)
{
    if ((r_old_sorted != nullptr) != (dpost != nullptr)) return hipErrorInvalidValue;
    int64_t nb = grid_for(n_ub);
    if (nb > sc.max_blocks) return hipErrorInvalidValue;
    unsigned ge = grid_for(n_ub);
    if (ge > 256) ge = 256;            // the excess-list kernels stride over the list
    kl<sv3_begin_kernel>(dim3(1), dim3(1), 0, st, s, sc.sv3_fallback);
    // the step calls this right after the potential, which left the node table current
    if (!dpost)
        if (hipError_t e = launch_mid_build(f, s, sc, n_ub, st)) return e;
    if (r_old_sorted) {
        const int64_t mb = (n_ub + MERGE_BLK - 1) / MERGE_BLK;
        if (2 * mb > sc.merge_ends_cap) return hipErrorInvalidValue;
        kl<merge_ends_kernel>(dim3((unsigned)((2 * mb + 255) / 256)), dim3(256), 0, st, r_old_sorted,
                              (const double2 *)sc.nodes, (const DevScalars *)s, sc.merge_ends);
        kl<old_node_potential_kernel>(dim3((unsigned)mb), dim3(MERGE_T), 0, st, r_old_sorted,
                                      (const double2 *)sc.nodes, s, sc.u_old, (const int64_t *)sc.merge_ends);
    }
    kl<sv3_kernel>(dim3((unsigned)nb), dim3(GT), 0, st, f, s, P, sc.excess_c, sc.sv3_blk,
                       (const double2 *)sc.mid, (const double2 *)sc.nodes,
                       r_old_sorted ? (const double2 *)sc.u_old : nullptr, (const uint32_t *)sc.vals_a, dpost);
    kl<sv3_scan_kernel>(dim3(1), dim3(1024), 0, st, s, sc.sv3_blk, nb);
    kl<sv3_compact_kernel>(dim3((unsigned)nb), dim3(GT), 0, st, s, (const double *)sc.excess_c,
                       (const int64_t *)sc.sv3_blk, sc.excess_list);
    kl<sv3_spec_kernel>(dim3(ge), dim3(GT), 0, st, f, s, (const double *)sc.excess_c,
                       (const int64_t *)sc.excess_list, sc.sv3_wend, sc.sv3_merged);
    kl<sv3_verify_kernel>(dim3(ge), dim3(GT), 0, st, s, (const int64_t *)sc.excess_list,
                       (const int32_t *)sc.sv3_merged, (const int64_t *)sc.sv3_wend, sc.sv3_fallback);
    kl<sv3_relink_kernel>(dim3(1), dim3(64), 0, st, f, s, (const double *)sc.excess_c,
                       (const int64_t *)sc.excess_list, (const int64_t *)sc.sv3_wend, sc.sv3_merged,
                       (const int64_t *)sc.sv3_fallback);
    kl<sv3_heads_kernel>(dim3(ge), dim3(GT), 0, st, f, s, (const double *)sc.excess_c,
                       (const int64_t *)sc.excess_list, (const int32_t *)sc.sv3_merged);
    return hipGetLastError();
}
hipError_t launch_compute_energy(const DevFields &f, DevScalars *s, const KParams &P, const Scratch &sc,
                                 int64_t n_ub, hipStream_t st)
{
    int64_t nb = grid_for(n_ub);
    if (nb > sc.max_blocks) return hipErrorInvalidValue;
    kl<energy_kernel>(dim3((unsigned)nb), dim3(GT), 0, st, f, s, P, sc.partials);
    kl<energy_end_kernel>(dim3(1), dim3(GT), 0, st, s, (const double *)sc.partials, nb);
    return hipGetLastError();
}
