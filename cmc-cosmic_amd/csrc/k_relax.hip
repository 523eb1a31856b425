// a4 two-body relaxation (src/cmc/cmc_dynamics.c:17-542, relaxation branch
// :427-498), a10 relaxation time step (src/cmc/cmc_dynamics_helper.c:2526-2579),
// calc_sigma_r (:2821-2986), calc_timestep (src/cmc/cmc_utils.c:2132-2148) and the
// per-pair / per-star RNG substream bases.
//
// Relaxation is one thread per radially adjacent pair (2p+1, 2p+2): a pure
// streaming map reading r, m, phi, vr, vt of both stars plus r[k-20], r[k+20]
// (cache-served) and writing vr, vt, E, J, Y: 72 B/star of compulsory HBM traffic.
// The scattering angles' sin/cos are correctly rounded double-double evaluations
// (dd_trig.hpp), so the kicks do not depend on the device math library.
#include "kernels.hpp"
#include "dd_trig.hpp"

namespace {
__device__ __forceinline__ double sqr(double x) { return x * x; }
__device__ __forceinline__ double cub(double x) { return x * x * x; }

// cmc_dynamics_helper.c:768-786 calc_n_local
__device__ __forceinline__ double calc_n_local(const double *__restrict__ r, int64_t k, int64_t p, int64_t N_LIMIT)
{
    int64_t kmin = k - p, kmax = k + p;
    if (kmin < 0) { kmin = 0; kmax = 2 * p + 1; }
    else if (kmax > N_LIMIT) { kmax = N_LIMIT; kmin = N_LIMIT - 2 * p - 1; }
    // clusters smaller than 2p+2 stars would index outside [0, N+1] (undefined in
    // the reference as well); clamp to the array
    if (kmax > N_LIMIT + 1) kmax = N_LIMIT + 1;
    if (kmin < 0) kmin = 0;
    return (2.0 * ((double)p)) * 3.0 / (4.0 * CMCG_PI * (cub(r[kmax]) - cub(r[kmin])));
}

// The step's substream bases: base[0] = seeded state advanced by step * 2^72 (the
// step's relaxation pairs), base[1] = base[0] + 2^70 (its orbit slots).  A block of
// RB_T threads (rng_jump_block); the offset's bits 72..127 are the step's bits 0..55.
constexpr int RB_T = 128;
constexpr int DT_SLOTS = 64;     // the step's time-step minimum: slots of sigma_dt_kernel (<= RB_T)
__device__ __forceinline__ void rng_bases(const RngTable *__restrict__ T, DevScalars *s, const Rng *__restrict__ seeded,
                                          Rng *__restrict__ base, int64_t step_override)
{
    __shared__ uint32_t cols[56 * 128];
    __shared__ Rng res;
    const uint64_t step = step_override >= 0 ? (uint64_t)step_override : s->step;
    Rng b = rng_jump_block(T, *seeded, step, 72, 56, cols, &res);
    if (threadIdx.x == 0) base[0] = b;
    b = rng_jump_block(T, b, 1ull, 70, 1, cols, &res);
    if (threadIdx.x == 0) base[1] = b;
}
struct rng_bases_kernel {
    static constexpr int LB = RB_T, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const RngTable *__restrict__ T, DevScalars *s,
                                                         const Rng *__restrict__ seeded, Rng *__restrict__ base,
                                                         int64_t step_override)
{
    rng_bases(T, s, seeded, base, step_override);
}
};

// state at the first pair (phase 0: offset 3 * 64 w) or slot (phase 1: offset
// (1 + 64 w) 2^20) of wave w.  With w = 256 b + t (block b, thread t) the offset
// splits into a block part, jumped once per block from the global tables, and
// 3 t 2^6 (phase 0) or t 2^26 (phase 1), jumped through LDS nibble tables of
// WB_NB consecutive powers; jumps by powers of one matrix commute.
constexpr int WB_T = 256, WB_NB = RNG_WB_POWERS;
struct rng_wave_bases_kernel {
    static constexpr int LB = WB_T, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const RngTable *__restrict__ T,
                                                              const Rng *__restrict__ base, int phase,
                                                              int64_t nwaves, Rng *__restrict__ out, int jb1,
                                                              const NibbleTables<WB_NB> *__restrict__ nib)
{
    __shared__ NibbleTables<WB_NB> lt;
    __shared__ uint32_t cols[64 * 128];
    __shared__ Rng res;
    const int t = threadIdx.x;
    const uint64_t b = kg_.bid;
    copy_nibble_tables<WB_NB>(nib, lt);     // powers 2^6.. (phase 0) or 2^26.. (phase 1), built on the host
    const Rng bb = phase == 0 ? rng_jump_block(T, base[0], 49152ull * b, 0, jb1, cols, &res)
                              : rng_jump_block(T, base[1], 1ull + (b << 14), 20, jb1 - 20, cols, &res);
    const int64_t w = (int64_t)b * WB_T + t;
    if (w >= nwaves) return;
    out[w] = rng_jump_lds<WB_NB>(lt, bb, phase == 0 ? 3u * (uint32_t)t : (uint32_t)t);
}
};

struct rng_draws_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const RngTable *__restrict__ T, const Rng *__restrict__ base, int phase,
                                 int64_t first, int64_t count, int ndraw, double *__restrict__ out, int jb1)
{
    int64_t i = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int64_t idx = first + i;
    Rng st;
    if (phase == 0) st = rng_jump(T, base[0], 3ull * (uint64_t)idx, 0, jb1);
    else st = rng_jump(T, base[1], (uint64_t)idx << 20, 20, jb1);
    for (int d = 0; d < ndraw; d++) out[i * ndraw + d] = rng_dbl(st);
}
};

struct dtrel_init_kernel { static constexpr int LB = 1024, MB = 0; static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s) { s->dtrel = __longlong_as_double(0x7ff0000000000000ll); } };

// simul_relax_new: one thread per window si = 1+p, 1+3p, ... (si < N - p), the
// window sums taken in the reference's order, then an order-independent min: the
// bit patterns of non-negative doubles order like the values, so the wave takes the
// minimum pattern with shuffles and publishes it with one atomicMin.  A block (one
// wave) stages its windows' contiguous range in LDS with coalesced loads, as the
// per-element terms m*madhoc and m*madhoc*(vr^2 + vt^2) (the same roundings as the
// reference's loop); windows of more than SR_LDS elements read global memory.
// SR_T windows at most per block, one per lane of the first wave; all SR_B threads
// stage the block's elements (more loads in flight than one wave issues alone).
constexpr int SR_T = 64, SR_B = 256, SR_LDS = 2560;
struct simul_relax_kernel {
    static constexpr int LB = SR_B, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ r, const double *__restrict__ m,
                                                           const double *__restrict__ vr,
                                                           const double *__restrict__ vt, DevScalars *s, KParams P)
{
    __shared__ double stmp[SR_LDS], sq[SR_LDS];
    const unsigned long long kInf = 0x7ff0000000000000ull;
    unsigned long long mine = kInf;
    const int64_t p = P.avekernel, N = s->n_max, len = 2 * p;
    const bool staged = len <= SR_LDS;
    const int G = staged ? (SR_LDS / len < SR_T ? (int)(SR_LDS / len) : SR_T) : SR_T;   // windows per block
    const int64_t w0 = (int64_t)kg_.bid * G;
    const int64_t e0 = 1 + w0 * len;
    if (staged) {
        for (int64_t q = threadIdx.x; q < G * len; q += SR_B) {
            const int64_t k = e0 + q;
            if (k <= N) {
                const double tmp = m[k] * P.madhoc;
                stmp[q] = tmp;
                sq[q] = tmp * (sqr(vr[k]) + sqr(vt[k]));
            }
        }
        __syncthreads();
    }
    const int64_t w = w0 + threadIdx.x;
    const int64_t si = 1 + p + w * 2 * p;
    if (threadIdx.x < G && si < N - p) {
        const int64_t simin = si - p;
        double Mv2ave = 0.0, Mave = 0.0, M2ave = 0.0;
        if (staged) {
            const int64_t q0 = simin - e0;
            for (int64_t q = q0; q < q0 + len; q++) {
                const double tmp = stmp[q];
                Mv2ave += sq[q];
                Mave += tmp;
                M2ave += sqr(tmp);
            }
        } else {
            for (int64_t k = simin; k < simin + len; k++) {
                double tmp = m[k] * P.madhoc;
                Mv2ave += tmp * (sqr(vr[k]) + sqr(vt[k]));
                Mave += tmp;
                M2ave += sqr(tmp);
            }
        }
        Mv2ave /= (double)(2 * p);
        Mave /= (double)(2 * p);
        M2ave /= (double)(2 * p);
        double sigma = sqrt(Mv2ave / Mave);
        double W = 4.0 * sigma / sqrt(3.0 * CMCG_PI);
        double n_local = calc_n_local(r, si, p, N);
        double dt = sqr(2.0 * P.theta_se_max / CMCG_PI) * (CMCG_PI / 32.0) * cub(W) /
                    (((double)P.n_star) * n_local * (4.0 * M2ave));
        if (dt >= 0.0) mine = (unsigned long long)__double_as_longlong(dt);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        unsigned long long other = __shfl_xor(mine, o, 64);
        mine = other < mine ? other : mine;
    }
    if (threadIdx.x == 0 && mine != kInf) atomicMin((unsigned long long *)&s->dtrel, mine);   // windows: wave 0 only
}
};

// calc_timestep (cmc_utils.c:2132-2148) and GetTimeStep's fallback (cmc_evolution_thr.c:430-432)
__device__ __forceinline__ void timestep_finalize(DevScalars *s)
{
    double Dt = s->dtrel;
    if (Dt == __longlong_as_double(0x7ff0000000000000ll)) Dt = 0.001;
    if ((s->TidalMassLoss_old - s->OldTidalMassLoss) > 0.01) {
        Dt = s->Prev_Dt * 0.8;
    } else if (Dt > 1.1 * s->Prev_Dt && s->Prev_Dt > 0 && (s->TidalMassLoss - s->OldTidalMassLoss) > 0.02) {
        Dt = s->Prev_Dt * 1.1;
    }
    s->Dt = Dt;
    s->TotalTime += Dt;
}
struct timestep_finalize_kernel { static constexpr int LB = 1024, MB = 0; static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s) { timestep_finalize(s); } };
// the step's time step (thread 0; first the minimum over sigma_dt_kernel's slots, which
// are reset for the next step, when dt_part is given), then its RNG substream bases
struct timestep_rng_bases_kernel {
    static constexpr int LB = RB_T, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const RngTable *__restrict__ T, DevScalars *s,
                                                                  const Rng *__restrict__ seeded,
                                                                  Rng *__restrict__ base,
                                                                  unsigned long long *__restrict__ dt_part)
{
    __shared__ unsigned long long wmin[RB_T / 64];
    if (dt_part) {
        unsigned long long mine = 0x7ff0000000000000ull;
        if (threadIdx.x < DT_SLOTS) {
            const unsigned long long c = ~dt_part[threadIdx.x * 16];
            dt_part[threadIdx.x * 16] = 0ull;
            mine = c < mine ? c : mine;
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const unsigned long long other = __shfl_xor(mine, o, 64);
            mine = other < mine ? other : mine;
        }
        if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = mine;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int q = 1; q < RB_T / 64; q++) mine = wmin[q] < mine ? wmin[q] : mine;
            const unsigned long long cur = (unsigned long long)__double_as_longlong(s->dtrel);
            s->dtrel = __longlong_as_double((long long)(mine < cur ? mine : cur));
        }
    }
    if (threadIdx.x == 0) timestep_finalize(s);
    rng_bases(T, s, seeded, base, -1);
}
};

// calc_sigma_r at one rank: sigma over the clamped window [si-p, si+p-1]; the
// reference slides running sums, here each window is summed directly.
// The same window sums with the block's range [first - 2p, last + 2p] staged in
// LDS (every window, including the shifted ones at both ends, lies inside it), in
// the same order and expression as sigma_kernel.
constexpr int SIG_T = 512, SIG_MAXP = 64, SIG_LDS = SIG_T + 4 * SIG_MAXP + 1;   // 12 KB

struct sigma_lds_kernel {
    static constexpr int LB = SIG_T, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ m, const double *__restrict__ vr,
                                                          const double *__restrict__ vt, double *__restrict__ sigma,
                                                          DevScalars *s, KParams P)
{
    __shared__ double a[SIG_LDS], b[SIG_LDS];
    const int64_t N = s->n_max, p = P.avekernel;
    const int64_t first = 1 + (int64_t)kg_.bid * SIG_T;
    if (first > N) return;
    const int64_t last = first + SIG_T - 1 < N ? first + SIG_T - 1 : N;
    const int64_t lo = first - 2 * p > 1 ? first - 2 * p : 1;
    const int64_t hi = last + 2 * p < N ? last + 2 * p : N;
    for (int64_t k = lo + threadIdx.x; k <= hi; k += SIG_T) {
        double mk = m[k] * P.madhoc;
        a[k - lo] = mk;
        b[k - lo] = mk * (sqr(vr[k]) + sqr(vt[k]));
    }
    __syncthreads();
    const int64_t si = first + threadIdx.x;
    if (si > N) return;
    int64_t simin = si - p, simax = simin + (2 * p - 1);
    if (simin < 1) { simin = 1; simax = simin + (2 * p - 1); }
    else if (simax > N) { simax = N; simin = simax - (2 * p - 1); }
    if (simin < 1) simin = 1;
    if (simax > N) simax = N;
    double Mv2ave = 0.0, Mave = 0.0;
    for (int64_t k = simin; k <= simax; k++) {
        Mv2ave += b[k - lo];
        Mave += a[k - lo];
    }
    sigma[si] = sqrt(Mv2ave / Mave);
}
};

struct sigma_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ m, const double *__restrict__ vr,
                             const double *__restrict__ vt, double *__restrict__ sigma, DevScalars *s, KParams P)
{
    int64_t si = 1 + (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    int64_t N = s->n_max, p = P.avekernel;
    if (si > N) return;
    int64_t simin = si - p, simax = simin + (2 * p - 1);
    if (simin < 1) { simin = 1; simax = simin + (2 * p - 1); }
    else if (simax > N) { simax = N; simin = simax - (2 * p - 1); }
    if (simin < 1) simin = 1;      // N < 2p: keep the window inside the array
    if (simax > N) simax = N;
    double Mv2ave = 0.0, Mave = 0.0;
    for (int64_t k = simin; k <= simax; k++) {
        double mk = m[k] * P.madhoc;
        Mv2ave += mk * (sqr(vr[k]) + sqr(vt[k]));
        Mave += mk;
    }
    sigma[si] = sqrt(Mv2ave / Mave);
}
};

// calc_sigma_r and simul_relax_new in one pass (the step's form): each block stages
// m*madhoc and m*madhoc*(vr^2 + vt^2) for its SIG_T stars and 2p on either side (the
// roundings of both references' loops, as sigma_lds_kernel and simul_relax_kernel
// take them), writes sigma of its stars, evaluates the relaxation-time windows that
// start among its stars (window w = [1 + 2pw, 1 + 2p(w+1)), the sums in the
// reference's order) and folds the block's smallest dt into one of DT_SLOTS slots
// (128 bytes apart) as an atomicMax of the complemented bit pattern (non-negative
// doubles order like their patterns; 0 is the identity, so the slots start and are
// left at 0); the one-block time-step kernel takes the minimum over the slots.  Needs
// p <= SIG_MAXP.
struct sigma_dt_kernel {
    static constexpr int LB = SIG_T, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ r, const double *__restrict__ m,
                                                         const double *__restrict__ vr, const double *__restrict__ vt,
                                                         double *__restrict__ sigma, DevScalars *s, KParams P,
                                                         unsigned long long *__restrict__ dt_part)
{
    __shared__ double a[SIG_LDS], b[SIG_LDS];
    __shared__ unsigned long long wmin[SIG_T / 64];
    const unsigned long long kInf = 0x7ff0000000000000ull;
    const int64_t N = s->n_max, p = P.avekernel, len = 2 * p;
    const int64_t first = 1 + (int64_t)kg_.bid * SIG_T;
    const int64_t last = first + SIG_T - 1 < N ? first + SIG_T - 1 : N;
    const int64_t lo = first - 2 * p > 1 ? first - 2 * p : 1;
    const int64_t hi = last + 2 * p < N ? last + 2 * p : N;
    if (first <= N)
        for (int64_t k = lo + threadIdx.x; k <= hi; k += SIG_T) {
            const double mk = m[k] * P.madhoc;
            a[k - lo] = mk;
            b[k - lo] = mk * (sqr(vr[k]) + sqr(vt[k]));
        }
    __syncthreads();
    unsigned long long mine = kInf;
    if (first <= N) {
        const int64_t si = first + threadIdx.x;
        if (si <= N) {
            int64_t simin = si - p, simax = simin + (2 * p - 1);
            if (simin < 1) { simin = 1; simax = simin + (2 * p - 1); }
            else if (simax > N) { simax = N; simin = simax - (2 * p - 1); }
            if (simin < 1) simin = 1;
            if (simax > N) simax = N;
            double Mv2ave = 0.0, Mave = 0.0;
            for (int64_t k = simin; k <= simax; k++) {
                Mv2ave += b[k - lo];
                Mave += a[k - lo];
            }
            sigma[si] = sqrt(Mv2ave / Mave);
        }
        // the windows starting at 1 + 2pw in [first, last]; each lies inside [lo, hi]
        const int64_t w = (first - 1 + len - 1) / len + threadIdx.x;
        const int64_t wstart = 1 + w * len, wsi = wstart + p;
        if (wstart <= last && wsi < N - p) {
            double Mv2ave = 0.0, Mave = 0.0, M2ave = 0.0;
            for (int64_t k = wstart; k < wstart + len; k++) {
                const double tmp = a[k - lo];
                Mv2ave += b[k - lo];
                Mave += tmp;
                M2ave += sqr(tmp);
            }
            Mv2ave /= (double)(2 * p);
            Mave /= (double)(2 * p);
            M2ave /= (double)(2 * p);
            const double sg = sqrt(Mv2ave / Mave);
            const double W = 4.0 * sg / sqrt(3.0 * CMCG_PI);
            const double n_local = calc_n_local(r, wsi, p, N);
            const double dt = sqr(2.0 * P.theta_se_max / CMCG_PI) * (CMCG_PI / 32.0) * cub(W) /
                              (((double)P.n_star) * n_local * (4.0 * M2ave));
            if (dt >= 0.0) mine = (unsigned long long)__double_as_longlong(dt);
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long other = __shfl_xor(mine, o, 64);
        mine = other < mine ? other : mine;
    }
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < SIG_T / 64; q++) mine = wmin[q] < mine ? wmin[q] : mine;
        if (mine != kInf) atomicMax(&dt_part[(kg_.bid % DT_SLOTS) * 16], ~mine);
    }
}
};

// Pair pr draws from offset 3 pr; a wave's first pair starts at the per-wave state
// and lane l jumps 3 l < 2^8 steps through the LDS tables.
// energy_conservation1 (cmc_utils.c:227-236) for star i from its values after the
// relaxation: vtold, vrold and Uoldrold = phi + MPI_PHI_S(r, i)
__device__ __forceinline__ void ec1_star(DevFields &f, const KParams &P, int64_t i, double r, double m, double phi,
                                         double vr, double vt)
{
    f.vtold[i] = vt;
    f.vrold[i] = vr;
    f.Uoldrold[i] = phi + phi_s_own(r, m, P.inv_nstar, r);
}

// EC1: energy_conservation1 fused into the pass (the step's form; the entry point of
// that name keeps its own kernel).  T: threads per block (an autotuned launch shape;
// pair p is thread p of the grid whatever T is, so results do not depend on it).
template <bool EC1, int T>
struct relax_kernel {
    static constexpr int LB = T, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P,
                                                    const Rng *__restrict__ wave_base,
                                                    const NibbleTables<8> *__restrict__ nib)
{
    __shared__ NibbleTables<8> lt;
    copy_nibble_tables<8>(nib, lt);
    int64_t pr = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    int64_t N = s->n_max;
    if (EC1 && pr == N / 2 && (N & 1)) {    // the unpaired last star keeps its velocities
        ec1_star(f, P, N, f.r[N], f.m[N], f.phi[N], f.vr[N], f.vt[N]);
        return;
    }
    if (pr >= N / 2) return;            // pairs while si <= N - N%2 - 1 (cmc_dynamics.c:203)
    const int64_t k = 2 * pr + 1, kp = k + 1;
    const double dt = s->Dt;
    Rng rng = rng_jump_lds<8>(lt, wave_base[pr >> 6], 3u * (uint32_t)(pr & 63));

    // calc_encounter_dyns (cmc_dynamics_helper.c:831-871)
    double Y = rng_dbl(rng);
    f.Y[k] = Y;
    f.Y[kp] = Y;
    double phi = Y * 2.0 * CMCG_PI;
    double vrk = f.vr[k], vtk = f.vt[k], vrkp = f.vr[kp], vtkp = f.vt[kp];
    double v[4], vp[4], w[4], w1[4], w2[4];
    v[1] = vtk; v[2] = 0.0; v[3] = vrk;
    double sphi, cphi;
    ddtrig::sincos(phi, &sphi, &cphi);
    vp[1] = vtkp * cphi;
    vp[2] = vtkp * sphi;
    vp[3] = vrkp;
    for (int j = 1; j <= 3; j++) w[j] = vp[j] - v[j];
    double W = sqrt(sqr(w[1]) + sqr(w[2]) + sqr(w[3]));
    if (W == 0.0) { dev_set_error(s, CMCG_E_WZERO, k); return; }

    double n_local = calc_n_local(f.r, k, P.avekernel, N);
    double mass_k = f.m[k], mass_kp = f.m[kp];
    (void)rng_dbl(rng);                 // P_enc test (cmc_dynamics.c:406); P_enc = 0 here

    double Trel12 = (CMCG_PI / 32.0) * cub(W) / (((double)P.n_star) * n_local * sqr((mass_k + mass_kp) * P.madhoc));
    double beta = (CMCG_PI / 2.0) * sqrt(dt / Trel12);
    if (beta > CMCG_PI / 2.0) beta = CMCG_PI / 2.0;
    double wp = sqrt(sqr(w[1]) + sqr(w[2]));
    if (wp == 0.0) { dev_set_error(s, CMCG_E_WPZERO, k); return; }
    w1[1] = -w[2] * W / wp;
    w1[2] = w[1] * W / wp;
    w1[3] = 0.0;
    w2[1] = -w[1] * w[3] / wp;
    w2[2] = -w[2] * w[3] / wp;
    w2[3] = wp;
    double psi = rng_dbl(rng) * 2 * CMCG_PI;
    double cb, sb, cp, sp;
    ddtrig::sincos(beta, &sb, &cb);
    ddtrig::sincos(psi, &sp, &cp);
    double v_new[4], vp_new[4];
    for (int j = 1; j <= 3; j++) {
        double w_new = w[j] * cb + w1[j] * sb * cp + w2[j] * sb * sp;
        v_new[j] = v[j] - mass_kp / (mass_k + mass_kp) * (w_new - w[j]);
        vp_new[j] = vp[j] + mass_k / (mass_k + mass_kp) * (w_new - w[j]);
    }
    double nvr_k = v_new[3], nvt_k = sqrt(sqr(v_new[1]) + sqr(v_new[2]));
    double nvr_kp = vp_new[3], nvt_kp = sqrt(sqr(vp_new[1]) + sqr(vp_new[2]));
    f.vr[k] = nvr_k; f.vt[k] = nvt_k;
    f.vr[kp] = nvr_kp; f.vt[kp] = nvt_kp;
    // set_star_EJ (cmc_dynamics_helper.c:878-882), old phi and r
    const double phik = f.phi[k], phikp = f.phi[kp], rk = f.r[k], rkp = f.r[kp];
    f.E[k] = phik + 0.5 * (sqr(nvr_k) + sqr(nvt_k));
    f.J[k] = rk * nvt_k;
    f.E[kp] = phikp + 0.5 * (sqr(nvr_kp) + sqr(nvt_kp));
    f.J[kp] = rkp * nvt_kp;
    if (EC1) {
        ec1_star(f, P, k, rk, mass_k, phik, nvr_k, nvt_k);
        ec1_star(f, P, kp, rkp, mass_kp, phikp, nvr_kp, nvt_kp);
    }
}
};

int bits_needed(uint64_t x)
{
    int b = 0;
    while (b < 64 && (x >> b)) b++;
    return b;
}
} // namespace

hipError_t launch_rng_bases(const RngTable *T, DevScalars *s, const Scratch &sc, hipStream_t st)
{
    kl<rng_bases_kernel>(dim3(1), dim3(RB_T), 0, st, T, s, sc.rng_seeded, sc.rng_base, (int64_t)-1);
    return hipGetLastError();
}

hipError_t launch_rng_bases_step(const RngTable *T, DevScalars *s, const Scratch &sc, int64_t step, hipStream_t st)
{
    kl<rng_bases_kernel>(dim3(1), dim3(RB_T), 0, st, T, s, sc.rng_seeded, sc.rng_base, step);
    return hipGetLastError();
}

hipError_t launch_rng_wave_bases(const RngTable *T, const Scratch &sc, int phase, int64_t count, hipStream_t st)
{
    int64_t nwaves = (count + 63) / 64;
    if (nwaves <= 0) return hipSuccess;
    if (nwaves > sc.max_waves) return hipErrorInvalidValue;
    int jb1 = phase == 0 ? bits_needed(192ull * (uint64_t)nwaves) : 20 + bits_needed((uint64_t)(64 * nwaves + 1));
    const NibbleTables<WB_NB> *nib = (const NibbleTables<WB_NB> *)sc.wb_nib + (phase == 0 ? 0 : 1);
    kl<rng_wave_bases_kernel>(dim3((unsigned)((nwaves + WB_T - 1) / WB_T)), dim3(WB_T), 0, st, T,
                       (const Rng *)sc.rng_base, phase, nwaves, sc.rng_wave, jb1, nib);
    return hipGetLastError();
}

// Packed (r, phi) nodes for the searches and the middle table (every MID_STEP-th node).
struct mid_build_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ r, const double *__restrict__ phi, DevScalars *s,
                                 double2 *__restrict__ nodes, double2 *__restrict__ mid)
{
    int64_t k = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (k > s->n_max + 1) return;
    double2 v = make_double2(r[k], phi[k]);
    nodes[k] = v;
    if (k % MID_STEP == 0) mid[k / MID_STEP] = v;
}
};

hipError_t launch_mid_build(const DevFields &f, DevScalars *s, const Scratch &sc, int64_t n_ub, hipStream_t st)
{
    int64_t nn = n_ub + 2;
    kl<mid_build_kernel>(dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, f.r, f.phi, s,
                       sc.nodes, sc.mid);
    return hipGetLastError();
}

hipError_t launch_rng_draws(const RngTable *T, const Rng *base, int phase, int64_t first, int64_t count,
                            int ndraw, double *out, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    uint64_t last = (uint64_t)(first + count);
    int jb1 = phase == 0 ? bits_needed(3ull * last) : 20 + bits_needed(last);
    kl<rng_draws_kernel>(dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, T, base, phase,
                       first, count, ndraw, out, jb1);
    return hipGetLastError();
}

hipError_t launch_sigma(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub, hipStream_t st)
{
    if (n_ub <= 0) return hipSuccess;
    if (P.avekernel <= SIG_MAXP)
        kl<sigma_lds_kernel>(dim3((unsigned)((n_ub + SIG_T - 1) / SIG_T)), dim3(SIG_T), 0, st, f.m,
                           f.vr, f.vt, f.sigma, s, P);
    else
        kl<sigma_kernel>(dim3((unsigned)((n_ub + 255) / 256)), dim3(256), 0, st, f.m, f.vr, f.vt,
                           f.sigma, s, P);
    return hipGetLastError();
}

hipError_t launch_simul_relax(const DevFields &f, DevScalars *s, const KParams &P, int64_t n_ub, hipStream_t st,
                              bool init_min)
{
    if (init_min) kl<dtrel_init_kernel>(dim3(1), dim3(1), 0, st, s);
    const int64_t nw = n_ub / (2 * P.avekernel) + 1;
    const int64_t len = 2 * P.avekernel;
    const int64_t G = len <= SR_LDS ? (SR_LDS / len < SR_T ? SR_LDS / len : SR_T) : SR_T;   // as in the kernel
    kl<simul_relax_kernel>(dim3((unsigned)((nw + G - 1) / G)), dim3(SR_B), 0, st, f.r, f.m, f.vr,
                       f.vt, s, P);
    return hipGetLastError();
}

hipError_t launch_timestep_rng_bases(const RngTable *T, DevScalars *s, const Scratch &sc, hipStream_t st)
{
    kl<timestep_rng_bases_kernel>(dim3(1), dim3(RB_T), 0, st, T, s, sc.rng_seeded, sc.rng_base,
                       (unsigned long long *)nullptr);
    return hipGetLastError();
}

// The step's time-step phase: calc_sigma_r and simul_relax_new in one pass (or, for a
// window longer than the pass stages, their own kernels), then Dt and the RNG bases.
hipError_t launch_timestep_sigma(const DevFields &f, DevScalars *s, const KParams &P, const RngTable *T,
                                 const Scratch &sc, int64_t n_ub, hipStream_t st)
{
    if (P.avekernel > SIG_MAXP || n_ub <= 0) {
        if (hipError_t e = launch_simul_relax(f, s, P, n_ub, st, false)) return e;
        if (hipError_t e = launch_sigma(f, s, P, n_ub, st)) return e;
        return launch_timestep_rng_bases(T, s, sc, st);
    }
    const int64_t nb = (n_ub + SIG_T - 1) / SIG_T;
    if (nb > sc.max_blocks) return hipErrorInvalidValue;
    kl<sigma_dt_kernel>(dim3((unsigned)nb), dim3(SIG_T), 0, st, f.r, f.m, f.vr, f.vt, f.sigma, s, P,
                       sc.dt_part);
    kl<timestep_rng_bases_kernel>(dim3(1), dim3(RB_T), 0, st, T, s, sc.rng_seeded, sc.rng_base,
                       sc.dt_part);
    return hipGetLastError();
}
hipError_t launch_timestep_finalize(DevScalars *s, hipStream_t st)
{
    kl<timestep_finalize_kernel>(dim3(1), dim3(1), 0, st, s);
    return hipGetLastError();
}

template <bool EC1, int TH>
void relax_launch(int64_t nth, hipStream_t st, const DevFields &f, DevScalars *s, const KParams &P, const Scratch &sc)
{
    kl<relax_kernel<EC1, TH>>(dim3((unsigned)((nth + TH - 1) / TH)), dim3(TH), 0, st, f, s, P,
                              (const Rng *)sc.rng_wave, (const NibbleTables<8> *)sc.relax_nib);
}

template <bool EC1>
hipError_t relax_launch_shape(int threads, int64_t nth, hipStream_t st, const DevFields &f, DevScalars *s,
                              const KParams &P, const Scratch &sc)
{
    switch (threads) {
    case 128: relax_launch<EC1, 128>(nth, st, f, s, P, sc); break;
    case 256: relax_launch<EC1, 256>(nth, st, f, s, P, sc); break;
    case 512: relax_launch<EC1, 512>(nth, st, f, s, P, sc); break;
    case 1024: relax_launch<EC1, 1024>(nth, st, f, s, P, sc); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_relax(const DevFields &f, DevScalars *s, const KParams &P, const RngTable *T,
                        const Scratch &sc, int64_t n_ub, hipStream_t st, bool with_ec1, int threads)
{
    int64_t npairs = n_ub / 2;
    if (with_ec1) {
        // one thread more for an unpaired last star
        const int64_t nth = (n_ub + 1) / 2;
        if (nth <= 0) return hipSuccess;
        if (npairs > 0)
            if (hipError_t e = launch_rng_wave_bases(T, sc, 0, npairs, st)) return e;
        return relax_launch_shape<true>(threads, nth, st, f, s, P, sc);
    }
    if (npairs <= 0) return hipSuccess;
    if (hipError_t e = launch_rng_wave_bases(T, sc, 0, npairs, st)) return e;
    return relax_launch_shape<false>(threads, npairs, st, f, s, P, sc);
}
