// a5-a8: get_positions_loop (src/cmc/cmc_evolution_thr.c:784-937) with
// calc_orbit_new (src/cmc/cmc_orbit.c:413-540), find_zero_Q (src/cmc/cmc_nr.c:316-383),
// find_root_vr (:500-603) and potential(r) (src/cmc/cmc_utils.c:112-165).
//
// One thread per star.  The rejection loop's potential(r) bisection is narrowed to
// [kmin, kmax+1] (the bracket the apsides already bound; the result index is the
// same as the reference's full-range FindZero_r, see potential_at).  Each star draws
// from its own taus113 substream (offset step*2^72 + 2^70 + slot*2^20).
//
// get_positive_Q_index (cmc_orbit.c:30-59) scans linearly from the centre when Q at
// the star's own index is not positive (v_r rounded to ~0).  That O(N) walk by one
// thread would stall the whole phase, so such stars are deferred: a grid-wide
// search finds the same first positive index, and a tail kernel finishes them.
//
// The phase is bound by dependent gathers into the sorted (r, phi) table (16 B per
// probe, ~(2 + tries) * log2 N probes per star), which at N = 4e6 is 64 MB and lives
// in L2/MALL; the compulsory HBM traffic is 96 B/star.
#include "kernels.hpp"

namespace {
constexpr int ORB_THREADS = 256;
constexpr int ORB_COARSE = 1024;     // LDS samples of the (r, phi) table per block

// Per-block partials of the orbit phase (fixed-order wave-shuffle tree, then the
// four waves in order): [0] max_r (max), [1] Eescaped, [2] Jescaped,
// [3] TidalMassLoss, [4] Etidal, [5..10] star counts by outcome (circular,
// removed E >= 0, removed r_apo >= Rtidal, destroyed, Brent, linear scan),
// [11] tries, [12] max tries (max), [13] learned brackets used, [14] learned
// brackets rejected.
constexpr int NRED = 15, RED_STRIDE = 16;
__device__ __forceinline__ bool red_is_max(int i) { return i == 0 || i == 12; }

template <int THREADS>
__device__ void block_reduce_store(double (&v)[NRED], double *__restrict__ out)
{
    __shared__ double wsum[THREADS / 64][NRED];
    for (int off = 32; off > 0; off >>= 1)
        for (int i = 0; i < NRED; i++) {
            double o = __shfl_xor(v[i], off);
            v[i] = red_is_max(i) ? (v[i] >= o ? v[i] : o) : v[i] + o;
        }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int i = 0; i < NRED; i++) wsum[w][i] = v[i];
    __syncthreads();
    if (threadIdx.x < NRED) {
        const int i = threadIdx.x;
        double a = wsum[0][i];
        for (int k = 1; k < THREADS / 64; k++) a = red_is_max(i) ? (a >= wsum[k][i] ? a : wsum[k][i]) : a + wsum[k][i];
        out[i] = a;
    }
}

// outcome counters of one star into the partials
__device__ __forceinline__ void count_star(double (&red)[NRED], int flags, int ntry, int hits, int misses,
                                           bool count_linscan)
{
    red[5] += (flags & CMCG_SF_CIRCULAR) ? 1.0 : 0.0;
    red[6] += (flags & CMCG_SF_REMOVED_E) ? 1.0 : 0.0;
    red[7] += (flags & CMCG_SF_REMOVED_RT) ? 1.0 : 0.0;
    red[8] += (flags & CMCG_SF_DESTROYED) ? 1.0 : 0.0;
    red[9] += (flags & CMCG_SF_BRENT) ? 1.0 : 0.0;
    red[10] += (count_linscan && (flags & CMCG_SF_LINSCAN)) ? 1.0 : 0.0;
    red[11] += (double)ntry;
    red[12] = red[12] >= (double)ntry ? red[12] : (double)ntry;
    red[13] += (double)hits;
    red[14] += (double)misses;
}

// remove_star (cmc_evolution_thr.c:655-716) contributions: Eescaped, Jescaped,
// TidalMassLoss, Etidal
__device__ __forceinline__ void removed(double (&red)[NRED], double Ej, double J, double mj, const KParams &P)
{
    red[1] = Ej * mj / (double)P.n_star;
    red[2] = J * mj / (double)P.n_star;
    red[3] = mj / (double)P.n_star;
    red[4] = Ej * mj / (double)P.n_star;
}

// The part of get_positions_loop after the positive-Q index is known
// (cmc_evolution_thr.c:832-934).
template <class N>
__device__ void sample_star(const N &nd, const DevFields &f, DevScalars *s, const KParams &P, Rng rng,
                            int64_t n, int64_t j, double rj, double mj, double E,
                            double Ej, double J, int64_t ktemp, int &flags, int &ntry, int &hits, int &misses,
                            double (&red)[NRED])
{
    const double *__restrict__ r = f.r;
    const double *__restrict__ phi = f.phi;
    OrbitRS o = calc_orbit_from_n(nd, r, phi, n, j, rj, mj, P, E, J, ktemp);
    if (o.err) dev_set_error(s, o.err, j);
    hits = o.hint_hits;
    misses = o.hint_misses;
    if (o.brent) flags |= CMCG_SF_BRENT;
    if (o.circular) {
        f.rnew[j] = rj; f.vrnew[j] = f.vr[j]; f.vtnew[j] = f.vt[j];
        flags |= CMCG_SF_CIRCULAR;
        return;
    }
    if (o.ra >= s->Rtidal) {
        f.r_apo[j] = o.ra; f.r_peri[j] = o.rp;
        f.rnew[j] = CMCG_SF_INFINITY; f.vrnew[j] = 0.0; f.vtnew[j] = 0.0;
        removed(red, Ej, J, mj, P);
        flags |= CMCG_SF_REMOVED_RT;
        return;
    }
    const double rmin = o.rp, rmax = o.ra;
    double g1 = sqrt(3.0 * (rmax - rmin) / o.dQdrp);
    double g2 = sqrt(-3.0 * (rmax - rmin) / o.dQdra);
    double F = 1.2 * (g1 >= g2 ? g1 : g2);
    // narrowed bracket for potential(r): r in [rmin, rmax] within [r_kmin, r_kmax+1]
    int64_t lo0 = o.kmin < 1 ? 1 : o.kmin;
    int64_t hi0 = o.kmax + 1 > n + 1 ? n + 1 : o.kmax + 1;
    const double r_lo0 = nd.R(lo0), r_hi0 = nd.R(hi0);
    double X = 0.0, rr = 0.0, vr = 0.0, potr = 0.0;
    int k;
    bool bad = false;
    for (k = 1; k <= CMCG_N_TRY; k++) {
        X = rng_dbl(rng);
        double s0 = 2.0 * X - 1.0;
        double g0 = F * rng_dbl(rng);
        rr = 0.5 * (rmin + rmax) + 0.25 * (rmax - rmin) * (3.0 * s0 - s0 * s0 * s0);
        int64_t lo = lo0, hi = hi0;
        if (!(r_lo0 < rr)) lo = 1;
        if (!(rr <= r_hi0)) hi = n + 1;
        potr = potential_at_n(nd, n, rr, lo, hi, &bad);
        double pot = potr + phi_s_own(rj, mj, P.inv_nstar, rr);
        double drds = 0.25 * (rmax - rmin) * (3.0 - 3.0 * s0 * s0);
        double Q = 2.0 * E - 2.0 * pot - J * J / rr / rr;
        if (Q >= 0.0) {
            vr = sqrt(Q);
        } else {
            if (isnan(Q)) { dev_set_error(s, CMCG_E_NAN, j); break; }
            vr = 0;
        }
        if (g0 < 1.0 / vr * drds) break;
    }
    if (bad) dev_set_error(s, CMCG_E_BRACKET, j);
    if (k == CMCG_N_TRY + 1) dev_set_error(s, CMCG_E_NTRY, j);
    ntry = k;
    f.X[j] = X;
    f.r_peri[j] = rmin;
    f.r_apo[j] = rmax;
    if (rng_dbl(rng) < 0.5) vr = -vr;
    f.rnew[j] = rr;
    f.vrnew[j] = vr;
    f.vtnew[j] = J / rr;
    f.pot_rnew[j] = potr;
    flags |= 256;      // sampled: pot_rnew holds potential(rnew)
    red[0] = rr;
}

// ------------------------------------------------------------------ wave-parallel tries
// The rejection loop of get_positions_loop (cmc_evolution_thr.c:858-906) takes a
// geometric number of tries (mean ~2), so a wave that ran one star per lane would
// idle most lanes while its slowest star (~8 tries) finishes.  Instead the wave
// evaluates tries in rounds: the stars still trying share the 64 lanes, each lane
// taking the next untried try number of one star, stepped to that try's draws in the
// star's own substream (try t uses draws 2t-2 and 2t-1, the sign draw follows the
// accepted try).  A star's result is its lowest-numbered terminating try, exactly as
// the sequential loop, so outputs do not depend on the grouping.

// what a star's tries need (set by prepare_star)
struct TryCtx {
    double rmin, rmax, F, E, J, rj, mj, r_lo0, r_hi0;
    int64_t lo0, hi0;
};

// result of a star's tries
struct TryOut {
    double X, rr, vr, potr;
    int k;             // tries taken; CMCG_N_TRY + 1 when none was accepted
    bool bad, nan;
};

// q-th (from 0) set bit of m
__device__ __forceinline__ int nth_set_bit(uint64_t m, int q)
{
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const uint64_t low = m & ((1ull << w) - 1ull);
        const int c = __popcll(low);
        if (q >= c) { q -= c; m >>= w; pos += w; }
        else m = low;
    }
    return pos;
}

// The part of get_positions_loop between the positive-Q index and the rejection loop
// (cmc_evolution_thr.c:832-857): apsides, removal at r_apo >= Rtidal, circular
// orbits, and the loop's constants.  Returns true if the star samples a new radius.
template <class N>
__device__ __forceinline__ bool prepare_star(const N &nd, const DevFields &f, DevScalars *s, const KParams &P,
                                             const Coarse &C, int64_t n, int64_t j, double rj, double mj,
                                             double E, double J, int64_t ktemp, int &flags, int &hits, int &misses,
                                             TryCtx &tc)
{
    OrbitRS o = calc_orbit_core<N, true>(nd, f.r, f.phi, n, j, rj, mj, P, E, J, ktemp, C);
    if (o.err) dev_set_error(s, o.err, j);
    hits = o.hint_hits;
    misses = o.hint_misses;
    if (o.brent) flags |= CMCG_SF_BRENT;
    if (o.circular) {
        f.rnew[j] = rj; f.vrnew[j] = f.vr[j]; f.vtnew[j] = f.vt[j];
        flags |= CMCG_SF_CIRCULAR;
        return false;
    }
    if (o.ra >= s->Rtidal) {
        f.r_apo[j] = o.ra; f.r_peri[j] = o.rp;
        f.rnew[j] = CMCG_SF_INFINITY; f.vrnew[j] = 0.0; f.vtnew[j] = 0.0;
        flags |= CMCG_SF_REMOVED_RT;      // its removal sums are added by the caller
        return false;
    }
    tc.rmin = o.rp;
    tc.rmax = o.ra;
    const double g1 = sqrt(3.0 * (tc.rmax - tc.rmin) / o.dQdrp);
    const double g2 = sqrt(-3.0 * (tc.rmax - tc.rmin) / o.dQdra);
    tc.F = 1.2 * (g1 >= g2 ? g1 : g2);
    tc.E = E; tc.J = J; tc.rj = rj; tc.mj = mj;
    tc.lo0 = o.kmin < 1 ? 1 : o.kmin;
    tc.hi0 = o.kmax + 1 > n + 1 ? n + 1 : o.kmax + 1;
    tc.r_lo0 = nd.R(tc.lo0);
    tc.r_hi0 = nd.R(tc.hi0);
    return true;
}

// One try (cmc_evolution_thr.c:862-886) from the state at its first draw; on return
// st is the state after the try's two draws.
template <class N>
__device__ __forceinline__ bool one_try(const N &nd, const Coarse &C, const KParams &P, int64_t n, const TryCtx &tc,
                                        Rng &st, double &X, double &rr, double &vr, double &potr, bool &bad,
                                        bool &nan)
{
    X = rng_dbl(st);
    const double s0 = 2.0 * X - 1.0;
    const double g0 = tc.F * rng_dbl(st);
    rr = 0.5 * (tc.rmin + tc.rmax) + 0.25 * (tc.rmax - tc.rmin) * (3.0 * s0 - s0 * s0 * s0);
    int64_t lo = tc.lo0, hi = tc.hi0;
    if (!(tc.r_lo0 < rr)) lo = 1;
    if (!(rr <= tc.r_hi0)) hi = n + 1;
    potr = C.ri.t ? potential_at_ridx_n(nd, n, rr, lo, hi, C.ri, &bad)
                  : potential_at_coarse_n(nd, n, rr, lo, hi, C, &bad);
    const double pot = potr + phi_s_own(tc.rj, tc.mj, P.inv_nstar, rr);
    const double drds = 0.25 * (tc.rmax - tc.rmin) * (3.0 - 3.0 * s0 * s0);
    const double Q = 2.0 * tc.E - 2.0 * pot - tc.J * tc.J / rr / rr;
    if (Q >= 0.0) {
        vr = sqrt(Q);
    } else {
        if (isnan(Q)) { nan = true; vr = 0.0; return true; }
        vr = 0;
    }
    return g0 < 1.0 / vr * drds;
}

// All lanes of the wave call this together (convergent).  need: this lane's star
// samples; wctx: the wave's 64 try constants in LDS (slot = lane, written before a
// block barrier); rng: this lane's substream state before its star's first try.
template <int CAP, class N>
__device__ void wave_tries(const N &nd, const Coarse &C, const KParams &P, int64_t n, bool need,
                           const TryCtx *__restrict__ wctx, Rng rng, TryOut &out)
{
    const int lane = threadIdx.x & 63;
    int next = 1;                         // owner: first try not yet evaluated
    bool bad_acc = false;
    out.k = 0; out.bad = false; out.nan = false;
    out.X = out.rr = out.vr = out.potr = 0.0;
    uint64_t pending = __ballot(need);
    while (pending) {
        const int R = __popcll(pending);
        int c = 64 / R, rem = 64 - c * R;
        if (c >= CAP) { c = CAP; rem = 0; }              // at most CAP lanes per star: fewer wasted tries
        const bool idle = lane >= c * R + rem;
        int q, o, cs;
        if (lane < (c + 1) * rem) { q = lane / (c + 1); o = lane % (c + 1); cs = c + 1; }
        else { const int l2 = lane - (c + 1) * rem; q = rem + l2 / c; o = l2 % c; cs = c; }
        const int sl = idle ? nth_set_bit(pending, 0) : nth_set_bit(pending, q);   // owner lane of the star helped
        const TryCtx t = wctx[sl];
        Rng st;
        st.z0 = __shfl(rng.z0, sl); st.z1 = __shfl(rng.z1, sl); st.z2 = __shfl(rng.z2, sl); st.z3 = __shfl(rng.z3, sl);
        const int k = __shfl(next, sl) + o;
        const bool eval = !idle && k <= CMCG_N_TRY;
        if (!idle)
            for (int d = 0; d < 2 * o; d++) (void)rng_int(st);
        double X = 0.0, rr = 0.0, vr = 0.0, potr = 0.0;
        bool bad = false, nan = false, term = false;
        if (eval) {
            term = one_try(nd, C, P, n, t, st, X, rr, vr, potr, bad, nan);
            Rng sg = st;
            if (term && !nan && rng_dbl(sg) < 0.5) vr = -vr;    // sign draw after the accepted try
        }
        // the last try of a round that ends without acceptance also signs its v_r:
        // it is the final try when the star runs out of tries
        if (eval && !term && k == CMCG_N_TRY) {
            Rng sg = st;
            if (rng_dbl(sg) < 0.5) vr = -vr;
        }
        const uint64_t tmask = __ballot(eval && term), bmask = __ballot(eval && bad);
        // owners collect their star's lanes [base, base + cs)
        const int my_q = __popcll(pending & ((1ull << lane) - 1ull));
        const bool owner = (pending >> lane) & 1ull;
        const int my_cs = my_q < rem ? c + 1 : c;
        const int base = my_q < rem ? my_q * (c + 1) : rem * (c + 1) + (my_q - rem) * c;
        const uint64_t span = (my_cs >= 64 ? ~0ull : ((1ull << my_cs) - 1ull)) << base;
        const uint64_t tw = tmask & span;
        const int last_k = next + my_cs - 1;               // owner: its last try this round
        int src = -1;
        bool done = false;
        if (owner) {
            if (tw) { src = __ffsll((unsigned long long)tw) - 1; done = true; }
            else if (last_k >= CMCG_N_TRY) { src = base + (CMCG_N_TRY - next); done = true; }
            else src = base + my_cs - 1;        // continue from the round's last try
        }
        const int srcl = src < 0 ? lane : src;
        const double sX = __shfl(X, srcl), srr = __shfl(rr, srcl), svr = __shfl(vr, srcl), spot = __shfl(potr, srcl);
        const int snan = __shfl((int)nan, srcl);
        Rng ns;
        ns.z0 = __shfl(st.z0, srcl); ns.z1 = __shfl(st.z1, srcl); ns.z2 = __shfl(st.z2, srcl); ns.z3 = __shfl(st.z3, srcl);
        if (owner) {
            // bad of the tries up to the final one (or all of this round's if none ended)
            const uint64_t upto = (src >= 63 ? ~0ull : ((1ull << (src + 1)) - 1ull)) & span;
            bad_acc |= (bmask & upto) != 0ull;
            if (done) {
                out.X = sX; out.rr = srr; out.vr = svr; out.potr = spot; out.nan = snan != 0;
                out.k = tw ? next + (src - base) : CMCG_N_TRY + 1;
                out.bad = bad_acc;
            } else {
                next += my_cs;
                rng = ns;
            }
        }
        pending &= ~__ballot(owner && done);
    }
}

// Slot j draws from offset j 2^20; the wave's first slot starts at the per-wave state
// and lane l jumps l 2^20 (l < 2^6) through the LDS tables.  With learned brackets
// the network's prediction proposes both find_zero_Q brackets.
// MINB: blocks per CU the register allocation must allow (1: no constraint; 1024 /
// THREADS: four waves per SIMD, at the cost of some spilled registers)
template <int THREADS, int COARSE, int CAP, int MINB>
struct orbits_kernel {
    static constexpr int LB = THREADS, MB = MINB;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P,
                                                             const RngTable *__restrict__ T,
                                                             const Rng *__restrict__ wave_base,
                                                             double *__restrict__ partials,
                                                             int64_t *__restrict__ defer_list,
                                                             const float2 *__restrict__ pred, float margin,
                                                             const double2 *__restrict__ mid,
                                                             const double2 *__restrict__ nodes,
                                                             const double2 *__restrict__ coarse,
                                                             const NibbleTables<ORB_NIB_POWERS> *__restrict__ nib,
                                                             const uint32_t *__restrict__ ridx, int ridx_bits)
{
    // the jump tables are needed only for the lanes' first states, the try constants
    // only after the apsides: one LDS area serves both
    __shared__ union alignas(16) Share {
        NibbleTables<ORB_NIB_POWERS> lt;
        TryCtx ctx[THREADS];
    } sh;
    __shared__ double cr[COARSE + 1], cphi[COARSE + 1];
    const int64_t n = s->n_max;
    Coarse C;
    C.cr = cr; C.cphi = cphi; C.cs = COARSE; C.n1 = n + 1;
    C.stride = coarse_stride(n, COARSE);
    C.hint = false; C.lrp = C.lra = 0.f; C.margin = margin;
    C.mid = mid;
    C.ri.t = nullptr;
    if (ridx && n >= 1) C.ri = ridx_of(ridx, nodes, n, ridx_bits);
    // both tables are built once per phase: contiguous copies
    for (int sI = threadIdx.x; sI <= COARSE; sI += blockDim.x) {
        const double2 v = coarse[sI];
        cr[sI] = v.x;
        cphi[sI] = v.y;
    }
    {
        constexpr int NQ = (int)(sizeof(NibbleTables<ORB_NIB_POWERS>) / sizeof(uint4));
        const uint4 *src = reinterpret_cast<const uint4 *>(nib);
        uint4 *dst = reinterpret_cast<uint4 *>(&sh.lt);
        for (int q = threadIdx.x; q < NQ; q += blockDim.x) dst[q] = src[q];
    }
    __syncthreads();
    const int64_t j = 1 + xcd_block(kg_) * blockDim.x + threadIdx.x;
    Rng rng{};
    if (j <= n) rng = rng_jump_lds<ORB_NIB_POWERS>(sh.lt, wave_base[(j - 1) >> 6], (uint32_t)((j - 1) & 63));
    __syncthreads();                           // the tables are dead from here
    bool need = false;
    int flags = 0, ntry = 0, hits = 0, misses = 0;
    if (j <= n) {
        const double rj = f.r[j], mj = f.m[j];
        const double Ej = f.E[j], J = f.J[j];
        const double E = Ej + phi_s_own(rj, mj, P.inv_nstar, rj);    // cmc_evolution_thr.c:806
        if (mj < CMCG_ZERO) {
            // destroy_obj (cmc_dynamics_helper.c:131-153) + remove_star_center
            f.vr[j] = 0.0; f.vt[j] = 0.0; f.E[j] = 0.0; f.J[j] = 0.0; f.id[j] = 0;
            f.X[j] = 0.0; f.Y[j] = 0.0; f.r_peri[j] = 0.0; f.r_apo[j] = 0.0;
            f.rnew[j] = CMCG_SF_INFINITY; f.m[j] = 2.2250738585072014e-308;
            f.vrnew[j] = 0.0; f.vtnew[j] = 0.0;
            flags = CMCG_SF_DESTROYED;
        } else if (E >= 0.0) {
            f.rnew[j] = CMCG_SF_INFINITY; f.vrnew[j] = 0.0; f.vtnew[j] = 0.0;
            flags = CMCG_SF_REMOVED_E;
        } else {
            if (q_le0(q_sign(rj, mj, P.inv_nstar, rj, f.phi[j], E, J))) {
                // linear-scan case: finished by the deferred path
                flags = CMCG_SF_LINSCAN;
                unsigned long long slot = atomicAdd((unsigned long long *)&s->n_defer, 1ull);
                if ((int64_t)slot < P.defer_cap) defer_list[slot] = j;
                else dev_set_error(s, CMCG_E_STATE, j);
            } else {
                if (pred) {
                    const float2 y = pred[j];
                    C.hint = true;
                    C.lrp = y.x;
                    C.lra = y.y;
                }
                need = prepare_star(PackedNodes{nodes}, f, s, P, C, n, j, rj, mj, E, J, j, flags, hits, misses,
                                    sh.ctx[threadIdx.x]);
            }
        }
    }
    __syncthreads();                           // try constants visible to the wave
    // every lane of the wave takes part, stars or not
    TryOut to;
    wave_tries<CAP>(PackedNodes{nodes}, C, P, n, need, sh.ctx + (threadIdx.x & ~63), rng, to);
    double red[NRED];
    for (int i = 0; i < NRED; i++) red[i] = 0.0;
    if (j <= n) {
        if (need) {
            const TryCtx &tc = sh.ctx[threadIdx.x];
            if (to.bad) dev_set_error(s, CMCG_E_BRACKET, j);
            if (to.nan) dev_set_error(s, CMCG_E_NAN, j);
            else if (to.k == CMCG_N_TRY + 1) dev_set_error(s, CMCG_E_NTRY, j);
            ntry = to.k;
            f.X[j] = to.X;
            f.r_peri[j] = tc.rmin;
            f.r_apo[j] = tc.rmax;
            f.rnew[j] = to.rr;
            f.vrnew[j] = to.vr;
            f.vtnew[j] = tc.J / to.rr;
            f.pot_rnew[j] = to.potr;
            flags |= 256;      // sampled: pot_rnew holds potential(rnew)
            red[0] = to.rr;
        }
        if (flags & (CMCG_SF_REMOVED_E | CMCG_SF_REMOVED_RT))
            removed(red, f.E[j], f.J[j], f.m[j], P);     // fields the removal left unchanged
        f.flags[j] = flags;
        f.ntry[j] = ntry;
        count_star(red, flags, ntry, hits, misses, true);
    }
    block_reduce_store<THREADS>(red, partials + RED_STRIDE * xcd_block(kg_));
}
};

// Deferred stars (one block of DEFER_T threads): order the list by star index, so
// the reduction of their sums does not depend on the order in which they were
// appended, and reset the first-positive-index slots.  A short list is sorted by one
// thread; a longer one by a bitonic sort over the next power of two (the list's
// capacity is one, above the number of stars), padded with INT64_MAX, in global memory.
constexpr int DEFER_T = 1024;
struct defer_prepare_kernel {
    static constexpr int LB = DEFER_T, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevScalars *s, int64_t *__restrict__ list,
                                                                int64_t *__restrict__ kfirst, int64_t cap)
{
    const int t = threadIdx.x;
    const int64_t nd = s->n_defer < cap ? s->n_defer : cap;
    if (nd <= 64) {
        if (t == 0)
            for (int64_t a = 1; a < nd; a++) {
                const int64_t v = list[a];
                int64_t b = a - 1;
                while (b >= 0 && list[b] > v) { list[b + 1] = list[b]; b--; }
                list[b + 1] = v;
            }
    } else {
        int64_t P = 1;
        while (P < nd) P <<= 1;
        for (int64_t i = nd + t; i < P; i += DEFER_T) list[i] = INT64_MAX;
        __syncthreads();
        for (int64_t k = 2; k <= P; k <<= 1)
            for (int64_t j = k >> 1; j > 0; j >>= 1) {
                for (int64_t i = t; i < P; i += DEFER_T) {
                    const int64_t l = i ^ j;
                    if (l > i) {
                        const int64_t a = list[i], b = list[l];
                        if ((a > b) == ((i & k) == 0)) { list[i] = b; list[l] = a; }
                    }
                }
                __syncthreads();
            }
    }
    const int64_t kf = s->n_max + 2;
    for (int64_t d = t; d < nd; d += DEFER_T) kfirst[d] = kf;
}
};

// Grid-wide equivalent of the linear scan: the first k in [0, N_MAX+1] with Q > 0.
struct defer_search_kernel {
    static constexpr int LB = 256, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P,
                                                           const int64_t *__restrict__ list,
                                                           int64_t *__restrict__ kfirst)
{
    const int64_t nd = s->n_defer < P.defer_cap ? s->n_defer : P.defer_cap;
    if (nd == 0) return;
    const int64_t n = s->n_max;
    const int64_t stride = (int64_t)kg_.nb * blockDim.x;
    for (int64_t d = kg_.by; d < nd; d += kg_.nby) {
        const int64_t j = list[d];
        const double rj = f.r[j], mj = f.m[j];
        const double E = f.E[j] + phi_s_own(rj, mj, P.inv_nstar, rj);
        const double J = f.J[j];
        for (int64_t k = (int64_t)kg_.bid * blockDim.x + threadIdx.x; k <= n + 1; k += stride) {
            if (k > (int64_t)kfirst[d]) break;
            if (function_q(rj, mj, P.inv_nstar, f.r[k], f.phi[k], E, J) > 0.0)
                atomicMin((unsigned long long *)&kfirst[d], (unsigned long long)k);
        }
    }
}
};

// The deferred stars, ORB_THREADS at a time in list order, and the phase's sums: the
// sampler blocks' records (the first level's RED2_BLOCKS records, in block order), then
// each deferred star's in list (= star) order, as a single thread would add them; then
// the end of new_orbits_calculate.
struct defer_tail_kernel {
    static constexpr int LB = ORB_THREADS, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P,
                                                                 const RngTable *__restrict__ T,
                                                                 const Rng *__restrict__ base, int jb1,
                                                                 const int64_t *__restrict__ list,
                                                                 const int64_t *__restrict__ kfirst,
                                                                 const double *__restrict__ red2)
{
    __shared__ double part[ORB_THREADS][NRED];
    __shared__ double acc[NRED];
    const int t = threadIdx.x;
    if (t < NRED) {
        double a = red2[t];
        for (int b = 1; b < RED2_BLOCKS; b++) {
            const double p = red2[RED_STRIDE * b + t];
            a = red_is_max(t) ? (a >= p ? a : p) : a + p;
        }
        acc[t] = a;
    }
    const int64_t nd = s->n_defer < P.defer_cap ? s->n_defer : P.defer_cap;
    const int64_t n = s->n_max;
    for (int64_t c0 = 0; c0 < nd; c0 += ORB_THREADS) {          // block-uniform
        const int64_t d = c0 + t;
        if (d < nd) {
            const int64_t j = list[d];
            const double rj = f.r[j], mj = f.m[j];
            const double Ej = f.E[j], J = f.J[j];
            const double E = Ej + phi_s_own(rj, mj, P.inv_nstar, rj);
            int64_t kt = kfirst[d];
            if (kt >= n) kt = -1;                      // cmc_orbit.c:52-56: circular
            double red[NRED];
            for (int i = 0; i < NRED; i++) red[i] = 0.0;
            int flags = CMCG_SF_LINSCAN, ntry = 0, hits = 0, misses = 0;
            sample_star(SplitNodes{f.r, f.phi}, f, s, P, rng_jump(T, base[1], (uint64_t)j << 20, 20, jb1), n, j,
                        rj, mj, E, Ej, J, kt, flags, ntry, hits, misses, red);
            f.flags[j] = flags;
            f.ntry[j] = ntry;
            count_star(red, flags, ntry, hits, misses, false);
            for (int i = 0; i < NRED; i++) part[t][i] = red[i];
        }
        __syncthreads();
        if (t < NRED) {
            const int cnt = nd - c0 < ORB_THREADS ? (int)(nd - c0) : ORB_THREADS;
            double a = acc[t];
            for (int u = 0; u < cnt; u++) {
                const double p = part[u][t];
                a = red_is_max(t) ? (a >= p ? a : p) : a + p;
            }
            acc[t] = a;
        }
        __syncthreads();
    }
    __syncthreads();
    if (t == 0) {                           // new_orbits_calculate's results
        s->max_r = acc[0];
        s->Eescaped += acc[1];
        s->Jescaped += acc[2];
        s->TidalMassLoss += acc[3];
        s->Etidal += acc[4];
        s->TidalMassLoss += s->TidalMassLoss_old;    // cmc_utils.c:2210
        s->n_circular = (int64_t)acc[5];
        s->n_removed_E = (int64_t)acc[6];
        s->n_removed_rt = (int64_t)acc[7];
        s->n_destroyed = (int64_t)acc[8];
        s->n_brent = (int64_t)acc[9];
        s->n_linscan = (int64_t)acc[10];
        s->tries_total = (int64_t)acc[11];
        s->tries_max = (int64_t)acc[12];
        s->n_bracket_hit = (int64_t)acc[13];
        s->n_bracket_fallback = (int64_t)acc[14];
    }
}
};

// First level of the orbit phase's reduction: block B of RED2_BLOCKS sums the partial
// records of sampler blocks [B C, (B + 1) C), C = ceil(nblocks / RED2_BLOCKS); 16
// consecutive threads read one record (coalesced), its 64 groups take records
// g, g + 64, ... of the range, and the groups are combined in group order.
struct orbits_part_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ partials, int64_t nblocks,
                                                           double *__restrict__ out)
{
    __shared__ double gsum[64][RED_STRIDE];
    const int t = threadIdx.x, i = t & 15, g = t >> 4;
    const int64_t C = (nblocks + RED2_BLOCKS - 1) / RED2_BLOCKS;
    const int64_t b0 = (int64_t)kg_.bid * C, b1 = b0 + C < nblocks ? b0 + C : nblocks;
    double acc = 0.0;
    if (i < NRED)
        for (int64_t b = b0 + g; b < b1; b += 64) {
            double p = partials[RED_STRIDE * b + i];
            acc = red_is_max(i) ? (acc >= p ? acc : p) : acc + p;
        }
    gsum[g][i] = acc;
    __syncthreads();
    if (t < NRED) {
        double a = gsum[0][t];
        for (int g2 = 1; g2 < 64; g2++) a = red_is_max(t) ? (a >= gsum[g2][t] ? a : gsum[g2][t]) : a + gsum[g2][t];
        out[RED_STRIDE * kg_.bid + t] = a;
    }
}
};

struct apsides_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P, int64_t first, int64_t count,
                               const double *__restrict__ E, const double *__restrict__ J,
                               double *__restrict__ out4, int64_t *__restrict__ kout, int32_t *__restrict__ circ)
{
    int64_t i = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int64_t j = first + i, n = s->n_max;
    OrbitRS o = calc_orbit(f.r, f.phi, n, j, f.r[j], f.m[j], P, E[i], J[i]);
    if (o.err) dev_set_error(s, o.err, j);
    out4[4 * i + 0] = o.rp; out4[4 * i + 1] = o.ra; out4[4 * i + 2] = o.dQdrp; out4[4 * i + 3] = o.dQdra;
    kout[2 * i + 0] = o.kmin; kout[2 * i + 1] = o.kmax;
    circ[i] = o.circular;
}
};

// The same per-star apsides written to one array per output (for callers that copy
// each output straight into its own host array).
struct apsides_soa_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P, int64_t first, int64_t count,
                                   const double *__restrict__ E, const double *__restrict__ J, ApsidesOut o)
{
    int64_t i = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int64_t j = first + i, n = s->n_max;
    OrbitRS r = calc_orbit(f.r, f.phi, n, j, f.r[j], f.m[j], P, E[i], J[i]);
    if (r.err) dev_set_error(s, r.err, j);
    o.rp[i] = r.rp; o.ra[i] = r.ra; o.dQdrp[i] = r.dQdrp; o.dQdra[i] = r.dQdra;
    o.kmin[i] = r.kmin; o.kmax[i] = r.kmax;
    o.circular[i] = r.circular;
}
};

struct fzq_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P, int64_t count, const int64_t *__restrict__ jj,
                           const int64_t *__restrict__ kmin, const int64_t *__restrict__ kmax,
                           const double *__restrict__ E, const double *__restrict__ J, int64_t *__restrict__ out)
{
    int64_t i = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int64_t j = jj[i];
    int64_t k = find_zero_Q(f.r, f.phi, f.r[j], f.m[j], P.inv_nstar, kmin[i], kmax[i], E[i], J[i]);
    if (k < 0) dev_set_error(s, CMCG_E_BRACKET, j);
    out[i] = k;
}
};

struct frv_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P, int64_t count, const int64_t *__restrict__ jj,
                           const int64_t *__restrict__ kk, const double *__restrict__ E,
                           const double *__restrict__ J, double *__restrict__ out)
{
    int64_t i = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int64_t j = jj[i];
    const RootVr rv = find_root_vr(f.r, f.phi, f.r[j], f.m[j], P.inv_nstar, P.cenma_m * P.madhoc, kk[i], E[i], J[i]);
    out[i] = rv.root;
    if (rv.bad) dev_set_error(s, CMCG_E_BRENT, j);
}
};

struct pred_from_apsides_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, float2 *__restrict__ pred)
{
    const int64_t j = 1 + (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (j > s->n_max) return;
    const double rp = f.r_peri[j], ra = f.r_apo[j];
    pred[j] = make_float2(rp > 0.0 ? (float)log(rp) : 0.f, ra > 0.0 ? (float)log(ra) : 0.f);
}
};

// ------------------------------------------------------------------ radial index
struct ridx_clear_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, uint32_t *__restrict__ t, int64_t count)
{
    for (int64_t i = (int64_t)kg_.bid * blockDim.x + threadIdx.x; i < count; i += (int64_t)kg_.nb * blockDim.x)
        t[i] = 0;
}
};

// The packed (r, phi) node table with its every MID_STEP-th entry (nodes 0..n+1), and
// the radial index's marks (nodes 0..n), in one pass over r and phi: node k (node 0
// stands for "below every bucket") writes its index at its bucket when it is the
// bucket's last node, and k at every multiple of RIDX_TILE in (bucket(r_k),
// bucket(r_{k+1})] of the tile table.  The index array must be cleared before.
struct node_index_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ r, const double *__restrict__ phi, DevScalars *s,
                                  double2 *__restrict__ nodes, double2 *__restrict__ mid, uint32_t *__restrict__ t,
                                  uint32_t *__restrict__ tb, int bits)
{
    const int64_t n = s->n_max;
    const int64_t k = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (k > n + 1) return;
    const double rk = r[k];
    const double2 v = make_double2(rk, phi[k]);
    nodes[k] = v;
    if (k % MID_STEP == 0) mid[k / MID_STEP] = v;
    if (k > n) return;
    const int64_t nb = (int64_t)1 << bits;
    int64_t bk = -1, bn = nb + 1;
    if (n >= 1) {
        RIdx x;
        x.t = t;
        x.nb = 1u << bits;
        x.lo = (uint64_t)__double_as_longlong(r[1]);
        x.shift = ridx_shift(x.lo, (uint64_t)__double_as_longlong(r[n]), bits);
        if (k >= 1) bk = ridx_bucket(x, rk);
        if (k < n) bn = ridx_bucket(x, r[k + 1]);
        if (k >= 1 && bn != bk) t[bk] = (uint32_t)k;
    }
    for (int64_t m = (bk + RIDX_TILE) / RIDX_TILE; m * RIDX_TILE <= bn; m++) tb[m] = (uint32_t)k;
}
};

// t[b] = the largest marked index below b in its tile, or the tile's entry value:
// an exclusive running maximum, in place.  16 buckets per thread (four 16-byte loads),
// the threads' maxima scanned by wave shuffles and the four wave totals through LDS.
struct ridx_scan_kernel {
    static constexpr int LB = 256, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, uint32_t *__restrict__ t, const uint32_t *__restrict__ tb,
                                                        int bits)
{
    __shared__ uint32_t wtot[4];
    const int64_t entries = ridx_entries(bits);
    const int64_t i0 = (int64_t)kg_.bid * RIDX_TILE + 16 * threadIdx.x;
    const bool whole = i0 + 16 <= entries;
    uint32_t v[16];
    if (whole) {
        const uint4 *p = reinterpret_cast<const uint4 *>(t + i0);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 x = p[q];
            v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = i0 + q < entries ? t[i0 + q] : 0u;
    }
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const uint32_t e = v[q];
        v[q] = m;
        m = e > m ? e : m;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = m;                                   // inclusive max-scan over the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc = o > inc ? o : inc;
    }
    uint32_t ex = __shfl_up(inc, 1, 64);
    if (lane == 0) ex = 0u;
    if (lane == 63) wtot[w] = inc;
    __syncthreads();
    uint32_t start = tb[kg_.bid];
    for (int q = 0; q < w; q++) start = wtot[q] > start ? wtot[q] : start;
    start = ex > start ? ex : start;
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = v[q] > start ? v[q] : start;
    if (whole) {
        uint4 *p = reinterpret_cast<uint4 *>(t + i0);
#pragma unroll
        for (int q = 0; q < 4; q++) p[q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    } else {
#pragma unroll
        for (int q = 0; q < 16; q++)
            if (i0 + q < entries) t[i0 + q] = v[q];
    }
}
};

// The LDS samples of the (r, phi) table for a COARSE-sample launch: node
// min(s * stride, n + 1) for s = 0..cs, from the packed nodes.
struct coarse_build_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double2 *__restrict__ nodes, DevScalars *s, int cs,
                                    double2 *__restrict__ out)
{
    const int i = kg_.bid * blockDim.x + threadIdx.x;
    if (i == 0) {
        // new_orbits_calculate (cmc_utils.c:2198) and the counters of this call
        s->OldTidalMassLoss = s->TidalMassLoss_old;
        s->n_defer = 0;
    }
    if (i > cs) return;
    const int64_t n = s->n_max;
    const int64_t st = coarse_stride(n, cs);
    const int64_t k = (int64_t)i * st;
    out[i] = nodes[k < n + 1 ? k : n + 1];
}
};

// Launch shapes of the orbit sampler that the autotuner chooses from: threads per
// block, LDS samples of the (r, phi) table, and the most lanes one star's tries get
// in a round (64: all idle lanes; fewer: less speculative work).  Every variant gives
// the same stars; only the block order of the escape sums differs.  (Asking the
// compiler for 5 waves per SIMD instead of 4 spills about 100 VGPRs and ran 40%
// slower.)
using OrbFn = void (*)(dim3, dim3, hipStream_t, DevFields, DevScalars *, KParams, const RngTable *, const Rng *,
                       double *, int64_t *, const float2 *, float, const double2 *, const double2 *, const double2 *,
                       const NibbleTables<ORB_NIB_POWERS> *, const uint32_t *, int);
template <int THREADS, int COARSE, int CAP, int MINB>
void orbits_launch(dim3 g, dim3 b, hipStream_t st, DevFields f, DevScalars *s, KParams P, const RngTable *T,
                   const Rng *wb, double *partials, int64_t *defer, const float2 *pred, float margin,
                   const double2 *mid, const double2 *nodes, const double2 *coarse,
                   const NibbleTables<ORB_NIB_POWERS> *nib, const uint32_t *ridx, int ridx_bits)
{
    kl<orbits_kernel<THREADS, COARSE, CAP, MINB>>(g, b, 0, st, f, s, P, T, wb, partials, defer, pred, margin, mid, nodes,
                                                  coarse, nib, ridx, ridx_bits);
}
struct OrbVariant { OrbFn fn; int threads, coarse, cap, minb; };
#define ORB_V(T, C, K, M) {orbits_launch<T, C, K, M>, T, C, K, M}
const OrbVariant kOrbVariants[] = {
    ORB_V(256, 1024, 64, 1), ORB_V(256, 512, 64, 1), ORB_V(256, 2048, 64, 1), ORB_V(128, 1024, 64, 1),
    ORB_V(512, 1024, 64, 1), ORB_V(128, 512, 64, 1), ORB_V(256, 1024, 2, 1),  ORB_V(256, 1024, 4, 1),
    ORB_V(256, 1024, 8, 1),  ORB_V(256, 512, 4, 1),  ORB_V(256, 1024, 4, 4),  ORB_V(256, 1024, 8, 4),
    ORB_V(256, 512, 4, 4)};
#undef ORB_V
constexpr int ORB_NVARIANTS = sizeof(kOrbVariants) / sizeof(kOrbVariants[0]);

int bits_needed(uint64_t x)
{
    int b = 0;
    while (b < 64 && (x >> b)) b++;
    return b;
}
} // namespace

hipError_t launch_orbits(const DevFields &f, DevScalars *s, const KParams &P, const RngTable *T,
                         const Scratch &sc, int64_t n_ub, hipStream_t st, const MlpState *mlp,
                         float margin, hipEvent_t ev_a, hipEvent_t ev_b, int variant, bool pred_fixed)
{
    const OrbVariant &V = kOrbVariants[variant >= 0 && variant < ORB_NVARIANTS ? variant : 0];
    int64_t nblocks = (n_ub + V.threads - 1) / V.threads;
    if (nblocks < 1) nblocks = 1;
    if (nblocks > sc.max_blocks) return hipErrorInvalidValue;
    int jb1 = 20 + bits_needed((uint64_t)n_ub);
    if (hipError_t e = launch_rng_wave_bases(T, sc, 1, n_ub, st)) return e;
    if (hipError_t e = launch_node_index(f, s, sc, n_ub, st)) return e;
    if (V.coarse > ORB_MAX_COARSE) return hipErrorInvalidValue;
    kl<coarse_build_kernel>(dim3((unsigned)((V.coarse + 256) / 256)), dim3(256), 0, st,
                       (const double2 *)sc.nodes, s, V.coarse, sc.coarse);
    const float2 *pred = nullptr;
    if (mlp) {
        if (!sc.pred) return hipErrorInvalidValue;
        if (!pred_fixed)
            if (hipError_t e = launch_mlp_infer_stars(f, s, P, *mlp, sc.pred, n_ub, st)) return e;
        pred = sc.pred;
    }
    if (ev_a) hipEventRecord(ev_a, st);
    V.fn(dim3((unsigned)nblocks), dim3(V.threads), st, f, s, P, T, (const Rng *)sc.rng_wave,
                       sc.partials, sc.defer_list, pred, margin, (const double2 *)sc.mid,
                       (const double2 *)sc.nodes, (const double2 *)sc.coarse,
                       (const NibbleTables<ORB_NIB_POWERS> *)sc.orb_nib, (const uint32_t *)sc.ridx, sc.ridx_bits);
    if (ev_b) hipEventRecord(ev_b, st);
    kl<defer_prepare_kernel>(dim3(1), dim3(DEFER_T), 0, st, s, sc.defer_list, sc.defer_kfirst, P.defer_cap);
    kl<defer_search_kernel>(dim3(256, 4), dim3(256), 0, st, f, s, P, (const int64_t *)sc.defer_list,
                       sc.defer_kfirst);
    kl<orbits_part_kernel>(dim3(RED2_BLOCKS), dim3(1024), 0, st, (const double *)sc.partials, nblocks,
                       sc.red2);
    kl<defer_tail_kernel>(dim3(1), dim3(ORB_THREADS), 0, st, f, s, P, T, sc.rng_base, jb1,
                       (const int64_t *)sc.defer_list, (const int64_t *)sc.defer_kfirst, (const double *)sc.red2);
    return hipGetLastError();
}

hipError_t launch_pred_from_apsides(const DevFields &f, DevScalars *s, const Scratch &sc, int64_t n_ub,
                                    hipStream_t st)
{
    if (!sc.pred) return hipErrorInvalidValue;
    if (n_ub > 0)
        kl<pred_from_apsides_kernel>(dim3((unsigned)((n_ub + 255) / 256)), dim3(256), 0, st, f, s,
                           sc.pred);
    return hipGetLastError();
}

hipError_t launch_node_index(const DevFields &f, DevScalars *s, const Scratch &sc, int64_t n_ub, hipStream_t st)
{
    const int64_t entries = ridx_entries(sc.ridx_bits);
    kl<ridx_clear_kernel>(dim3(512), dim3(256), 0, st, sc.ridx, entries);
    kl<node_index_kernel>(dim3((unsigned)((n_ub + 2 + 255) / 256)), dim3(256), 0, st, f.r, f.phi, s,
                       (double2 *)sc.nodes, (double2 *)sc.mid, sc.ridx, sc.ridx_tb, sc.ridx_bits);
    kl<ridx_scan_kernel>(dim3((unsigned)ridx_tiles(sc.ridx_bits)), dim3(256), 0, st, sc.ridx,
                       (const uint32_t *)sc.ridx_tb, sc.ridx_bits);
    return hipGetLastError();
}

int orbit_variant_count() { return ORB_NVARIANTS; }
void orbit_variant_shape(int v, int *threads, int *coarse, int *cap, int *minb)
{
    *threads = kOrbVariants[v].threads;
    *coarse = kOrbVariants[v].coarse;
    if (cap) *cap = kOrbVariants[v].cap;
    if (minb) *minb = kOrbVariants[v].minb;
}

hipError_t launch_apsides(const DevFields &f, DevScalars *s, const KParams &P, int64_t first, int64_t count,
                          const double *E, const double *J, double *out4, int64_t *kout, int32_t *circ,
                          hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    kl<apsides_kernel>(dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, f, s, P, first,
                       count, E, J, out4, kout, circ);
    return hipGetLastError();
}

hipError_t launch_apsides_soa(const DevFields &f, DevScalars *s, const KParams &P, int64_t first, int64_t count,
                              const double *E, const double *J, const ApsidesOut &o, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    kl<apsides_soa_kernel>(dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, f, s, P, first,
                       count, E, J, o);
    return hipGetLastError();
}

hipError_t launch_find_root_vr(const DevFields &f, DevScalars *s, const KParams &P, int64_t count,
                               const int64_t *j, const int64_t *k, const double *E, const double *J, double *out,
                               hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    kl<frv_kernel>(dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, f, s, P, count, j, k, E,
                       J, out);
    return hipGetLastError();
}

hipError_t launch_find_zero_Q(const DevFields &f, DevScalars *s, const KParams &P, int64_t count,
                              const int64_t *j, const int64_t *kmin, const int64_t *kmax,
                              const double *E, const double *J, int64_t *out, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    kl<fzq_kernel>(dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, f, s, P, count, j, kmin,
                       kmax, E, J, out);
    return hipGetLastError();
}
