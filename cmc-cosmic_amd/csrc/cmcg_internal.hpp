// Internal declarations of the cmcg HIP library (gfx950).
//
// Device data layout (all in HBM, one allocation per field, SoA):
//   r, m, phi           : N+2 doubles, the reference's star_r/star_m/star_phi with
//                         sentinels at 0 and N+1 (src/cmc/cmc_utils.c:713-714,
//                         src/cmc/cmc_fits.c:128-129)
//   vr, vt, E, J, ...   : per-star fields at slots 1..N (the AoS star_t fields the
//                         path touches, include/cmc/cmc.h:409-601)
// Every double expression below is written in the reference's operand order and the
// library is compiled with -ffp-contract=off, so the results equal the CPU
// reference's on the same inputs (except libm sin/cos, ~1 ulp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/cmcg.h"
#include "seg.hpp"

#define CMCG_SF_INFINITY 1.0e10
#define CMCG_ZERO 1.0e-20
#define CMCG_PI 3.14159265358979323
#define CMCG_N_TRY 50000

// ---------------------------------------------------------------- device scalars
// Lives in device memory so that a whole step can run without host round trips.
struct DevScalars {
    int64_t n_max, n_max_new;
    double Mtotal, Rtidal, max_r, Dt, Prev_Dt, TotalTime;
    double TidalMassLoss, TidalMassLoss_old, OldTidalMassLoss;
    double Eescaped, Jescaped, Etidal, Eoops;
    double Etot, Ekin, Epot;
    uint64_t step;
    double Eescaped_old, Jescaped_old, Etidal_old;
    // per-call work
    double DTidalMassLoss;
    int32_t error; int32_t pad0;
    int64_t error_index;
    int64_t reserved0;
    int64_t n_excess;            // set_velocities3 stars with vnew2 <= 0
    int64_t first_inf;           // potential_calculate: first slot with r >= SF_INFINITY
    int64_t n_defer;             // orbit phase: stars whose own-index Q <= 0 (linear-scan path)
    int64_t sort_n;              // sort: live stars after dropping the r == SF_INFINITY tail
    int64_t sort_in;             // sort: slots sorted (N_MAX_NEW before the sort)
    // orbit statistics
    int64_t n_circular, n_removed_E, n_removed_rt, n_destroyed, n_brent, n_linscan;
    int64_t tries_total, tries_max, n_bracket_hit, n_bracket_fallback;
    double dtrel;
};

// RNG jump table: g_pow[b][j][i] = column i of T_j^(2^b), b < 128.
struct RngTable { uint32_t col[128][4][32]; };

struct Rng { uint32_t z0, z1, z2, z3; };

// ---------------------------------------------------------------- device fields
struct DevFields {
    double *r, *m, *phi;
    double *vr, *vt, *E, *J;
    int64_t *id;
    double *rnew, *vrnew, *vtnew, *X, *Y, *r_peri, *r_apo;
    double *rOld, *Uoldrold, *Uoldrnew, *vrold, *vtold, *EI;
    double *sigma, *pot_rnew;
    int32_t *ntry, *flags;
};

struct KParams {
    int64_t n_star;
    double inv_nstar;      // 1.0/N_STAR exactly as the reference's (1.0/clus.N_STAR)
    double madhoc;         // 1.0/((double)N_STAR), cmc_utils.c:1183
    double cenma_m;
    double orbit_r;
    double theta_se_max;
    double gamma;
    int64_t avekernel;
    int32_t chain_k_drop;  // test hook: potential chains' error bound widened by 2^chain_k_drop
    int32_t pad0;
    int64_t defer_cap;     // capacity of the deferred-star list: a power of two above the star capacity
};

__device__ __forceinline__ void dev_set_error(DevScalars *s, int code, int64_t idx)
{
    if (atomicCAS(&s->error, 0, code) == 0) s->error_index = idx;
}

// ---------------------------------------------------------------- taus113
// One step of src/libs/taus113-v2.c:166-177 (32-bit arithmetic replaces the MASK).
__device__ __forceinline__ uint32_t rng_int(Rng &s)
{
    uint32_t b;
    b = ((s.z0 << 6) ^ s.z0) >> 13;  s.z0 = ((s.z0 & 4294967294u) << 18) ^ b;
    b = ((s.z1 << 2) ^ s.z1) >> 27;  s.z1 = ((s.z1 & 4294967288u) << 2) ^ b;
    b = ((s.z2 << 13) ^ s.z2) >> 21; s.z2 = ((s.z2 & 4294967280u) << 7) ^ b;
    b = ((s.z3 << 3) ^ s.z3) >> 12;  s.z3 = ((s.z3 & 4294967168u) << 13) ^ b;
    return s.z0 ^ s.z1 ^ s.z2 ^ s.z3;
}
// taus113-v2.c:187-189
__device__ __forceinline__ double rng_dbl(Rng &s) { return (double)rng_int(s) / 4294967296.0; }

__device__ __forceinline__ uint32_t gf2_matvec(const uint32_t *__restrict__ col, uint32_t v)
{
    uint32_t out = 0;
#pragma unroll
    for (int i = 0; i < 32; i++) out ^= col[i] & (0u - ((v >> i) & 1u));
    return out;
}

// Jump the state forward by `off` steps, using bits [b0, b1) of `off` (the caller
// knows the rest are zero).  Loops are uniform over bit positions, so the matrix
// columns are wave-uniform loads.
__device__ __forceinline__ Rng rng_jump(const RngTable *__restrict__ T, Rng s, uint64_t off, int b0, int b1)
{
    for (int b = b0; b < b1; b++) {
        if (!((off >> b) & 1ull)) continue;
        s.z0 = gf2_matvec(T->col[b][0], s.z0);
        s.z1 = gf2_matvec(T->col[b][1], s.z1);
        s.z2 = gf2_matvec(T->col[b][2], s.z2);
        s.z3 = gf2_matvec(T->col[b][3], s.z3);
    }
    return s;
}

// Block-cooperative jump: s advanced by 2^(row0 + k) steps for every k < nbits (<= 64)
// with bit k of `bits` set.  Every thread of the block calls it (blockDim.x >= 128).
// The block first stages the columns of those matrix powers in LDS (cols: 128 * nbits
// words), then thread (c, i) = (t / 32, t % 32), t < 128, keeps word c of the state
// and contributes column i of a power when bit i of the word is set: an XOR over the
// 32 lanes is the product, and the chain of products never waits on memory.  The
// four words advance independently.  res: an LDS slot for the result.
__device__ __forceinline__ Rng rng_jump_block(const RngTable *__restrict__ T, Rng s, uint64_t bits, int row0,
                                              int nbits, uint32_t *cols, Rng *res)
{
    const int t = threadIdx.x;
    const uint32_t *src = &T->col[row0][0][0];
    for (int e = t; e < nbits * 128; e += blockDim.x)
        if ((bits >> (e >> 7)) & 1ull) cols[e] = src[e];
    __syncthreads();
    if (t < 128) {
        const int c = t >> 5, i = t & 31;
        uint32_t z = c == 0 ? s.z0 : c == 1 ? s.z1 : c == 2 ? s.z2 : s.z3;
        for (int k = 0; k < nbits; k++) {
            if (!((bits >> k) & 1ull)) continue;
            uint32_t v = cols[k * 128 + t] & (0u - ((z >> i) & 1u));
#pragma unroll
            for (int d = 16; d > 0; d >>= 1) v ^= __shfl_xor(v, d, 32);
            z = v;
        }
        if (i == 0) {
            if (c == 0) res->z0 = z;
            else if (c == 1) res->z1 = z;
            else if (c == 2) res->z2 = z;
            else res->z3 = z;
        }
    }
    __syncthreads();
    const Rng out = *res;
    __syncthreads();          // cols and res free again
    return out;
}

// In-wave jumps through LDS: for the NB matrix powers T^(2^b), b = b0..b0+NB-1, and
// each state word, tab[b][c][q][v] is the product of the matrix with the nibble v
// placed at bits 4q..4q+3, so a 32x32 GF(2) product is 8 table lookups.  Lanes of a
// wave read 16 consecutive words per lookup: no bank conflicts.
template <int NB>
struct NibbleTables { uint32_t t[NB][4][8][16]; };

template <int NB>
__device__ void build_nibble_tables(const RngTable *__restrict__ T, int b0, NibbleTables<NB> &lt)
{
    for (int idx = threadIdx.x; idx < NB * 4 * 8 * 16; idx += blockDim.x) {
        const int v = idx & 15, q = (idx >> 4) & 7, c = (idx >> 7) & 3, b = idx >> 9;
        const uint32_t *col = T->col[b0 + b][c];
        uint32_t o = 0;
        for (int i = 0; i < 4; i++)
            if ((v >> i) & 1) o ^= col[4 * q + i];
        lt.t[b][c][q][v] = o;
    }
    __syncthreads();
}

// copy of tables built once per context (same contents as build_nibble_tables): 16-byte
// loads, no column XORs per block
template <int NB>
__device__ void copy_nibble_tables(const NibbleTables<NB> *__restrict__ src, NibbleTables<NB> &lt)
{
    constexpr int NQ = (int)(sizeof(NibbleTables<NB>) / sizeof(uint4));
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
    uint4 *d4 = reinterpret_cast<uint4 *>(&lt);
    for (int q = threadIdx.x; q < NQ; q += blockDim.x) d4[q] = s4[q];
    __syncthreads();
}

__device__ __forceinline__ uint32_t nib_matvec(const uint32_t (&tab)[8][16], uint32_t v)
{
    uint32_t o = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) o ^= tab[q][(v >> (4 * q)) & 15u];
    return o;
}

// advance by off * 2^b0 steps, off < 2^NB
template <int NB>
__device__ __forceinline__ Rng rng_jump_lds(const NibbleTables<NB> &lt, Rng s, uint32_t off)
{
    for (int b = 0; b < NB; b++) {
        if (!((off >> b) & 1u)) continue;
        s.z0 = nib_matvec(lt.t[b][0], s.z0);
        s.z1 = nib_matvec(lt.t[b][1], s.z1);
        s.z2 = nib_matvec(lt.t[b][2], s.z2);
        s.z3 = nib_matvec(lt.t[b][3], s.z3);
    }
    return s;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  This
// bijection gives XCD x a contiguous range of logical blocks, so the blocks that
// share an L2 work on neighbouring stars and their search paths overlap.
__device__ __forceinline__ int64_t xcd_block(const KGrid &kg)
{
    const int64_t nb = kg.nb, b = kg.bid;
    const int64_t x = b & 7, local = b >> 3, q = nb >> 3, rem = nb & 7;
    return x * q + (x < rem ? x : rem) + local;
}

// ---------------------------------------------------------------- potential helpers
// include/cmc/cmc.h:2175 MPI_PHI_S(rad, j)
__device__ __forceinline__ double phi_s(const double *__restrict__ r, const double *__restrict__ m,
                                        double inv_nstar, double rad, int64_t j)
{
    return (rad >= r[j] ? m[j] / rad : m[j] / r[j]) * inv_nstar;
}
__device__ __forceinline__ double phi_s_own(double rj, double mj, double inv_nstar, double rad)
{
    return (rad >= rj ? mj / rad : mj / rj) * inv_nstar;
}

// cmc_utils.c:1654-1662 function_q evaluated in double (the reference uses x87
// long double here; the oracle measures how often the sign differs).
__device__ __forceinline__ double function_q(double rj, double mj, double inv_nstar,
                                             double r, double pot, double E, double J)
{
    double Jor = J / r;
    double Jr = Jor * Jor;
    double phis = (r >= rj ? mj / r : mj / rj) * inv_nstar;
    return 2.0 * (E - (pot + phis)) - Jr;
}

// Sign of function_q(rj, mj, inv_nstar, r, pot, E, J) as that function computes it
// in double: -1 (Q < 0), 0 (Q == 0), 1 (Q > 0), 2 (NaN).  A division-free value q~
// (1/r from the hardware reciprocal refined by two Newton steps, accurate to 2^-45
// for any seed accurate to 12 bits) differs from the computed Q by less than
// (2 2^-45 + 10 2^-53) S, S = 2(|E| + |pot| + |phis|) + (J/r)^2, since each of the two
// evaluations rounds at most six times; when |q~| exceeds 2^-42 S the signs agree,
// otherwise function_q itself decides.  qa receives q~ (a guide for searches only).
__device__ __forceinline__ int q_sign(double rj, double mj, double inv_nstar, double r, double pot, double E,
                                      double J, double *qa = nullptr)
{
    double y = __builtin_amdgcn_rcp(r);
    double e = __builtin_fma(-r, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-r, y, 1.0);
    y = __builtin_fma(y, e, y);
    const double a = J * y, T = a * a;
    const double ph = (r >= rj ? mj * y : mj / rj) * inv_nstar;
    const double q = 2.0 * (E - (pot + ph)) - T;
    if (qa) *qa = q;
    const double S = 2.0 * (__builtin_fabs(E) + __builtin_fabs(pot) + __builtin_fabs(ph)) + T;
    if (__builtin_fabs(q) > 0x1p-42 * S) return q < 0.0 ? -1 : 1;
    const double qe = function_q(rj, mj, inv_nstar, r, pot, E, J);
    return qe < 0.0 ? -1 : (qe > 0.0 ? 1 : (qe == 0.0 ? 0 : 2));
}
__device__ __forceinline__ bool q_lt0(int sg) { return sg == -1; }
__device__ __forceinline__ bool q_gt0(int sg) { return sg == 1; }
__device__ __forceinline__ bool q_ge0(int sg) { return sg == 0 || sg == 1; }
__device__ __forceinline__ bool q_le0(int sg) { return sg == -1 || sg == 0; }

// Access to the (r, phi) nodes: the two global arrays, or a packed copy of 16-byte
// (r, phi) nodes, whose Q evaluations and interpolations touch one cache line per
// node instead of two.  Both hold the same values.
struct SplitNodes {
    const double *__restrict__ r;
    const double *__restrict__ phi;
    __device__ __forceinline__ double R(int64_t k) const { return r[k]; }
    __device__ __forceinline__ double P(int64_t k) const { return phi[k]; }
    __device__ __forceinline__ void RP(int64_t k, double &a, double &b) const { a = r[k]; b = phi[k]; }
};
struct PackedNodes {
    const double2 *__restrict__ rp;
    __device__ __forceinline__ double R(int64_t k) const { return rp[k].x; }
    __device__ __forceinline__ double P(int64_t k) const { return rp[k].y; }
    __device__ __forceinline__ void RP(int64_t k, double &a, double &b) const
    {
        double2 v = rp[k];
        a = v.x;
        b = v.y;
    }
};

// cmc_nr.c:62-82 FindZero_r: largest k in [kmin, kmax] with r[k] < x (for sorted r).
template <class N>
__device__ __forceinline__ int64_t find_zero_r_n(const N &nd, int64_t kmin, int64_t kmax, double x)
{
    int guard = 0;
    do {
        int64_t ktry = (kmin + kmax + 1) / 2;
        if (nd.R(ktry) < x) kmin = ktry;
        else kmax = ktry - 1;
    } while (kmax != kmin && ++guard < 96);
    return kmin;
}
__device__ __forceinline__ int64_t find_zero_r(const double *__restrict__ r, int64_t kmin, int64_t kmax, double x)
{
    return find_zero_r_n(SplitNodes{r, nullptr}, kmin, kmax, x);
}

// cmc_utils.c:112-165 potential(r); the bisection bracket [lo, hi] may be narrowed
// by the caller to any [lo, hi] with r[lo] < x <= r[hi] (lo >= 1): the result is the
// unique largest k with r[k] < x, identical to the full-range search.
template <class N>
__device__ __forceinline__ double potential_at_n(const N &nd, int64_t n_max, double x, int64_t lo, int64_t hi,
                                                 bool *bad)
{
    if (x == 0.0) return nd.P(0);
    double r1 = nd.R(1);
    if (x < r1) {
        double p0 = nd.P(0);
        return p0 - (p0 - nd.P(1)) * r1 / x;
    }
    int64_t i;
    if (lo == hi - 1) i = lo;
    else i = find_zero_r_n(nd, lo, hi, x);
    if (i > n_max) { *bad = true; return 0.0; }   // x beyond the outer sentinel
    double ri, pi, ri1, pi1;
    nd.RP(i, ri, pi);
    nd.RP(i + 1, ri1, pi1);
    if (ri > x || ri1 < x) { *bad = true; return 0.0; }
    if (i == 0) return pi1;
    return pi + (pi1 - pi) * (1.0 / ri - 1.0 / x) / (1.0 / ri - 1.0 / ri1);
}
__device__ __forceinline__ double potential_at(const double *__restrict__ r, const double *__restrict__ phi,
                                               int64_t n_max, double x, int64_t lo, int64_t hi, bool *bad)
{
    return potential_at_n(SplitNodes{r, phi}, n_max, x, lo, hi, bad);
}

// cmc_nr.c:316-383 find_zero_Q.  Returns -1 if the bracket cannot close.
template <class N>
__device__ __forceinline__ int64_t find_zero_Q_n(const N &nd, double rj, double mj, double inv_nstar,
                                                 int64_t kmin, int64_t kmax, double E, double J)
{
    double a, b, a2, b2;
    nd.RP(kmin, a, b);
    nd.RP(kmax, a2, b2);
    // the direction test qmin < qmax: decided by the signs when they differ
    const int smin = q_sign(rj, mj, inv_nstar, a, b, E, J), smax = q_sign(rj, mj, inv_nstar, a2, b2, E, J);
    bool up;
    if ((smin == -1 || smin == 1) && (smax == -1 || smax == 1) && smin != smax) up = smin < smax;
    else up = function_q(rj, mj, inv_nstar, a, b, E, J) < function_q(rj, mj, inv_nstar, a2, b2, E, J);
    int guard = 0;
    if (up) {
        do {
            int64_t ktry = (kmin + kmax + 1) / 2;
            nd.RP(ktry, a, b);
            if (q_lt0(q_sign(rj, mj, inv_nstar, a, b, E, J))) kmin = ktry;
            else kmax = ktry - 1;
        } while (kmax != kmin && ++guard < 96);
    } else {
        do {
            int64_t ktry = (kmin + kmax + 1) / 2;
            nd.RP(ktry, a, b);
            if (q_gt0(q_sign(rj, mj, inv_nstar, a, b, E, J))) kmin = ktry;
            else kmax = ktry - 1;
        } while (kmax != kmin && ++guard < 96);
    }
    return guard >= 96 ? -1 : kmin;
}
__device__ __forceinline__ int64_t find_zero_Q(const double *__restrict__ r, const double *__restrict__ phi,
                                               double rj, double mj, double inv_nstar,
                                               int64_t kmin, int64_t kmax, double E, double J)
{
    return find_zero_Q_n(SplitNodes{r, phi}, rj, mj, inv_nstar, kmin, kmax, E, J);
}

// cmc_orbit.c:722-740 set_a_b
template <class N>
__device__ __forceinline__ void set_a_b_n(const N &nd, double rj, double mj, const KParams &P, int64_t k,
                                          double *a, double *b)
{
    if (k == 0) {
        double r1, p1;
        nd.RP(1, r1, p1);
        *a = -P.cenma_m * P.madhoc;
        *b = p1 + P.cenma_m * P.madhoc / r1 + mj * P.madhoc / rj;
    } else {
        double rk, pk, rk1, pk1;
        nd.RP(k, rk, pk);
        nd.RP(k + 1, rk1, pk1);
        double Uk = pk + phi_s_own(rj, mj, P.inv_nstar, rk);
        double Uk1 = pk1 + phi_s_own(rj, mj, P.inv_nstar, rk1);
        *a = (Uk1 - Uk) / (1. / rk1 - 1. / rk);
        *b = (Uk / rk1 - Uk1 / rk) / (1. / rk1 - 1. / rk);
    }
}
__device__ __forceinline__ void set_a_b(const double *__restrict__ r, const double *__restrict__ phi,
                                        double rj, double mj, const KParams &P, int64_t k,
                                        double *a, double *b)
{
    set_a_b_n(SplitNodes{r, phi}, rj, mj, P, k, a, b);
}

// cmc_nr.c:239-269 + :393-403: v_r^2 inside [r_k, r_{k+1}]; cen = cenma.m * madhoc
__device__ __forceinline__ double vr2_in_interval(const double *__restrict__ r, const double *__restrict__ phi,
                                                  double rj, double mj, double inv_nstar, double cen, int64_t k,
                                                  double E, double J, double x)
{
    const double eps = 2.220446049250313e-16;
    double pot;
    if (fabs(x - r[k]) < eps) pot = phi[k];
    else if (fabs(x - r[k + 1]) < eps) pot = phi[k + 1];
    else if (x < r[1]) pot = phi[0] - cen / x;
    else pot = (phi[k] + (phi[k + 1] - phi[k]) * (1.0 / r[k] - 1.0 / x) / (1.0 / r[k] - 1.0 / r[k + 1]));
    return function_q(rj, mj, inv_nstar, x, pot, E, J);
}

__device__ __forceinline__ bool interval_converged(double lo, double hi, double epsabs, double epsrel)
{
    double alo = fabs(lo), ahi = fabs(hi), min_abs;
    if ((lo > 0.0 && hi > 0.0) || (lo < 0.0 && hi < 0.0)) min_abs = alo < ahi ? alo : ahi;
    else min_abs = 0.0;
    return fabs(hi - lo) < epsabs + epsrel * min_abs;
}

// cmc_nr.c:500-603 find_root_vr by the Brent-Dekker method (Brent 1973, ch. 4),
// the same iteration as the CPU oracle's.  Rare fallback path; scalar arguments and a
// returned pair keep its caller's objects out of private memory.
struct RootVr { double root; int bad; };
__device__ __noinline__ RootVr find_root_vr(const double *__restrict__ r, const double *__restrict__ phi,
                                            double rj, double mj, double inv_nstar, double cen, int64_t k,
                                            double E, double J)
{
    bool bad_ = false, *bad = &bad_;
    const double prec = 1e-11, deps = 2.220446049250313e-16;
    int iter = 100;
    double a = r[k], b = r[k + 1];
    double fa = vr2_in_interval(r, phi, rj, mj, inv_nstar, cen, k, E, J, a);
    double fb = vr2_in_interval(r, phi, rj, mj, inv_nstar, cen, k, E, J, b);
    if ((fa > 0) == (fb > 0) && (fa < 0) == (fb < 0)) return RootVr{0.0, 1};
    double c = a, fc = fa, d = b - a, e = d;
    double lo = a < b ? a : b, hi = a < b ? b : a;
    bool converged = false;
    double prev_apsis = -1.0;
    while (!converged && iter-- > 0) {
        if ((fb > 0 && fc > 0) || (fb < 0 && fc < 0)) { c = a; fc = fa; d = b - a; e = d; }
        if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
        double tol = 2.0 * deps * fabs(b);
        double mm = 0.5 * (c - b);
        if (fb == 0.0) { lo = hi = b; converged = true; break; }
        if (fabs(e) < tol || fabs(fa) <= fabs(fb)) {
            d = mm; e = mm;
        } else {
            double p, q, r2, s2 = fb / fa;
            if (a == c) {
                p = 2.0 * mm * s2; q = 1.0 - s2;
            } else {
                q = fa / fc; r2 = fb / fc;
                p = s2 * (2.0 * mm * q * (q - r2) - (b - a) * (r2 - 1.0));
                q = (q - 1.0) * (r2 - 1.0) * (s2 - 1.0);
            }
            if (p > 0) q = -q; else p = -p;
            if (2.0 * p < fmin(3.0 * mm * q - fabs(tol * q), fabs(e * q))) { e = d; d = p / q; }
            else { d = mm; e = mm; }
        }
        a = b; fa = fb;
        if (fabs(d) > tol) b += d;
        else b += (mm > 0 ? tol : -tol);
        fb = vr2_in_interval(r, phi, rj, mj, inv_nstar, cen, k, E, J, b);
        lo = b < c ? b : c; hi = b < c ? c : b;
        if ((fb > 0 && fc > 0) || (fb < 0 && fc < 0)) { lo = a < b ? a : b; hi = a < b ? b : a; }
        converged = interval_converged(lo, hi, prec, prec);
        if (!converged && iter < 10) {
            if (prev_apsis >= 0.0 && fabs(b - prev_apsis) < 5e-13 + 5e-13 * fabs(b)) converged = true;
            prev_apsis = b;
        }
    }
    if (!converged) *bad = true;
    double apsis = b;
    if (vr2_in_interval(r, phi, rj, mj, inv_nstar, cen, k, E, J, apsis) < 0.) {
        if (vr2_in_interval(r, phi, rj, mj, inv_nstar, cen, k, E, J, lo) < 0.) apsis = hi;
        if (vr2_in_interval(r, phi, rj, mj, inv_nstar, cen, k, E, J, hi) < 0.) apsis = lo;
    }
    if (vr2_in_interval(r, phi, rj, mj, inv_nstar, cen, k, E, J, apsis) < 0.) *bad = true;
    return RootVr{apsis, bad_ ? 1 : 0};
}

// Top levels of the sorted-potential search tree staged in LDS: cr[s] = r[k(s)],
// cphi[s] = phi[k(s)] with k(s) = min(s * stride, n + 1), s = 0..cs (k(cs) = n + 1).
// Radial index of the sorted radii: bucket b of a radius x > 0 is
// (bits(x) - bits(r_1)) >> shift (the IEEE bits of a positive double order like the
// values), clamped to nb; t[b] = number of nodes k in 1..n whose bucket is below b
// (t has nb + 2 entries).  Every node up to t[b] lies below x and every node after
// t[b + 1] above it, so [t[b], t[b + 1] + 1] brackets the largest k with r_k < x.
struct RIdx {
    const uint32_t *t;
    uint64_t lo;
    int shift;
    uint32_t nb;
};
__device__ __forceinline__ int ridx_shift(uint64_t lo, uint64_t hi, int bits)
{
    const uint64_t d = hi > lo ? hi - lo : 0;
    const int len = d ? 64 - __clzll((long long)d) : 0;
    return len > bits ? len - bits : 0;
}
__device__ __forceinline__ uint32_t ridx_bucket(const RIdx &x, double r)
{
    const uint64_t u = (uint64_t)__double_as_longlong(r);
    const uint64_t d = u > x.lo ? (u - x.lo) >> x.shift : 0;
    return d < x.nb ? (uint32_t)d : x.nb;
}
// the index of nodes 1..n (packed (r, phi), n >= 1) with 2^bits buckets
__device__ __forceinline__ RIdx ridx_of(const uint32_t *t, const double2 *__restrict__ nodes, int64_t n, int bits)
{
    RIdx x;
    x.t = t;
    x.nb = 1u << bits;
    x.lo = (uint64_t)__double_as_longlong(nodes[1].x);
    x.shift = ridx_shift(x.lo, (uint64_t)__double_as_longlong(nodes[n].x), bits);
    return x;
}

// [lo, hi] with r_lo < x <= r_hi (x >= r_1, x <= r_{n+1})
__device__ __forceinline__ void ridx_bracket(const RIdx &x, double r, int32_t &lo, int32_t &hi)
{
    const uint32_t b = ridx_bucket(x, r);
    lo = (int32_t)x.t[b];
    hi = (int32_t)x.t[b + 1] + 1;
}

struct Coarse {
    const double *cr, *cphi;
    int64_t stride, n1;          // stride is a multiple of MID_STEP
    int cs;
    // middle level in global memory (L2-resident): (r, phi) at every MID_STEP-th node
    const double2 *mid;
    // learned bracket proposal: predicted log r_peri, log r_apo and the half-width
    bool hint;
    float lrp, lra, margin;
    // radial index (t == nullptr: not built)
    RIdx ri;
};

constexpr int MID_STEP = 16;

__device__ __forceinline__ int64_t coarse_stride(int64_t n, int cs)
{
    int64_t st = (n + 1 + cs - 1) / cs;
    st = (st + MID_STEP - 1) / MID_STEP * MID_STEP;
    return st < MID_STEP ? MID_STEP : st;
}

__device__ __forceinline__ int64_t coarse_k(const Coarse &c, int s)
{
    int64_t k = (int64_t)s * c.stride;
    return k < c.n1 ? k : c.n1;
}

// Index near radius x: the LDS sample interval holding x, then linear interpolation
// in log r inside it (r is sorted; only an estimate, every use is verified).
__device__ __forceinline__ int64_t coarse_index_estimate(const Coarse &c, double x)
{
    int a = 0, b = c.cs;
    while (a < b) {
        int m = (a + b + 1) >> 1;
        if (c.cr[m] < x) a = m;
        else b = m - 1;
    }
    if (a >= c.cs) return c.n1;
    int64_t ka = coarse_k(c, a), kb = coarse_k(c, a + 1);
    float la = __logf((float)c.cr[a]), lb = __logf((float)c.cr[a + 1]), lx = __logf((float)x);
    float fr = lb > la ? (lx - la) / (lb - la) : 0.f;
    fr = fr >= 0.f ? (fr > 1.f ? 1.f : fr) : 0.f;      // also maps NaN to 0
    return ka + (int64_t)(fr * (float)(kb - ka));
}

// Largest t in [t0, t1] with eval(mid[t]) true, or t0 - 1 if there is none, for a
// predicate that is true then false over [t0, t1].  eval also returns a key, negative
// where the predicate holds, that varies smoothly with t; k_lo and k_hi are the keys
// just outside the range (at the bracket ends, k_lo < 0 <= k_hi).  Up to two probes
// read a whole 128-byte line of the middle table (8 entries) at the index where the
// keys interpolate to zero; the rest is bisection.  The answer is the same for any
// guesses, since the predicate alone decides it.
template <class Eval>
__device__ __forceinline__ int64_t mid_search(const double2 *__restrict__ mid, int64_t t0, int64_t t1, double k_lo,
                                              double k_hi, Eval eval)
{
    int64_t lo = t0 - 1, hi = t1 + 1;         // eval true at lo, false at hi (virtually)
    for (int probe = 0; probe < 2 && hi - lo > 1; probe++) {
        double f = k_lo / (k_lo - k_hi);
        if (!(f > 0.0 && f < 1.0)) f = 0.5;   // also NaN
        int64_t g = lo + 1 + (int64_t)(f * (double)(hi - lo - 1));
        if (g >= hi) g = hi - 1;
        const int64_t ls = g & ~(int64_t)7;
        int64_t last_true = lo, first_false = hi;
        double kt = k_lo, kf = k_hi;
        double2 e[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int64_t t = ls + u;
            if (t > lo && t < hi) e[u] = mid[t];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int64_t t = ls + u;
            if (t > lo && t < hi && first_false == hi) {
                double k;
                if (eval(e[u], k)) { last_true = t; kt = k; }
                else { first_false = t; kf = k; }
            }
        }
        lo = last_true; k_lo = kt;
        hi = first_false; k_hi = kf;
    }
    while (hi - lo > 1) {
        const int64_t m = lo + ((hi - lo) >> 1);
        double k;
        if (eval(mid[m], k)) lo = m;
        else hi = m;
    }
    return lo;
}

// potential(x) with the index search started from the LDS samples: the result is the
// unique largest k with r[k] < x, so any bracket [lo, hi] with r[lo] < x <= r[hi]
// gives the same index as the full-range search.
template <class N>
__device__ __forceinline__ double potential_at_coarse_n(const N &nd, int64_t n_max, double x, int64_t lo, int64_t hi,
                                                        const Coarse &c, bool *bad)
{
    if (x == 0.0 || x < nd.R(1)) return potential_at_n(nd, n_max, x, lo, hi, bad);
    int a = 0, b = c.cs;                    // largest s with cr[s] < x (cr[0] = r[0] < x)
    while (a < b) {
        int m = (a + b + 1) >> 1;
        if (c.cr[m] < x) a = m;
        else b = m - 1;
    }
    const float lx = __logf((float)x);
    double k_lo = -1.0, k_hi = 1.0;
    if (a < c.cs) {
        int64_t ka = coarse_k(c, a), kb = coarse_k(c, a + 1);
        if (ka > lo) lo = ka;
        if (kb < hi) hi = kb;
        k_lo = (double)(__logf((float)c.cr[a]) - lx);
        k_hi = (double)(__logf((float)c.cr[a + 1]) - lx);
    }
    if (c.mid) {
        // largest mid node strictly inside (lo, hi) with r < x
        const int64_t t0 = (lo + MID_STEP) / MID_STEP, t1 = (hi - 1) / MID_STEP;
        if (t0 <= t1) {
            const int64_t t = mid_search(c.mid, t0, t1, k_lo, k_hi, [&](double2 e, double &k) {
                k = (double)(__logf((float)e.x) - lx);
                return e.x < x;
            });
            if (t < t0) {
                hi = MID_STEP * t0;
            } else {
                lo = MID_STEP * t;
                if (MID_STEP * (t + 1) < hi) hi = MID_STEP * (t + 1);
            }
        }
    }
    return potential_at_n(nd, n_max, x, lo, hi, bad);
}
__device__ __forceinline__ double potential_at_coarse(const double *__restrict__ r, const double *__restrict__ phi,
                                                      int64_t n_max, double x, int64_t lo, int64_t hi,
                                                      const Coarse &c, bool *bad)
{
    return potential_at_coarse_n(SplitNodes{r, phi}, n_max, x, lo, hi, c, bad);
}

// potential(x) with the index search started from the radial index: one bucket read
// gives a bracket of a few nodes (the answer is the unique largest k with r_k < x,
// so any valid bracket gives the full-range search's index), then the caller's
// bracket [lo, hi] (valid when r_lo < x <= r_hi) narrows it further.
template <class N>
__device__ __forceinline__ double potential_at_ridx_n(const N &nd, int64_t n_max, double x, int64_t lo, int64_t hi,
                                                      const RIdx &ri, bool *bad)
{
    if (x == 0.0 || x < nd.R(1)) return potential_at_n(nd, n_max, x, lo, hi, bad);
    int32_t a, b;
    ridx_bracket(ri, x, a, b);
    if (a > lo) lo = a;
    if (b < hi) hi = b;
    return potential_at_n(nd, n_max, x, lo, hi, bad);
}

// Narrow a find_zero_Q bracket [lo, hi] with pred(Q(lo)) true and pred(Q(hi)) false
// (pred monotone over the bracket) to the mid-node interval holding the sign change.
// q_lo, q_hi: Q at the bracket ends when known (they guide the first probes only).
template <bool POSITIVE>
__device__ __forceinline__ void mid_narrow_q(const Coarse &c, double rj, double mj, double inv, double E, double J,
                                             int64_t &lo, int64_t &hi, double q_lo = 0.0, double q_hi = 0.0)
{
    if (!c.mid) return;
    const int64_t t0 = (lo + MID_STEP) / MID_STEP, t1 = (hi - 1) / MID_STEP;
    if (t0 > t1) return;
    // key negative where the predicate holds: Q for Q < 0, -Q for Q > 0
    double k_lo = POSITIVE ? -q_lo : q_lo, k_hi = POSITIVE ? -q_hi : q_hi;
    if (!(k_lo < 0.0 && k_hi >= 0.0)) { k_lo = -1.0; k_hi = 1.0; }
    const int64_t t = mid_search(c.mid, t0, t1, k_lo, k_hi, [&](double2 e, double &k) {
        double q;
        const int sg = q_sign(rj, mj, inv, e.x, e.y, E, J, &q);
        k = POSITIVE ? -q : q;
        return POSITIVE ? q_gt0(sg) : q_lt0(sg);
    });
    if (t < t0) { hi = MID_STEP * t0; return; }
    lo = MID_STEP * t;
    if (MID_STEP * (t + 1) < hi) hi = MID_STEP * (t + 1);
}

// A guess of the radius where Q changes sign between nodes (ra, pa) and (rb, pb),
// ra < rb: the potential between them taken as the Henon interpolant A/r + B through
// the two nodes (exactly the reference's interpolation when they are adjacent) plus
// the star's own term (m_j/r_j below r_j, m_j/r above; judged at ra), so that
// Q r^2 = 2(E - B - s0) r^2 - 2(A + s1) r - J^2.  The smaller positive root for the
// pericentre, the larger for the apocentre; -1 when there is none.  Only a guess:
// every index it leads to is verified by the signs of Q.
__device__ __forceinline__ double interp_root(double ra, double pa, double rb, double pb, double rj, double mj,
                                              double inv, double E, double J, bool peri)
{
    const double ia = 1.0 / ra, ib = 1.0 / rb;
    const double A = (pb - pa) / (ib - ia), B = pa - A * ia;
    double s0 = 0.0, s1 = 0.0;
    if (ra >= rj) s1 = mj * inv;
    else s0 = mj * inv / rj;
    const double a2 = 2.0 * (E - B - s0), a1 = -2.0 * (A + s1), a0 = -J * J;
    const double disc = a1 * a1 - 4.0 * a2 * a0;
    if (!(disc >= 0.0) || a2 == 0.0) return -1.0;
    const double sq = sqrt(disc);
    const double q = -0.5 * (a1 + (a1 >= 0.0 ? sq : -sq));
    if (q == 0.0) return -1.0;
    double x1 = q / a2, x2 = a0 / q;
    if (x1 > x2) { const double t = x1; x1 = x2; x2 = t; }
    if (x2 <= 0.0) return -1.0;
    if (x1 <= 0.0) return x2;
    return peri ? x1 : x2;
}

// find_zero_Q's answer on a bracket [lo, hi] whose ends have the bisection's
// predicate true and false (POSITIVE: the apocentre, predicate Q > 0; otherwise the
// pericentre, Q < 0), for a Q with one sign change on the bracket: interpolation
// steps (interp_root on the current bracket ends, turned into a node bracket by the
// radial index and probed at its two ends) narrow it, then a bisection in the nodes
// finishes.  Every index kept is a probed node, so the answer is the unique sign
// change, the index the full-range bisection returns.  (rl, pl), (rh, ph): nodes lo
// and hi.
template <bool POSITIVE, class N>
__device__ __forceinline__ int64_t ridx_zero_Q(const N &nd, const RIdx &ri, double rj, double mj, double inv,
                                               double E, double J, int64_t lo, int64_t hi, double rl, double pl,
                                               double rh, double ph)
{
    auto pred = [&](double r, double p) {
        const int sg = q_sign(rj, mj, inv, r, p, E, J);
        return POSITIVE ? q_gt0(sg) : q_lt0(sg);
    };
    int32_t l = (int32_t)lo, h = (int32_t)hi;
    for (int it = 0; it < 3 && h - l > 8; it++) {
        const double rg = interp_root(rl, pl, rh, ph, rj, mj, inv, E, J, !POSITIVE);
        int32_t a, b;
        if (rg > rl && rg < rh) {
            ridx_bracket(ri, rg, a, b);
            a = a < l ? l : (a > h ? h : a);
            b = b > h ? h : (b < l ? l : b);
        } else {
            a = b = l + ((h - l) >> 1);        // no usable guess: a bisection step
        }
        if (a > l && a < h) {
            double r, p;
            nd.RP(a, r, p);
            if (pred(r, p)) { l = a; rl = r; pl = p; }
            else { h = a; rh = r; ph = p; }
        }
        if (b > l && b < h) {
            double r, p;
            nd.RP(b, r, p);
            if (pred(r, p)) { l = b; rl = r; pl = p; }
            else { h = b; rh = r; ph = p; }
        }
    }
    while (h - l > 1) {
        const int32_t m = l + ((h - l) >> 1);
        double r, p;
        nd.RP(m, r, p);
        if (pred(r, p)) l = m;
        else h = m;
    }
    return l;
}

struct OrbitRS { double rp, ra, dQdrp, dQdra; int64_t kmin, kmax; int circular; int brent; int linscan; int err;
                 int hint_hits, hint_misses; };

// cmc_orbit.c:413-540 calc_orbit_new after get_positive_Q_index has produced ktemp
// (an index with Q > 0; ktemp < 0 means "no positive Q": circular orbit).
// HAS_C: the LDS samples, radial index and learned proposal of Cr are available.  (A
// template flag rather than a null test of a pointer: comparing the address of a
// private object with null is not folded on this target and keeps the object in
// scratch memory.)
template <class N, bool HAS_C>
__device__ __forceinline__ OrbitRS calc_orbit_core(const N &nd, const double *__restrict__ r,
                                                   const double *__restrict__ phi, int64_t n_max, int64_t j,
                                                   double rj, double mj, const KParams &P, double E, double J,
                                                   int64_t ktemp, const Coarse &Cr)
{
    OrbitRS o;
    o.brent = 0; o.linscan = 0; o.err = 0; o.hint_hits = 0; o.hint_misses = 0;
    if (ktemp < 0) {
        o.rp = o.ra = rj; o.kmin = o.kmax = j; o.dQdrp = o.dQdra = 0.0; o.circular = 1;
        return o;
    }
    // Q is negative at the centre, positive on one interval around ktemp and negative
    // far out (r^3 dPhi/dr = r M(r) increases), so the index bisection's result is
    // the unique sign change of its bracket; the LDS samples narrow both brackets
    // to one sample interval before the bisection in global memory.
    int64_t kp_lo = 0, kp_hi = ktemp, ka_lo = ktemp, ka_hi = n_max + 1;
    bool got_p = false, got_a = false;
    bool vp = false, va = false;       // bracket ends checked: Q(lo), Q(hi) have the bisection's signs
    double qp_lo = 0.0, qp_hi = 0.0, qa_lo = 0.0, qa_hi = 0.0;   // Q at sample ends (search guidance only)
    if (HAS_C && Cr.hint) {
        // learned proposals: the index range of the predicted apsis +- margin (in
        // log r), padded by 2, accepted only when Q has the bisection's signs at its ends
        const double inv = P.inv_nstar;
        const float ep = __expf(Cr.margin);
        // node index just below / at or above radius x: exact bounds from the radial
        // index, estimates padded by two nodes from the LDS samples otherwise
        auto below = [&](double x) -> int64_t {
            if (Cr.ri.t) { int32_t a, b; ridx_bracket(Cr.ri, x, a, b); return a; }
            return coarse_index_estimate(Cr, x) - 2;
        };
        auto above = [&](double x) -> int64_t {
            if (Cr.ri.t) { int32_t a, b; ridx_bracket(Cr.ri, x, a, b); return b; }
            return coarse_index_estimate(Cr, x) + 2;
        };
        int64_t lo = below((double)(__expf(Cr.lrp) / ep));
        int64_t hi = above((double)(__expf(Cr.lrp) * ep));
        lo = lo < 0 ? 0 : lo;
        hi = hi > ktemp ? ktemp : hi;
        if (lo < hi && q_lt0(q_sign(rj, mj, inv, nd.R(lo), nd.P(lo), E, J)) &&
            (hi == ktemp || q_ge0(q_sign(rj, mj, inv, nd.R(hi), nd.P(hi), E, J)))) {
            kp_lo = lo; kp_hi = hi; got_p = vp = true; o.hint_hits++;
        } else {
            o.hint_misses++;
        }
        lo = below((double)(__expf(Cr.lra) / ep));
        hi = above((double)(__expf(Cr.lra) * ep));
        lo = lo < ktemp ? ktemp : lo;
        hi = hi > n_max + 1 ? n_max + 1 : hi;
        if (lo < hi && (lo == ktemp || q_gt0(q_sign(rj, mj, inv, nd.R(lo), nd.P(lo), E, J))) &&
            !q_gt0(q_sign(rj, mj, inv, nd.R(hi), nd.P(hi), E, J))) {
            ka_lo = lo; ka_hi = hi; got_a = va = true; o.hint_hits++;
        } else {
            o.hint_misses++;
        }
    }
    if (HAS_C) {
        const double inv = P.inv_nstar;
        // a verified learned bracket is narrowed by the LDS samples inside it: the
        // sample bisection runs over those samples only
        auto narrow = [&](int64_t &lo, int64_t &hi, bool positive) {
            const int s0 = (int)(lo / Cr.stride) + 1, s1 = (int)((hi - 1) / Cr.stride);
            if (s0 > s1 || s1 > Cr.cs) return;
            int a = s0 - 1, b = s1;                    // sample a: predicate true (a = s0 - 1: node lo)
            while (a < b) {
                const int m = (a + b + 1) >> 1;
                const int sg = q_sign(rj, mj, inv, Cr.cr[m], Cr.cphi[m], E, J);
                if (positive ? q_gt0(sg) : q_lt0(sg)) a = m;
                else b = m - 1;
            }
            if (a >= s0) lo = coarse_k(Cr, a);
            if (a < s1) hi = coarse_k(Cr, a + 1);
        };
        if (got_p) narrow(kp_lo, kp_hi, false);
        if (got_a) narrow(ka_lo, ka_hi, true);
        if (!got_p && ktemp >= 1 && q_lt0(q_sign(rj, mj, inv, Cr.cr[0], Cr.cphi[0], E, J))) {
            int a = 0, b = (int)((ktemp - 1) / Cr.stride);
            const int top = b;
            while (a < b) {
                int m = (a + b + 1) >> 1;
                if (q_lt0(q_sign(rj, mj, inv, Cr.cr[m], Cr.cphi[m], E, J))) a = m;
                else b = m - 1;
            }
            kp_lo = coarse_k(Cr, a);
            kp_hi = a < top ? coarse_k(Cr, a + 1) : ktemp;
            vp = true;
            (void)q_sign(rj, mj, inv, Cr.cr[a], Cr.cphi[a], E, J, &qp_lo);
            if (a < top) (void)q_sign(rj, mj, inv, Cr.cr[a + 1], Cr.cphi[a + 1], E, J, &qp_hi);
        }
        const int sb = (int)(ktemp / Cr.stride) + 1;
        if (!got_a && sb <= Cr.cs && !q_gt0(q_sign(rj, mj, inv, Cr.cr[Cr.cs], Cr.cphi[Cr.cs], E, J))) {
            va = true;
            if (!q_gt0(q_sign(rj, mj, inv, Cr.cr[sb], Cr.cphi[sb], E, J))) {
                ka_hi = coarse_k(Cr, sb);
            } else {
                int a = sb, b = Cr.cs - 1;
                while (a < b) {
                    int m = (a + b + 1) >> 1;
                    if (q_gt0(q_sign(rj, mj, inv, Cr.cr[m], Cr.cphi[m], E, J))) a = m;
                    else b = m - 1;
                }
                ka_lo = coarse_k(Cr, a);
                ka_hi = coarse_k(Cr, a + 1);
                (void)q_sign(rj, mj, inv, Cr.cr[a], Cr.cphi[a], E, J, &qa_lo);
                (void)q_sign(rj, mj, inv, Cr.cr[a + 1], Cr.cphi[a + 1], E, J, &qa_hi);
            }
        }
    }
    // verified brackets: interpolation search through the radial index; otherwise the
    // middle table and a bisection (which also settles the direction itself)
    const bool ri = HAS_C && Cr.ri.t;
    int64_t kmin, kmax;
    if (ri && vp) {
        double rl, pl, rh, ph;
        nd.RP(kp_lo, rl, pl);
        nd.RP(kp_hi, rh, ph);
        kmin = ridx_zero_Q<false>(nd, Cr.ri, rj, mj, P.inv_nstar, E, J, kp_lo, kp_hi, rl, pl, rh, ph);
    } else {
        if (vp && HAS_C) mid_narrow_q<false>(Cr, rj, mj, P.inv_nstar, E, J, kp_lo, kp_hi, qp_lo, qp_hi);
        kmin = find_zero_Q_n(nd, rj, mj, P.inv_nstar, kp_lo, kp_hi, E, J);
    }
    if (ri && va) {
        double rl, pl, rh, ph;
        nd.RP(ka_lo, rl, pl);
        nd.RP(ka_hi, rh, ph);
        kmax = ridx_zero_Q<true>(nd, Cr.ri, rj, mj, P.inv_nstar, E, J, ka_lo, ka_hi, rl, pl, rh, ph);
    } else {
        if (va && HAS_C) mid_narrow_q<true>(Cr, rj, mj, P.inv_nstar, E, J, ka_lo, ka_hi, qa_lo, qa_hi);
        kmax = find_zero_Q_n(nd, rj, mj, P.inv_nstar, ka_lo, ka_hi, E, J);
    }
    if (kmin < 0 || kmax < 0) {
        o.err = CMCG_E_BRACKET;
        o.rp = o.ra = rj; o.kmin = o.kmax = j; o.dQdrp = o.dQdra = 0.0; o.circular = 1;
        return o;
    }
    double a, b, rmin, rmax, inside_sqrt;
    bool bad = false;
    set_a_b_n(nd, rj, mj, P, kmin, &a, &b);
    inside_sqrt = a * a - 2.0 * J * J * (b - E);
    if (inside_sqrt < 0.0) {
        rmin = -1.0 * J * J / a;
        if ((rmin < nd.R(kmin)) || (rmin > nd.R(kmin + 1))) {
            const RootVr rv = find_root_vr(r, phi, rj, mj, P.inv_nstar, P.cenma_m * P.madhoc, kmin, E, J);
            rmin = rv.root;
            bad |= rv.bad != 0;
            o.brent = 1;
        }
    } else {
        rmin = J * J / (-a + sqrt(inside_sqrt));
    }
    double dQdr_min = 2.0 * J * J / (rmin * rmin * rmin) + 2.0 * a / (rmin * rmin);
    set_a_b_n(nd, rj, mj, P, kmax, &a, &b);
    inside_sqrt = a * a - 2.0 * J * J * (b - E);
    if (inside_sqrt < 0.0) {
        rmax = -a / (2.0 * (b - E));
        if ((rmax < nd.R(kmax)) || (rmax > nd.R(kmax + 1))) {
            const RootVr rv = find_root_vr(r, phi, rj, mj, P.inv_nstar, P.cenma_m * P.madhoc, kmax, E, J);
            rmax = rv.root;
            bad |= rv.bad != 0;
            o.brent = 1;
        }
    } else {
        rmax = (-a + sqrt(inside_sqrt)) / (2.0 * (b - E));
    }
    double dQdr_max = 2.0 * J * J / (rmax * rmax * rmax) + 2.0 * a / (rmax * rmax);
    if (bad) o.err = CMCG_E_BRENT;
    if ((rmin > rmax) || (dQdr_min <= 0.) || (dQdr_max >= 0.)) {
        o.rp = o.ra = rj; o.kmin = o.kmax = j; o.dQdrp = o.dQdra = 0.0; o.circular = 1;
    } else {
        o.rp = rmin; o.ra = rmax; o.kmin = kmin; o.kmax = kmax;
        o.dQdrp = dQdr_min; o.dQdra = dQdr_max; o.circular = 0;
    }
    return o;
}

template <class N>
__device__ __forceinline__ OrbitRS calc_orbit_from_n(const N &nd, const double *__restrict__ r,
                                                     const double *__restrict__ phi, int64_t n_max, int64_t j,
                                                     double rj, double mj, const KParams &P, double E, double J,
                                                     int64_t ktemp)
{
    const Coarse none{};
    return calc_orbit_core<N, false>(nd, r, phi, n_max, j, rj, mj, P, E, J, ktemp, none);
}

__device__ __forceinline__ OrbitRS calc_orbit_from(const double *__restrict__ r, const double *__restrict__ phi,
                                                   int64_t n_max, int64_t j, double rj, double mj,
                                                   const KParams &P, double E, double J, int64_t ktemp)
{
    return calc_orbit_from_n(SplitNodes{r, phi}, r, phi, n_max, j, rj, mj, P, E, J, ktemp);
}

// get_positive_Q_index (cmc_orbit.c:30-59) with its linear scan, then calc_orbit_from.
// Used by the small batch entry points; the orbit phase defers the scan instead.
__device__ __forceinline__ OrbitRS calc_orbit(const double *__restrict__ r, const double *__restrict__ phi,
                                              int64_t n_max, int64_t j, double rj, double mj,
                                              const KParams &P, double E, double J)
{
    int64_t ktemp = j;
    double Qtemp = function_q(rj, mj, P.inv_nstar, r[ktemp], phi[ktemp], E, J);
    int lin = 0;
    if (Qtemp <= 0.0) {
        lin = 1;
        ktemp = -1;
        do {
            ktemp++;
            Qtemp = function_q(rj, mj, P.inv_nstar, r[ktemp], phi[ktemp], E, J);
        } while (Qtemp <= 0.0 && ktemp <= n_max);
        if (ktemp >= n_max) ktemp = -1;            // cmc_orbit.c:52-56
    }
    OrbitRS o = calc_orbit_from(r, phi, n_max, j, rj, mj, P, E, J, ktemp);
    o.linscan = lin;
    return o;
}

// ---------------------------------------------------------------- host side
struct cmcg_ctx_impl;   // defined in cmcg_api.hip

#define CMCG_HIP_CHECK(expr)                                                       \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) { cmcg_report_hip(_e, #expr, __FILE__, __LINE__); return CMCG_E_HIP; } \
    } while (0)

void cmcg_report_hip(hipError_t e, const char *expr, const char *file, int line);
