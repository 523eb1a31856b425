/*
 * Oracle restatement of the reference RNG (TEST INFRASTRUCTURE ONLY).
 *
 * Reference: src/libs/taus113-v2.c — L'Ecuyer's maximally equidistributed combined
 * Tausworthe generator (Math. Comp. 68 (1999) 261-269), four 32-bit components.
 *
 * Pinned by: the reference's own testjumping() known answer (seed 123456 -> state
 * and the 2^20-th draw 3511752030, committed in tests/golden/rng_golden.json) and by
 * draw-for-draw comparison with oracle/_ref/libtaus113_ref.so (the reference file
 * compiled from source) for several seeds and both jump polynomials.
 */
#include "cmc_oracle.h"
#include <string.h>

/* One step of the four component recurrences, taus113-v2.c:166-177.  The mask
 * 0xffffffff of the reference is implicit in uint32_t arithmetic. */
static inline void t113_advance(orc_t113 *s)
{
    uint32_t b;
    b = ((s->z[0] << 6) ^ s->z[0]) >> 13;
    s->z[0] = ((s->z[0] & 4294967294u) << 18) ^ b;
    b = ((s->z[1] << 2) ^ s->z[1]) >> 27;
    s->z[1] = ((s->z[1] & 4294967288u) << 2) ^ b;
    b = ((s->z[2] << 13) ^ s->z[2]) >> 21;
    s->z[2] = ((s->z[2] & 4294967280u) << 7) ^ b;
    b = ((s->z[3] << 3) ^ s->z[3]) >> 12;
    s->z[3] = ((s->z[3] & 4294967168u) << 13) ^ b;
}

uint32_t orc_t113_int(orc_t113 *st)
{
    t113_advance(st);
    return st->z[0] ^ st->z[1] ^ st->z[2] ^ st->z[3];
}

/* taus113-v2.c:187-189: integer / 2^32 */
double orc_t113_dbl(orc_t113 *st) { return orc_t113_int(st) / 4294967296.0; }

/* taus113-v2.c:197-216: LCG(69069) seeding with lower bounds 2, 8, 16, 128 and ten
 * warm-up draws; seed 0 means 1. */
void orc_t113_seed(uint64_t s, orc_t113 *st)
{
    uint64_t z;
    if (s == 0) s = 1;
    z = (69069ull * s) & 0xffffffffull;            if (z < 2) z += 2;     st->z[0] = (uint32_t)z;
    z = (69069ull * st->z[0]) & 0xffffffffull;     if (z < 8) z += 8;     st->z[1] = (uint32_t)z;
    z = (69069ull * st->z[1]) & 0xffffffffull;     if (z < 16) z += 16;   st->z[2] = (uint32_t)z;
    z = (69069ull * st->z[2]) & 0xffffffffull;     if (z < 128) z += 128; st->z[3] = (uint32_t)z;
    for (int i = 0; i < 10; i++) orc_t113_int(st);
}

/* taus113-v2.c:253-270: XOR of the next 32 states of each component selected by the
 * bits of that component's jump polynomial. */
orc_t113 orc_t113_jump_poly(orc_t113 st, const uint32_t jpoly[4])
{
    orc_t113 out = {{0, 0, 0, 0}};
    uint32_t bit = 1;
    for (int i = 0; i < 32; i++) {
        for (int j = 0; j < 4; j++)
            if (jpoly[j] & bit) out.z[j] ^= st.z[j];
        bit <<= 1;
        t113_advance(&st);
    }
    return out;
}

/* --- arbitrary jumps: each component map is GF(2)-linear on 32 bits, so T^n is a
 * 32x32 bit matrix.  Columns are stored as words: col[i] = T(e_i).  This is an
 * independent construction from the device library's (which is checked against it). */
static uint32_t g_pow[128][4][32];   /* g_pow[b][j] = T_j^(2^b) */
static int g_pow_ready = 0;

static uint32_t comp_step(int j, uint32_t z)
{
    orc_t113 s = {{0, 0, 0, 0}};
    s.z[j] = z;
    t113_advance(&s);
    return s.z[j];
}

static uint32_t matvec(const uint32_t col[32], uint32_t v)
{
    uint32_t out = 0;
    for (int i = 0; i < 32; i++)
        if ((v >> i) & 1u) out ^= col[i];
    return out;
}

static void build_powers(void)
{
    if (g_pow_ready) return;
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 32; i++) g_pow[0][j][i] = comp_step(j, 1u << i);
    for (int b = 1; b < 128; b++)
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 32; i++)
                g_pow[b][j][i] = matvec(g_pow[b - 1][j], g_pow[b - 1][j][i]);
    g_pow_ready = 1;
}

orc_t113 orc_t113_jump(orc_t113 st, uint64_t lo, uint64_t hi)
{
    build_powers();
    for (int b = 0; b < 128; b++) {
        uint64_t w = b < 64 ? lo : hi;
        if (!((w >> (b & 63)) & 1ull)) continue;
        for (int j = 0; j < 4; j++) st.z[j] = matvec(g_pow[b][j], st.z[j]);
    }
    return st;
}

/* T^(2^b): the jump by one power of two (used to walk a layout whose substreams are
 * 2^b apart one substream at a time). */
orc_t113 orc_t113_jump_pow2(orc_t113 st, int b)
{
    build_powers();
    for (int j = 0; j < 4; j++) st.z[j] = matvec(g_pow[b][j], st.z[j]);
    return st;
}

orc_t113 orc_stream_relax(const orc_t113 *seeded, uint64_t step, uint64_t pair)
{
    /* offset = step*2^72 + 3*pair  (3*pair < 2^64 for any realistic N) */
    return orc_t113_jump(*seeded, 3ull * pair, step << 8);
}

orc_t113 orc_stream_orbit(const orc_t113 *seeded, uint64_t step, uint64_t slot)
{
    /* offset = step*2^72 + 2^70 + slot*2^20 */
    return orc_t113_jump(*seeded, slot << 20, (step << 8) | (1ull << 6));
}
