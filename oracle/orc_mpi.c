/*
 * The reference's MPI decomposition of the Henon step on P threads of one host —
 * TEST INFRASTRUCTURE ONLY (the CPU baseline timed beside the device, and its tests).
 *
 * The reference runs one MPI rank per core.  Each rank owns a contiguous chunk of the
 * radially sorted stars (findLimits, cmc_utils.c:2299-2333, chunks in multiples of
 * MIN_CHUNK_SIZE = 20, cmc_io.c:1750, the remainder to the last rank) and draws from
 * its own sequential taus113 stream, the seeded state jumped rank * 2^80
 * (set_rng_states, cmc_utils.c:2364-2374).  dynamics_apply and get_positions work on
 * the rank's chunk (cmc_dynamics.c:17-542, cmc_evolution_thr.c:944-1023) and their
 * sums are reduced over the ranks (MPI_Allreduce); the sort is a sample sort
 * (cmc_sort.c:375-629: local sort, splitters from samples, exchange, merge); the
 * potential is computed by every rank over all stars (cmc_utils.c:675-736), which
 * costs what one rank's computation costs.  Here the ranks are threads sharing the
 * arrays, each with its own copy of the scalars.
 *
 * Pinned by: at P = 1 the step is orc_step in REF_SEQ mode bit for bit; at P > 1 the
 * rank streams are the reference's 2^80 jumps (tests/test_oracle_ranks.py).
 */
#define _GNU_SOURCE
#include "cmc_oracle.h"
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* cmc_utils.c:2299-2333 findIndices + findLimits (1-based slots, inclusive ends) */
void orc_find_limits(int64_t N, int64_t blk, int P, int64_t *start, int64_t *end)
{
    const int64_t blocks = N / blk, remain = N % blk;
    const int64_t chunk = (blocks / P) * blk;
    for (int i = 0; i < P; i++) {
        if (i < blocks % P) {
            start[i] = i * chunk + i * blk + 1;
            end[i] = start[i] + chunk + blk - 1;
        } else {
            start[i] = i * chunk + (blocks % P) * blk + 1;
            end[i] = start[i] + chunk - 1;
        }
    }
    end[P - 1] += remain;
}

int orc_ranks_init(orc_ranks *R, orc_state *s, int P)
{
    if (P < 1 || P > ORC_MAX_RANKS) return -1;
    memset(R, 0, sizeof(*R));
    R->s = s;
    R->P = P;
    /* rank i: i jumps by 2^80 = 2^64 * 2^16 */
    for (int i = 0; i < P; i++) R->st[i] = orc_t113_jump(s->seq, 0, (uint64_t)i << 16);
    return 0;
}

/* ---------------------------------------------------------------- rank threads */
enum { PH_RELAX, PH_ORBITS };
typedef struct {
    orc_ranks *R;
    int rank, phase;
    double dt;
    orc_state t;                 /* the rank's copy of the scalars (arrays shared) */
} rank_job;

static void *rank_main(void *arg)
{
    rank_job *jb = (rank_job *)arg;
    orc_ranks *R = jb->R;
    orc_state *t = &jb->t;
    *t = *R->s;
    t->range_lo = R->start[jb->rank];
    t->range_hi = R->end[jb->rank];
    t->rng_mode = ORC_RNG_REF_SEQ;
    t->seq = R->st[jb->rank];
    t->error = 0;
    if (t->range_hi < t->range_lo) return NULL;     /* an empty chunk */
    if (jb->phase == PH_RELAX) {
        t->skip_sigma = 1;
        orc_dynamics_apply(t, jb->dt);
    } else {
        t->Eescaped = t->Jescaped = t->TidalMassLoss = t->Etidal = 0.0;
        t->n_brent = t->n_circular = t->n_removed_E = t->n_removed_tidal = t->n_removed_mass = 0;
        t->n_linear_scan = 0;
        orc_get_positions(t);
    }
    R->st[jb->rank] = t->seq;
    return NULL;
}

static void run_ranks(orc_ranks *R, int phase, double dt, rank_job *jobs)
{
    pthread_t th[ORC_MAX_RANKS];
    for (int i = 0; i < R->P; i++) {
        jobs[i].R = R; jobs[i].rank = i; jobs[i].phase = phase; jobs[i].dt = dt;
        if (i > 0) pthread_create(&th[i], NULL, rank_main, &jobs[i]);
    }
    rank_main(&jobs[0]);
    for (int i = 1; i < R->P; i++) pthread_join(th[i], NULL);
}

static void take_error(orc_state *s, const orc_state *t)
{
    if (!s->error && t->error) { s->error = t->error; s->error_index = t->error_index; }
}

/* ---------------------------------------------------------------- sample sort */
typedef struct {
    const orc_state *s;
    int64_t *perm;            /* slots, sorted per chunk, then merged into out */
    int64_t *out;
    int P, rank;
    int64_t *cs, *ce;         /* chunk bounds (0-based into perm) */
    double *split_r;          /* P-1 splitters (r, id) */
    int64_t *split_id;
    int64_t *obeg;            /* output offset of each rank's key range */
    double **fields; int nf; double *tmp; int64_t *tmpi;
    int64_t n;
} sort_job;

static int key_less(const orc_state *s, int64_t a, int64_t b)
{
    if (s->r[a] != s->r[b]) return s->r[a] < s->r[b];
    return s->id[a] < s->id[b];
}
static int key_less_split(const orc_state *s, int64_t a, double r, int64_t id)
{
    if (s->r[a] != r) return s->r[a] < r;
    return s->id[a] < id;
}

static __thread const orc_state *t_sort_state;
static int cmp_slots(const void *pa, const void *pb)
{
    int64_t a = *(const int64_t *)pa, b = *(const int64_t *)pb;
    if (key_less(t_sort_state, a, b)) return -1;
    if (key_less(t_sort_state, b, a)) return 1;
    return 0;
}

/* first position in perm[lo, hi) whose key is not below the splitter */
static int64_t lower_bound(const orc_state *s, const int64_t *perm, int64_t lo, int64_t hi, double r, int64_t id)
{
    while (lo < hi) {
        int64_t m = lo + (hi - lo) / 2;
        if (key_less_split(s, perm[m], r, id)) lo = m + 1;
        else hi = m;
    }
    return lo;
}

static void *sort_local(void *arg)
{
    sort_job *j = (sort_job *)arg;
    t_sort_state = j->s;
    qsort(j->perm + j->cs[j->rank], (size_t)(j->ce[j->rank] - j->cs[j->rank]), sizeof(int64_t), cmp_slots);
    return NULL;
}

/* rank r merges the parts of every chunk that fall in its key range */
static void *sort_merge(void *arg)
{
    sort_job *j = (sort_job *)arg;
    const orc_state *s = j->s;
    int64_t lo[ORC_MAX_RANKS], hi[ORC_MAX_RANKS];
    for (int c = 0; c < j->P; c++) {
        lo[c] = j->rank == 0 ? j->cs[c]
                             : lower_bound(s, j->perm, j->cs[c], j->ce[c], j->split_r[j->rank - 1], j->split_id[j->rank - 1]);
        hi[c] = j->rank == j->P - 1 ? j->ce[c]
                                    : lower_bound(s, j->perm, j->cs[c], j->ce[c], j->split_r[j->rank], j->split_id[j->rank]);
    }
    int64_t o = j->obeg[j->rank];
    for (;;) {
        int best = -1;
        for (int c = 0; c < j->P; c++)
            if (lo[c] < hi[c] && (best < 0 || key_less(s, j->perm[lo[c]], j->perm[lo[best]]))) best = c;
        if (best < 0) break;
        j->out[o++] = j->perm[lo[best]++];
    }
    return NULL;
}

/* output slots [a, b) of each field permuted by out */
static void *sort_permute(void *arg)
{
    sort_job *j = (sort_job *)arg;
    const int64_t a = j->n * j->rank / j->P, b = j->n * (j->rank + 1) / j->P;
    for (int f = 0; f < j->nf; f++) {
        double *src = j->fields[f];
        if (!src) continue;
        for (int64_t i = a; i < b; i++) j->tmp[(int64_t)f * j->n + i] = src[j->out[i]];
    }
    for (int64_t i = a; i < b; i++) j->tmpi[i] = j->s->id[j->out[i]];
    return NULL;
}

static void run_sort_phase(sort_job *jobs, int P, void *(*fn)(void *))
{
    pthread_t th[ORC_MAX_RANKS];
    for (int i = 1; i < P; i++) pthread_create(&th[i], NULL, fn, &jobs[i]);
    fn(&jobs[0]);
    for (int i = 1; i < P; i++) pthread_join(th[i], NULL);
}

/* qsorts_new on P ranks: the same (r, id) order as orc_sort_by_r, built as a sample
 * sort (cmc_sort.c:375-629), then remove_stripped_stars and post_sort_comm. */
int64_t orc_sort_by_r_ranks(orc_state *s, int P)
{
    const int64_t n = s->N_MAX_NEW;
    if (P <= 1 || n < 64 * (int64_t)P) return orc_sort_by_r(s);
    double *fields[] = { s->r, s->m, s->vr, s->vt, s->E, s->J, s->rnew, s->vrnew, s->vtnew,
                         s->X, s->Y, s->r_peri, s->r_apo, s->rOld, s->Uoldrold, s->Uoldrnew,
                         s->vrold, s->vtold, s->EI };
    const int nf = (int)(sizeof(fields) / sizeof(fields[0]));
    int64_t *perm = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *out = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    double *tmp = (double *)malloc((size_t)n * nf * sizeof(double));
    int64_t *tmpi = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t cs[ORC_MAX_RANKS], ce[ORC_MAX_RANKS], obeg[ORC_MAX_RANKS];
    double split_r[ORC_MAX_RANKS];
    int64_t split_id[ORC_MAX_RANKS];
    sort_job jobs[ORC_MAX_RANKS];
    for (int64_t i = 0; i < n; i++) perm[i] = i + 1;
    for (int i = 0; i < P; i++) { cs[i] = n * i / P; ce[i] = n * (i + 1) / P; }
    for (int i = 0; i < P; i++) {
        sort_job jb = { s, perm, out, P, i, cs, ce, split_r, split_id, obeg, fields, nf, tmp, tmpi, n };
        jobs[i] = jb;
    }
    run_sort_phase(jobs, P, sort_local);
    /* splitters: P - 1 regular samples of each sorted chunk, sorted, every P-th */
    int64_t ns = (int64_t)P * (P - 1);
    int64_t *samp = (int64_t *)malloc((size_t)(ns > 0 ? ns : 1) * sizeof(int64_t));
    int64_t q = 0;
    for (int c = 0; c < P; c++)
        for (int k = 1; k < P; k++) samp[q++] = perm[cs[c] + (ce[c] - cs[c]) * k / P];
    t_sort_state = s;
    qsort(samp, (size_t)ns, sizeof(int64_t), cmp_slots);
    for (int k = 1; k < P; k++) { split_r[k - 1] = s->r[samp[k * (P - 1)]]; split_id[k - 1] = s->id[samp[k * (P - 1)]]; }
    free(samp);
    /* each rank's output offset: the sizes of the key ranges before it */
    int64_t acc = 0;
    for (int r = 0; r < P; r++) {
        obeg[r] = acc;
        for (int c = 0; c < P; c++) {
            int64_t a = r == 0 ? cs[c] : lower_bound(s, perm, cs[c], ce[c], split_r[r - 1], split_id[r - 1]);
            int64_t b = r == P - 1 ? ce[c] : lower_bound(s, perm, cs[c], ce[c], split_r[r], split_id[r]);
            acc += b - a;
        }
    }
    run_sort_phase(jobs, P, sort_merge);
    run_sort_phase(jobs, P, sort_permute);
    for (int f = 0; f < nf; f++)
        if (fields[f]) memcpy(fields[f] + 1, tmp + (int64_t)f * n, (size_t)n * sizeof(double));
    memcpy(s->id + 1, tmpi, (size_t)n * sizeof(int64_t));
    int64_t nn = n;
    while (nn > 0 && s->r[nn] == ORC_SF_INFINITY) nn--;
    s->N_MAX_NEW = nn;
    s->N_MAX = nn;
    free(perm); free(out); free(tmp); free(tmpi);
    return nn;
}

/* ---------------------------------------------------------------- the step */
/* One Henon step in the order of cmc.c:240-407 (orc_step) with the relaxation, the
 * orbit phase and the sort on the ranks. */
void orc_step_ranks(orc_ranks *R)
{
    orc_state *s = R->s;
    const int P = R->P;
    rank_job *jobs = (rank_job *)malloc((size_t)P * sizeof(rank_job));
    double Eescaped_old = s->Eescaped, Jescaped_old = s->Jescaped, Etidal_old = s->Etidal;
    s->TidalMassLoss_old = s->TidalMassLoss;
    s->Eescaped = 0.0; s->Jescaped = 0.0; s->TidalMassLoss = 0.0; s->Etidal = 0.0;

    orc_calc_timestep(s);
    s->N_MAX_NEW = s->N_MAX;
    orc_find_limits(s->N_MAX, 20, P, R->start, R->end);
    orc_calc_sigma_r(s);
    run_ranks(R, PH_RELAX, s->Dt, jobs);
    for (int i = 0; i < P; i++) take_error(s, &jobs[i].t);
    if (s->error) { free(jobs); return; }
    s->Prev_Dt = s->Dt;
    orc_energy_conservation1(s);
    s->OldTidalMassLoss = s->TidalMassLoss_old;
    run_ranks(R, PH_ORBITS, 0.0, jobs);
    double max_r = 0.0;
    for (int i = 0; i < P; i++) {                 /* MPI_Allreduce over the ranks, in rank order */
        const orc_state *t = &jobs[i].t;
        take_error(s, t);
        if (t->range_hi < t->range_lo) continue;
        if (t->max_r > max_r) max_r = t->max_r;
        s->Eescaped += t->Eescaped; s->Jescaped += t->Jescaped;
        s->TidalMassLoss += t->TidalMassLoss; s->Etidal += t->Etidal;
        s->n_brent += t->n_brent; s->n_circular += t->n_circular; s->n_removed_E += t->n_removed_E;
        s->n_removed_tidal += t->n_removed_tidal; s->n_removed_mass += t->n_removed_mass;
        s->n_linear_scan += t->n_linear_scan;
    }
    s->max_r = max_r;
    free(jobs);
    if (s->error) return;
    s->TidalMassLoss += s->TidalMassLoss_old;
    orc_energy_conservation2(s);
    orc_tidally_strip_stars(s);
    orc_compute_intermediate_energy(s);
    orc_sort_by_r_ranks(s, P);
    orc_potential_calculate(s);
    orc_set_velocities3(s);
    s->Eescaped += Eescaped_old;
    s->Jescaped += Jescaped_old;
    s->Etidal += Etidal_old;
    orc_compute_energy(s);
    s->step++;
    /* at one rank the sequential stream is the state's own cursor */
    if (P == 1) s->seq = R->st[0];
}

orc_ranks *orc_ranks_new(orc_state *s, int P)
{
    orc_ranks *R = (orc_ranks *)malloc(sizeof(orc_ranks));
    if (!R) return NULL;
    if (orc_ranks_init(R, s, P)) { free(R); return NULL; }
    return R;
}
void orc_ranks_free(orc_ranks *R) { free(R); }
