/*
 * covsearch_oracle.c -- CPU oracle for the covering-code tabu search.
 *
 * TEST INFRASTRUCTURE ONLY: the checker, never the product.  It restates the
 * search of cov/engine/covsearch2e.c (reference file; line numbers below refer
 * to it) so that the HIP engine can be compared with it bit for bit, and it is
 * itself pinned against the reference solver compiled from source (golden
 * fixtures under tests/golden/).
 *
 * Semantics restated (all integer, all deterministic):
 *   rng          xoshiro256++ seeded by four splitmix64 steps; a draw modulo m
 *                is next() % m                                   (:201-220)
 *   word index   MSB-first base q: w = sum_j d_j q^(n-1-j)      (:916-921)
 *   cnt[w]       number of codewords within distance R (uint16)  (:69)
 *   ball         radius-R ball of a codeword, +-1 on cnt          (:353-380)
 *   move delta   pos - neg over the C(n-1,R)(q-1)^R patterns      (:450-558)
 *   shared eval  leaving T - H_p from the full distance-R sphere, entering
 *                from walks or from the uncovered set            (:696-760)
 *   commit       --cnt[a], ++cnt[b] over the same patterns        (:562-590)
 *   pick         64 probes, then a cyclic scan from a random start (:973-1016)
 *   step         candidates, partial Fisher-Yates, reservoir tie-break,
 *                tabu with aspiration, tabu stamp, kick          (:1304-1580)
 *   init         random digits + rebuild, greedy placement, or remove-and-
 *                repair descent from an oversized seed           (:1156-1239)
 *
 * The evaluator here is written for clarity first (per-subset counts over an
 * odometer), with an exact uncovered-word set to keep large cells testable.
 */
#include "covsearch_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN 32

struct cso {
    int q, n, R, M;
    int64_t N;
    int64_t pw[MAXN];
    uint16_t *cnt;
    int64_t uncovered;
    uint8_t *code, *best, *tmpw;
    int64_t *cidx;
    int64_t best_unc;
    int32_t *tabu;
    uint64_t rs[4];
    int64_t ndraw;       /* rng draws consumed */
    int64_t it, since, nkicks;
    int kick_order;
    int threads;
    int64_t target;      /* stop once best_unc <= target (< 0: none)        */
    int64_t ttt_it;      /* iteration that first reached the target, or -1  */

    /* --wide (:1353-1376): every move of a candidate codeword is a
     * candidate while the reference's uncovered-word list is valid and
     * short.  Whether it is depends on that list's own bookkeeping (valid
     * flag, length counting stale entries), which rl_* restate
     * (:286-338, :1326-1329); the entries themselves are not needed. */
    int wide;
    int rl_valid;
    int64_t rl_n;
    uint64_t *rl_bit;    /* N bits: the word is in the list          */
    int64_t *rl_idx;     /* RL_CAP entries                           */
    int64_t *dvecw;      /* M * n * q deltas of every move           */

    /* position subsets: subs[r] lists all r-subsets of 0..n-1 (lexicographic),
     * nsub[r] of them; psub[p] lists the R-subsets avoiding p. */
    uint8_t *subs[MAXN + 1];
    int nsub[MAXN + 1];
    uint8_t *psub;       /* n * npsub * R */
    int npsub;

    /* exact uncovered set (only when N < 2^31) */
    int use_set;
    int32_t *upos;       /* N, -1 if covered */
    int64_t *uidx;       /* members */
    uint8_t *udig;       /* member digits, n each */
    int64_t nu;

    /* step scratch */
    int *cand_i, *cand_p, *cand_v, *dist, *cw_list, *cw_mark;
    int64_t *cand_d, *dvec;
    int64_t ncap;
};

/* ------------------------------------------------------------------ rng */
/* xoshiro256++ (Blackman & Vigna), seeded with splitmix64 (:201-220) */
static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(uint64_t *s)
{
    const uint64_t out = rotl64(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return out;
}

static void rng_seed(uint64_t *s, uint64_t seed)
{
    for (int i = 0; i < 4; i++) {
        seed += 0x9E3779B97F4A7C15ULL;
        uint64_t z = seed;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        s[i] = z ^ (z >> 31);
    }
}

static inline uint64_t draw(cso_t *s) { s->ndraw++; return rng_next(s->rs); }
static inline uint64_t rng_mod(cso_t *s, uint64_t m) { return draw(s) % m; }

/* ------------------------------------------------------------ helpers */
static int64_t binom64(int a, int b)
{
    if (b < 0 || b > a) return 0;
    if (b > a - b) b = a - b;
    int64_t r = 1;
    for (int i = 0; i < b; i++) r = r * (a - i) / (i + 1);
    return r;
}

/* all r-subsets of the positions in `pos` (np of them), lexicographic */
static uint8_t *enum_subsets(const int *pos, int np, int r, int *count)
{
    int64_t cnt = binom64(np, r);
    *count = (int)cnt;
    uint8_t *out = malloc((size_t)(cnt ? cnt : 1) * (size_t)(r ? r : 1));
    if (cnt == 0) return out;
    int c[MAXN];
    for (int i = 0; i < r; i++) c[i] = i;
    for (int64_t k = 0; k < cnt; k++) {
        for (int i = 0; i < r; i++) out[k * r + i] = (uint8_t)pos[c[i]];
        int i = r - 1;
        while (i >= 0 && c[i] == np - r + i) i--;
        if (i < 0) break;
        c[i]++;
        for (int j = i + 1; j < r; j++) c[j] = c[j - 1] + 1;
    }
    return out;
}

static int64_t widx(const cso_t *s, const uint8_t *w)
{
    int64_t x = 0;
    for (int j = 0; j < s->n; j++) x += (int64_t)w[j] * s->pw[j];
    return x;
}

static void word_of(const cso_t *s, int64_t idx, uint8_t *w)
{
    for (int j = 0; j < s->n; j++) w[j] = (uint8_t)((idx / s->pw[j]) % s->q);
}

/* offs[j*(q-1)+k] = (v_k - c_j) * pw[j] with v_k the k-th value != c_j (:340-347) */
static void make_offs(const cso_t *s, const uint8_t *c, int64_t *offs)
{
    const int q1 = s->q - 1;
    for (int j = 0; j < s->n; j++) {
        int k = 0;
        for (int v = 0; v < s->q; v++)
            if (v != c[j]) offs[j * q1 + k++] = (int64_t)(v - c[j]) * s->pw[j];
    }
}

/* ------------------------------------------------ uncovered-set upkeep */
static inline void set_add(cso_t *s, int64_t w)
{
    if (!s->use_set || s->upos[w] >= 0) return;
    s->upos[w] = (int32_t)s->nu;
    s->uidx[s->nu] = w;
    word_of(s, w, s->udig + (size_t)s->nu * s->n);
    s->nu++;
}

static inline void set_del(cso_t *s, int64_t w)
{
    if (!s->use_set) return;
    int64_t k = s->upos[w];
    if (k < 0) return;
    int64_t last = --s->nu;
    if (k != last) {
        int64_t lw = s->uidx[last];
        s->uidx[k] = lw;
        memcpy(s->udig + (size_t)k * s->n, s->udig + (size_t)last * s->n, s->n);
        s->upos[lw] = (int32_t)k;
    }
    s->upos[w] = -1;
}

/* the reference's list bookkeeping: insert on cnt -> 0 (:294-302),
 * compaction (:305-320) and rebuild in ascending order (:323-338) */
#define RL_CAP ((int64_t)1 << 20)
static inline void rl_insert(cso_t *s, int64_t w)
{
    if (!s->wide || !s->rl_valid || ((s->rl_bit[w >> 6] >> (w & 63)) & 1)) return;
    if (s->rl_n >= RL_CAP) { s->rl_valid = 0; return; }
    s->rl_bit[w >> 6] |= 1ULL << (w & 63);
    s->rl_idx[s->rl_n++] = w;
}

static void rl_compact(cso_t *s)
{
    int64_t m = 0;
    for (int64_t e = 0; e < s->rl_n; e++) {
        const int64_t w = s->rl_idx[e];
        if (s->cnt[w] == 0) s->rl_idx[m++] = w;
        else s->rl_bit[w >> 6] &= ~(1ULL << (w & 63));
    }
    s->rl_n = m;
}

static void rl_rebuild(cso_t *s)
{
    memset(s->rl_bit, 0, (size_t)((s->N + 63) / 64) * 8);
    s->rl_n = 0;
    s->rl_valid = 1;
    for (int64_t w = 0; w < s->N; w++) {
        if (s->cnt[w]) continue;
        if (s->rl_n >= RL_CAP) { s->rl_valid = 0; return; }
        s->rl_bit[w >> 6] |= 1ULL << (w & 63);
        s->rl_idx[s->rl_n++] = w;
    }
}

/* cnt transitions: every +-1 on cnt goes through these two */
static inline int64_t cnt_inc(cso_t *s, int64_t w)
{
    if (++s->cnt[w] == 1) { set_del(s, w); return -1; }
    return 0;
}

static inline int64_t cnt_dec(cso_t *s, int64_t w)
{
    if (--s->cnt[w] == 0) { set_add(s, w); rl_insert(s, w); return 1; }
    return 0;
}

/* --------------------------------------------------- pattern walking */
/* Visit the (q-1)^r words base + sum_i offs[ps[i]][k_i]; the odometer runs the
 * last position fastest.  BODY sees the word index as `w_`. */
#define WALK(s, offs, ps, r, base, BODY)                                       \
    do {                                                                       \
        const int q1_ = (s)->q - 1;                                            \
        if ((r) == 0) { int64_t w_ = (base); BODY; break; }                    \
        int64_t part_[MAXN + 1];                                               \
        int k_[MAXN];                                                          \
        part_[0] = (base);                                                     \
        for (int i_ = 0; i_ < (r) - 1; i_++) {                                 \
            k_[i_] = 0;                                                        \
            part_[i_ + 1] = part_[i_] + (offs)[(ps)[i_] * q1_];                \
        }                                                                      \
        const int64_t *last_ = (offs) + (ps)[(r) - 1] * q1_;                   \
        for (;;) {                                                             \
            const int64_t b_ = part_[(r) - 1];                                 \
            for (int kk_ = 0; kk_ < q1_; kk_++) {                              \
                int64_t w_ = b_ + last_[kk_];                                  \
                BODY;                                                          \
            }                                                                  \
            int i_ = (r) - 2;                                                  \
            while (i_ >= 0 && k_[i_] == q1_ - 1) { k_[i_] = 0; i_--; }         \
            if (i_ < 0) break;                                                 \
            k_[i_]++;                                                          \
            for (int j_ = i_; j_ < (r) - 1; j_++)                              \
                part_[j_ + 1] = part_[j_] + (offs)[(ps)[j_] * q1_ + k_[j_]];   \
        }                                                                      \
    } while (0)

/* ------------------------------------------------------------- balls */
/* +-1 over the radius-R ball around codeword word `c` at index base (:353-380) */
static int64_t ball_apply(cso_t *s, const uint8_t *c, int64_t base, int delta)
{
    int64_t offs[MAXN * 64];
    make_offs(s, c, offs);
    int64_t dunc = 0;
    for (int r = 0; r <= s->R && r <= s->n; r++) {
        for (int k = 0; k < s->nsub[r]; k++) {
            const uint8_t *ps = s->subs[r] + (size_t)k * (r ? r : 1);
            if (delta > 0) WALK(s, offs, ps, r, base, dunc += cnt_inc(s, w_));
            else           WALK(s, offs, ps, r, base, dunc += cnt_dec(s, w_));
        }
    }
    s->uncovered += dunc;
    return dunc;
}

void cso_ball(cso_t *s, int i, int delta)
{
    ball_apply(s, s->code + (size_t)i * s->n, s->cidx[i], delta);
}

/* the exact uncovered set, rebuilt from cnt in ascending word order */
static void set_from_cnt(cso_t *s)
{
    if (!s->use_set) return;
    s->nu = 0;
    for (int64_t w = 0; w < s->N; w++) {
        s->upos[w] = -1;
        if (s->cnt[w] == 0) set_add(s, w);
    }
}

/* zero the counters and mark the first m codewords' balls (:958-967) */
static void mark_from_scratch(cso_t *s, const uint8_t *cw, int64_t *idx, int m)
{
    const int saved = s->use_set;
    memset(s->cnt, 0, sizeof(uint16_t) * (size_t)s->N);
    s->uncovered = s->N;
    s->use_set = 0;
    for (int i = 0; i < m; i++) {
        idx[i] = widx(s, cw + (size_t)i * s->n);
        ball_apply(s, cw + (size_t)i * s->n, idx[i], +1);
    }
    s->use_set = saved;
    set_from_cnt(s);
}

static void rebuild(cso_t *s)
{
    mark_from_scratch(s, s->code, s->cidx, s->M);
}

int64_t cso_unique_cover(cso_t *s, int i)
{
    const uint8_t *c = s->code + (size_t)i * s->n;
    int64_t offs[MAXN * 64];
    make_offs(s, c, offs);
    int64_t u = 0;
    for (int r = 0; r <= s->R && r <= s->n; r++)
        for (int k = 0; k < s->nsub[r]; k++) {
            const uint8_t *ps = s->subs[r] + (size_t)k * (r ? r : 1);
            WALK(s, offs, ps, r, s->cidx[i], u += (s->cnt[w_] == 1));
        }
    return u;
}

/* ----------------------------------------------------------- create */
cso_t *cso_create(int q, int n, int R, int M, uint64_t seed)
{
    if (q < 2 || q > 64 || n < 1 || n > MAXN - 1 || R < 0 || M < 1 || M > 65535)
        return NULL;
    cso_t *s = calloc(1, sizeof *s);
    s->q = q; s->n = n; s->R = R; s->M = M;
    s->N = 1;
    for (int i = 0; i < n; i++) {
        if (s->N > (int64_t)4e10 / q) { free(s); return NULL; }
        s->N *= q;
    }
    for (int p = 0; p < n; p++) {
        s->pw[p] = 1;
        for (int e = 0; e < n - 1 - p; e++) s->pw[p] *= q;
    }
    s->target = -1;
    s->ttt_it = -1;
    s->cnt = calloc((size_t)s->N, sizeof(uint16_t));
    s->code = calloc((size_t)M * n, 1);
    s->best = calloc((size_t)M * n, 1);
    s->tmpw = calloc((size_t)n, 1);
    s->cidx = calloc((size_t)M, sizeof(int64_t));
    s->tabu = calloc((size_t)M * n * q, sizeof(int32_t));
    rng_seed(s->rs, seed);
    s->threads = 1;

    int all[MAXN];
    for (int j = 0; j < n; j++) all[j] = j;
    for (int r = 0; r <= R && r <= n; r++) s->subs[r] = enum_subsets(all, n, r, &s->nsub[r]);
    s->npsub = (int)binom64(n - 1, R);
    s->psub = malloc((size_t)n * (s->npsub ? s->npsub : 1) * (R ? R : 1));
    for (int p = 0; p < n; p++) {
        int rest[MAXN], nr = 0, cntp;
        for (int j = 0; j < n; j++) if (j != p) rest[nr++] = j;
        uint8_t *t = enum_subsets(rest, nr, R, &cntp);
        memcpy(s->psub + (size_t)p * s->npsub * (R ? R : 1), t,
               (size_t)s->npsub * (R ? R : 1));
        free(t);
    }

    s->use_set = s->N < ((int64_t)1 << 31);
    if (s->use_set) {
        s->upos = malloc(sizeof(int32_t) * (size_t)s->N);
        s->uidx = malloc(sizeof(int64_t) * (size_t)s->N);
        s->udig = malloc((size_t)s->N * n);
        for (int64_t w = 0; w < s->N; w++) s->upos[w] = -1;
        s->nu = 0;
    }
    s->ncap = (int64_t)M * n + 1;
    s->cand_i = malloc(sizeof(int) * s->ncap);
    s->cand_p = malloc(sizeof(int) * s->ncap);
    s->cand_v = malloc(sizeof(int) * s->ncap);
    s->cand_d = malloc(sizeof(int64_t) * s->ncap);
    s->dist = malloc(sizeof(int) * M);
    s->cw_list = malloc(sizeof(int) * M);
    s->cw_mark = calloc((size_t)M, sizeof(int));
    s->dvec = malloc(sizeof(int64_t) * (size_t)M * n);
    s->uncovered = s->N;
    return s;
}

void cso_destroy(cso_t *s)
{
    if (!s) return;
    free(s->cnt); free(s->code); free(s->best); free(s->tmpw); free(s->cidx);
    free(s->tabu);
    for (int r = 0; r <= s->R && r <= s->n; r++) free(s->subs[r]);
    free(s->psub);
    free(s->upos); free(s->uidx); free(s->udig);
    free(s->cand_i); free(s->cand_p); free(s->cand_v); free(s->cand_d);
    free(s->dist); free(s->cw_list); free(s->cw_mark); free(s->dvec);
    free(s->rl_bit); free(s->rl_idx); free(s->dvecw);
    free(s);
}

/* ------------------------------------------------------------- pick */
/* (:973-1016 with --upick/--ucache off) */
int64_t cso_pick(cso_t *s)
{
    const uint64_t N = (uint64_t)s->N;
    for (int t = 0; t < 64; t++) {
        int64_t w = (int64_t)(draw(s) % N);
        if (s->cnt[w] == 0) return w;
    }
    const int64_t start = (int64_t)(draw(s) % N);
    /* first word with cnt == 0 in cyclic order from start; a short direct scan,
     * then the exact set (same answer, less work when uncovered words are rare) */
    int64_t lim = s->N < 4096 ? s->N : 4096;
    for (int64_t d = 0; d < lim; d++) {
        int64_t w = start + d;
        if (w >= s->N) w -= s->N;
        if (s->cnt[w] == 0) return w;
    }
    if (s->use_set) {
        if (s->nu == 0) return -1;
        int64_t bestd = -1, bw = -1;
        for (int64_t e = 0; e < s->nu; e++) {
            int64_t w = s->uidx[e];
            int64_t d = w - start;
            if (d < 0) d += s->N;
            if (bestd < 0 || d < bestd) { bestd = d; bw = w; }
        }
        return bw;
    }
    for (int64_t d = lim; d < s->N; d++) {
        int64_t w = start + d;
        if (w >= s->N) w -= s->N;
        if (s->cnt[w] == 0) return w;
    }
    return -1;
}

/* ------------------------------------------------------------- init */
int cso_init(cso_t *s, const uint8_t *seed_code, int mseed, int greedy)
{
    const int n = s->n, M = s->M;
    if (mseed < 0) return -1;
    for (int64_t k = 0; k < (int64_t)mseed * n; k++)
        if (seed_code[k] >= s->q) return -1;
    int have = 0;
    if (mseed > M) {
        /* remove-and-repair: drop the codeword with the fewest uniquely covered
         * words, first such index on ties (:1156-1196) */
        uint8_t *sc = malloc((size_t)mseed * n);
        int64_t *si = malloc(sizeof(int64_t) * mseed);
        memcpy(sc, seed_code, (size_t)mseed * n);
        mark_from_scratch(s, sc, si, mseed);
        while (mseed > M) {
            int worst = 0;
            int64_t wu = -1;
            for (int i = 0; i < mseed; i++) {
                int64_t offs[MAXN * 64];
                make_offs(s, sc + (size_t)i * n, offs);
                int64_t u2 = 0;
                for (int r = 0; r <= s->R && r <= n; r++)
                    for (int k = 0; k < s->nsub[r]; k++) {
                        const uint8_t *ps = s->subs[r] + (size_t)k * (r ? r : 1);
                        WALK(s, offs, ps, r, si[i], u2 += (s->cnt[w_] == 1));
                    }
                if (wu < 0 || u2 < wu) { wu = u2; worst = i; }
            }
            ball_apply(s, sc + (size_t)worst * n, si[worst], -1);
            mseed--;
            memmove(sc + (size_t)worst * n, sc + (size_t)(worst + 1) * n,
                    (size_t)(mseed - worst) * n);
            memmove(si + worst, si + worst + 1, sizeof(int64_t) * (size_t)(mseed - worst));
        }
        memcpy(s->code, sc, (size_t)M * n);
        free(sc);
        free(si);
        have = M;
    } else if (mseed > 0) {
        memcpy(s->code, seed_code, (size_t)mseed * n);
        have = mseed;
    }
    if (have < M && greedy) {
        /* greedy placement on uncovered words (:1197-1234) */
        mark_from_scratch(s, s->code, s->cidx, have);
        for (int i = have; i < M; i++) {
            int64_t w = cso_pick(s);
            if (w < 0) w = (int64_t)rng_mod(s, (uint64_t)s->N);
            word_of(s, w, s->code + (size_t)i * n);
            s->cidx[i] = w;
            ball_apply(s, s->code + (size_t)i * n, w, +1);
        }
    } else {
        /* uniform random digits, row-major, then a full rebuild (:1235-1239) */
        for (int i = have; i < M; i++)
            for (int j = 0; j < n; j++)
                s->code[(size_t)i * n + j] = (uint8_t)rng_mod(s, (uint64_t)s->q);
        rebuild(s);
    }
    memcpy(s->best, s->code, (size_t)M * n);
    s->best_unc = s->uncovered;
    s->it = 0;
    s->since = 0;
    s->nkicks = 0;
    memset(s->tabu, 0, sizeof(int32_t) * (size_t)M * n * s->q);
    return 0;
}

int cso_set_code(cso_t *s, const uint8_t *code)
{
    for (int64_t k = 0; k < (int64_t)s->M * s->n; k++)
        if (code[k] >= s->q) return -1;
    memcpy(s->code, code, (size_t)s->M * s->n);
    rebuild(s);
    return 0;
}

/* --------------------------------------------------------- evaluation */
int64_t cso_move_delta(cso_t *s, int i, int p, int v)
{
    /* per-move semantics: pos = #{a : cnt[a] == 1}, neg = #{b : cnt[b] == 0}
     * over the S patterns of the subsets avoiding p (:412-558) */
    const uint8_t *c = s->code + (size_t)i * s->n;
    int64_t offs[MAXN * 64];
    make_offs(s, c, offs);
    const int64_t dv = (int64_t)(v - c[p]) * s->pw[p];
    const int64_t base = s->cidx[i];
    const int R = s->R;
    int64_t pos = 0, neg = 0;
    for (int k = 0; k < s->npsub; k++) {
        const uint8_t *ps = s->psub + ((size_t)p * s->npsub + k) * (R ? R : 1);
        WALK(s, offs, ps, R, base, {
            pos += (s->cnt[w_] == 1);
            neg += (s->cnt[w_ + dv] == 0);
        });
    }
    return pos - neg;
}

void cso_eval_codeword(cso_t *s, int i, const uint8_t *u, int64_t *dvec, int use_list)
{
    const int n = s->n, R = s->R;
    const uint8_t *c = s->code + (size_t)i * n;
    const int64_t base = s->cidx[i];
    int64_t offs[MAXN * 64];
    make_offs(s, c, offs);

    /* leaving side: per R-subset of all positions, the words of the full
     * distance-R sphere with cnt == 1; hit[j] sums the subsets holding j */
    int64_t t1 = 0, hit[MAXN] = {0};
    for (int k = 0; k < s->nsub[R]; k++) {
        const uint8_t *ps = s->subs[R] + (size_t)k * (R ? R : 1);
        int64_t c1 = 0;
        WALK(s, offs, ps, R, base, c1 += (s->cnt[w_] == 1));
        t1 += c1;
        for (int t = 0; t < R; t++) hit[ps[t]] += c1;
    }

    /* entering side: words x with cnt == 0, d(x,c) == R+1, x_p == u_p */
    int64_t neg[MAXN] = {0};
    if (use_list && s->use_set) {
        for (int64_t e = 0; e < s->nu; e++) {
            const uint8_t *xd = s->udig + (size_t)e * n;
            int d = 0;
            for (int j = 0; j < n && d <= R + 1; j++) d += (xd[j] != c[j]);
            if (d != R + 1) continue;
            for (int j = 0; j < n; j++)
                if (xd[j] != c[j] && xd[j] == u[j]) neg[j]++;
        }
    } else {
        for (int p = 0; p < n; p++) {
            if (c[p] == u[p]) continue;
            const int64_t dv = (int64_t)(u[p] - c[p]) * s->pw[p];
            int64_t nn = 0;
            for (int k = 0; k < s->npsub; k++) {
                const uint8_t *ps = s->psub + ((size_t)p * s->npsub + k) * (R ? R : 1);
                WALK(s, offs, ps, R, base + dv, nn += (s->cnt[w_] == 0));
            }
            neg[p] = nn;
        }
    }
    for (int p = 0; p < n; p++)
        if (c[p] != u[p]) dvec[p] = (t1 - hit[p]) - neg[p];
}

/* Every move (p, v != c_p) of codeword i from one evaluation (:696-760):
 * dvec[p*q + v] = (T - H_p) - neg[p][v], the entering counts neg[j][x_j]
 * over the uncovered words x at distance R+1 (j running over the
 * positions where x and c differ), or walked per move without the set. */
void cso_eval_codeword_wide(cso_t *s, int i, int64_t *dvec)
{
    const int n = s->n, R = s->R, q = s->q;
    const uint8_t *c = s->code + (size_t)i * n;
    const int64_t base = s->cidx[i];
    int64_t offs[MAXN * 64];
    make_offs(s, c, offs);
    int64_t t1 = 0, hit[MAXN] = {0};
    for (int k = 0; k < s->nsub[R]; k++) {
        const uint8_t *ps = s->subs[R] + (size_t)k * (R ? R : 1);
        int64_t c1 = 0;
        WALK(s, offs, ps, R, base, c1 += (s->cnt[w_] == 1));
        t1 += c1;
        for (int t = 0; t < R; t++) hit[ps[t]] += c1;
    }
    int64_t neg[MAXN * 64];
    memset(neg, 0, sizeof(int64_t) * (size_t)n * q);
    if (s->use_set) {
        for (int64_t e = 0; e < s->nu; e++) {
            const uint8_t *xd = s->udig + (size_t)e * n;
            int d = 0;
            for (int j = 0; j < n && d <= R + 1; j++) d += (xd[j] != c[j]);
            if (d != R + 1) continue;
            for (int j = 0; j < n; j++)
                if (xd[j] != c[j]) neg[j * q + xd[j]]++;
        }
    } else {
        for (int p = 0; p < n; p++)
            for (int v = 0; v < q; v++) {
                if (v == c[p]) continue;
                const int64_t dv = (int64_t)(v - c[p]) * s->pw[p];
                int64_t nn = 0;
                for (int k = 0; k < s->npsub; k++) {
                    const uint8_t *ps = s->psub + ((size_t)p * s->npsub + k) * (R ? R : 1);
                    WALK(s, offs, ps, R, base + dv, nn += (s->cnt[w_] == 0));
                }
                neg[p * q + v] = nn;
            }
    }
    for (int p = 0; p < n; p++)
        for (int v = 0; v < q; v++)
            dvec[p * q + v] = v == c[p] ? CSO_DSEL_RANDOM : (t1 - hit[p]) - neg[p * q + v];
}

int64_t cso_commit(cso_t *s, int i, int p, int v)
{
    /* --cnt[a], ++cnt[b] with b = a + dv over the S patterns; a and b range
     * over disjoint sets, every word once (:562-590) */
    uint8_t *c = s->code + (size_t)i * s->n;
    int64_t offs[MAXN * 64];
    make_offs(s, c, offs);
    const int64_t dv = (int64_t)(v - c[p]) * s->pw[p];
    const int64_t base = s->cidx[i];
    const int R = s->R;
    int64_t dunc = 0;
    for (int k = 0; k < s->npsub; k++) {
        const uint8_t *ps = s->psub + ((size_t)p * s->npsub + k) * (R ? R : 1);
        WALK(s, offs, ps, R, base, {
            dunc += cnt_dec(s, w_);
            dunc += cnt_inc(s, w_ + dv);
        });
    }
    s->cidx[i] += dv;
    c[p] = (uint8_t)v;
    s->uncovered += dunc;
    return dunc;
}

/* -------------------------------------------------------------- run */
int64_t cso_run(cso_t *s, int64_t iters, int tabu_len, int64_t stall,
                int kick, int cand, int64_t *trace)
{
    const int n = s->n, q = s->q, M = s->M, R = s->R;
    uint8_t u[MAXN];
    int64_t done = 0;
    int64_t P = 1;
    for (int r = 0; r < R; r++) P *= (q - 1);
    const int64_t TOTPAT = (int64_t)s->npsub * P;   /* S, the per-move sphere */

    while (s->uncovered > 0) {
        /* the top of the loop (:1322-1324): iteration budget, then target */
        if (iters >= 0 && done >= iters) break;
        if (s->target >= 0 && s->best_unc <= s->target) break;
        s->it++;
        done++;
        const int64_t it = s->it;
        if (s->wide) {
            if (!s->rl_valid && s->uncovered <= RL_CAP / 2) rl_rebuild(s);
            else if (s->rl_valid && s->rl_n > 2 * s->uncovered + 64) rl_compact(s);
        }

        int64_t uw = cso_pick(s);
        if (uw < 0) break;
        word_of(s, uw, u);

        /* distances and candidates (:1339-1376) */
        int dmin = n + 1;
        for (int i = 0; i < M; i++) {
            const uint8_t *c = s->code + (size_t)i * n;
            int d = 0;
            for (int j = 0; j < n; j++) d += (c[j] != u[j]);
            s->dist[i] = d;
            if (d < dmin) dmin = d;
        }
        const int dtar = (dmin <= R + 1) ? R + 1 : dmin;
        /* --wide takes every move of the candidate codewords while the
         * reference's list is valid and short (:1353-1357) */
        const int wide_now = s->wide && s->rl_valid && s->rl_n * n < (int64_t)(R + 1) * TOTPAT;
        int ncand = 0;
        for (int i = 0; i < M; i++) {
            if (s->dist[i] != dtar) continue;
            const uint8_t *c = s->code + (size_t)i * n;
            for (int p = 0; p < n; p++) {
                if (wide_now) {
                    for (int v = 0; v < q; v++) {
                        if (v == c[p]) continue;
                        s->cand_i[ncand] = i;
                        s->cand_p[ncand] = p;
                        s->cand_v[ncand] = v;
                        ncand++;
                    }
                    continue;
                }
                if (c[p] == u[p]) continue;
                s->cand_i[ncand] = i;
                s->cand_p[ncand] = p;
                s->cand_v[ncand] = u[p];
                ncand++;
            }
        }
        if (ncand == 0) {
            int i = (int)rng_mod(s, (uint64_t)M);
            int p = (int)rng_mod(s, (uint64_t)n);
            s->cand_i[0] = i; s->cand_p[0] = p; s->cand_v[0] = u[p];
            ncand = 1;
        }
        /* partial Fisher-Yates over the first `take` slots (:1383-1389) */
        const int take = ncand < cand ? ncand : cand;
        for (int t = 0; t < take; t++) {
            int j = t + (int)rng_mod(s, (uint64_t)(ncand - t));
            int ti = s->cand_i[t], tp = s->cand_p[t], tv = s->cand_v[t];
            s->cand_i[t] = s->cand_i[j]; s->cand_p[t] = s->cand_p[j]; s->cand_v[t] = s->cand_v[j];
            s->cand_i[j] = ti; s->cand_p[j] = tp; s->cand_v[j] = tv;
        }

        /* shared-sphere evaluation per distinct codeword */
        int ncw = 0;
        for (int t = 0; t < take; t++) {
            int i = s->cand_i[t];
            if (s->cw_mark[i] != (int)(it & 0x3fffffff) + 1) {
                s->cw_mark[i] = (int)(it & 0x3fffffff) + 1;
                s->cw_list[ncw++] = i;
            }
        }
        const int use_list = s->use_set && (s->nu * n < (int64_t)(R + 1) * TOTPAT);
        if (wide_now) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(s->threads) if (ncw > 1 && s->threads > 1)
            for (int z = 0; z < ncw; z++) {
                int i = s->cw_list[z];
                cso_eval_codeword_wide(s, i, s->dvecw + (size_t)i * n * q);
            }
            for (int t = 0; t < take; t++)
                s->cand_d[t] = s->dvecw[((size_t)s->cand_i[t] * n + s->cand_p[t]) * q + s->cand_v[t]];
        } else {
#pragma omp parallel for schedule(dynamic, 1) num_threads(s->threads) if (ncw > 1 && s->threads > 1)
            for (int z = 0; z < ncw; z++) {
                int i = s->cw_list[z];
                cso_eval_codeword(s, i, u, s->dvec + (size_t)i * n, use_list);
            }
            for (int t = 0; t < take; t++)
                s->cand_d[t] = s->dvec[(size_t)s->cand_i[t] * n + s->cand_p[t]];
        }

        /* selection: tabu with aspiration, reservoir tie-break (:1503-1518) */
        int64_t bestd = CSO_DSEL_RANDOM;
        int nties = 0, bi = -1, bp = -1, bv = -1;
        for (int t = 0; t < take; t++) {
            const int i = s->cand_i[t], p = s->cand_p[t], v = s->cand_v[t];
            const int64_t d = s->cand_d[t];
            const int is_tabu = s->tabu[((size_t)i * n + p) * q + v] > it;
            if (is_tabu && s->uncovered + d >= s->best_unc) continue;
            if (d < bestd) { bestd = d; bi = i; bp = p; bv = v; nties = 1; }
            else if (d == bestd) {
                nties++;
                if (rng_mod(s, (uint64_t)nties) == 0) { bi = i; bp = p; bv = v; }
            }
        }
        int64_t dsel = bestd;
        if (bi < 0) {
            int t = (int)rng_mod(s, (uint64_t)take);
            bi = s->cand_i[t]; bp = s->cand_p[t]; bv = s->cand_v[t];
            dsel = CSO_DSEL_RANDOM;
            nties = 0;
        }

        const int oldv = s->code[(size_t)bi * n + bp];
        const int64_t dunc = cso_commit(s, bi, bp, bv);
        s->tabu[((size_t)bi * n + bp) * q + oldv] =
            (int32_t)(it + tabu_len + (int64_t)rng_mod(s, 4));

        int kicked = 0;
        int stop = 0;
        if (s->uncovered < s->best_unc) {
            s->best_unc = s->uncovered;
            memcpy(s->best, s->code, (size_t)M * n);
            s->since = 0;
            /* time-to-target record (:1536-1541) */
            if (s->target >= 0 && s->ttt_it < 0 && s->best_unc <= s->target) s->ttt_it = it;
            if (s->best_unc == 0) stop = 1;
        } else if (++s->since > stall) {
            /* kick: teleport `kick` random codewords near uncovered words (:1553-1578) */
            for (int t = 0; t < kick; t++) {
                int i = (int)rng_mod(s, (uint64_t)M);
                uint8_t *ci = s->code + (size_t)i * n;
                ball_apply(s, ci, s->cidx[i], -1);
                int64_t w = cso_pick(s);
                if (w >= 0) {
                    word_of(s, w, s->tmpw);
                    memcpy(ci, s->tmpw, n);
                    for (int t2 = 0; t2 < R; t2++) {
                        int pos, val;
                        if (s->kick_order == 0) {
                            val = (int)rng_mod(s, (uint64_t)q);
                            pos = (int)rng_mod(s, (uint64_t)n);
                        } else {
                            pos = (int)rng_mod(s, (uint64_t)n);
                            val = (int)rng_mod(s, (uint64_t)q);
                        }
                        ci[pos] = (uint8_t)val;
                    }
                } else {
                    for (int j = 0; j < n; j++) ci[j] = (uint8_t)rng_mod(s, (uint64_t)q);
                }
                s->cidx[i] = widx(s, ci);
                ball_apply(s, ci, s->cidx[i], +1);
            }
            s->nkicks++;
            memset(s->tabu, 0, sizeof(int32_t) * (size_t)M * n * q);
            s->since = 0;
            kicked = 1;
        }
        if (trace) {
            int64_t *tr = trace + (done - 1) * CSO_TR_W;
            tr[CSO_TR_IT] = it;
            tr[CSO_TR_U] = uw;
            tr[CSO_TR_NCAND] = ncand;
            tr[CSO_TR_NCW] = ncw;
            tr[CSO_TR_BI] = bi;
            tr[CSO_TR_BP] = bp;
            tr[CSO_TR_BV] = bv;
            tr[CSO_TR_OLDV] = oldv;
            tr[CSO_TR_DSEL] = dsel;
            tr[CSO_TR_DUNC] = dunc;
            tr[CSO_TR_UNC] = s->uncovered;
            tr[CSO_TR_BEST] = s->best_unc;
            tr[CSO_TR_KICK] = kicked;
            tr[CSO_TR_DRAWS] = s->ndraw;
            tr[CSO_TR_SINCE] = s->since;
            tr[CSO_TR_NTIES] = nties;
        }
        if (stop) break;
    }
    return done;
}

/* -------------------------------------------------------- accessors */
void cso_set_kick_order(cso_t *s, int order) { s->kick_order = order; }
/* --wide (:1068): switch the wide neighbourhood on before cso_init */
int cso_set_wide(cso_t *s, int on)
{
    if (!on) { s->wide = 0; return 0; }
    if (!s->rl_bit) {
        s->rl_bit = calloc((size_t)((s->N + 63) / 64), 8);
        s->rl_idx = malloc(sizeof(int64_t) * (size_t)RL_CAP);
        s->dvecw = malloc(sizeof(int64_t) * (size_t)s->M * s->n * s->q);
        const int64_t cap = (int64_t)s->M * s->n * s->q + 1;
        s->cand_i = realloc(s->cand_i, sizeof(int) * cap);
        s->cand_p = realloc(s->cand_p, sizeof(int) * cap);
        s->cand_v = realloc(s->cand_v, sizeof(int) * cap);
        s->cand_d = realloc(s->cand_d, sizeof(int64_t) * cap);
        s->ncap = cap;
        if (!s->rl_bit || !s->rl_idx || !s->dvecw || !s->cand_i || !s->cand_p || !s->cand_v || !s->cand_d)
            return -1;
    }
    s->wide = 1;
    s->rl_valid = 0;
    s->rl_n = 0;
    return 0;
}

/* --target (:1039,1078): a target the best code already meets is met at the
 * current iteration (:1291-1293) */
void cso_set_target(cso_t *s, int64_t target)
{
    s->target = target < 0 ? -1 : target;
    s->ttt_it = -1;
    if (s->target >= 0 && s->best_unc <= s->target) s->ttt_it = s->it;
}
int64_t cso_ttt_it(const cso_t *s) { return s->ttt_it; }
void cso_set_threads(cso_t *s, int threads) { s->threads = threads > 0 ? threads : 1; }
int64_t cso_uncovered(const cso_t *s) { return s->uncovered; }
int64_t cso_best_uncovered(const cso_t *s) { return s->best_unc; }
int64_t cso_iters(const cso_t *s) { return s->it; }
int64_t cso_kicks(const cso_t *s) { return s->nkicks; }
int64_t cso_space(const cso_t *s) { return s->N; }
const uint8_t *cso_code(const cso_t *s) { return s->code; }
const uint8_t *cso_best(const cso_t *s) { return s->best; }
const uint16_t *cso_cnt(const cso_t *s) { return s->cnt; }
void cso_rng_state(const cso_t *s, uint64_t out[4]) { memcpy(out, s->rs, sizeof s->rs); }
void cso_set_rng_state(cso_t *s, const uint64_t in[4]) { memcpy(s->rs, in, sizeof s->rs); }
int64_t cso_draws(const cso_t *s) { return s->ndraw; }

typedef struct { int64_t x; int pos; } xp_t;

static int cmp_xp(const void *a, const void *b)
{
    const xp_t *u = a, *v = b;
    if (u->x != v->x) return (u->x > v->x) - (u->x < v->x);
    return (u->pos > v->pos) - (u->pos < v->pos);
}

/* distinctness repair of the published code (:791-853): keep first
 * occurrences in order, then top up with the smallest unused word indices */
void cso_distinct_fill(const cso_t *s, const uint8_t *cw, int nw, uint8_t *out)
{
    const int n = s->n, M = s->M;
    if (nw > M) nw = M;
    xp_t *a = malloc(sizeof(xp_t) * (size_t)(nw + 1));
    char *keep = calloc((size_t)nw + 1, 1);
    for (int i = 0; i < nw; i++) { a[i].x = widx(s, cw + (size_t)i * n); a[i].pos = i; }
    qsort(a, nw, sizeof(xp_t), cmp_xp);
    int nd = 0;
    for (int i = 0; i < nw; i++)
        if (i == 0 || a[i].x != a[i - 1].x) { keep[a[i].pos] = 1; a[nd++].x = a[i].x; }
    int m = 0;
    for (int i = 0; i < nw; i++)
        if (keep[i]) { memcpy(out + (size_t)m * n, cw + (size_t)i * n, n); m++; }
    int k = 0;
    for (int64_t x = 0; m < M && x < s->N; x++) {
        while (k < nd && a[k].x < x) k++;
        if (k < nd && a[k].x == x) continue;
        word_of(s, x, out + (size_t)m * n);
        m++;
    }
    free(a);
    free(keep);
}
