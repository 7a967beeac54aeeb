/*
 * covsearch_oracle.h -- CPU oracle for the covering-code tabu search.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the search that
 * cov/engine/covsearch2e.c runs, used to check the HIP engine bit for bit.  It
 * is never the product: only tests/, the smoke check and the CPU-baseline leg
 * of bench.py may load it.
 *
 * A run is a pure function of (q, n, R, M, seed, init, tabu, stall, kick,
 * cand, iters).  The per-step trace layout (CSO_TR_*) is shared with the HIP
 * engine so the two can be compared record by record.
 */
#ifndef COVSEARCH_ORACLE_H
#define COVSEARCH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cso cso_t;

/* one int64 record per search step */
enum {
    CSO_TR_IT = 0,    /* iteration number, 1-based                            */
    CSO_TR_U,         /* uncovered word picked at the start of the step       */
    CSO_TR_NCAND,     /* candidate moves collected                            */
    CSO_TR_NCW,       /* distinct candidate codewords evaluated               */
    CSO_TR_BI,        /* chosen codeword                                      */
    CSO_TR_BP,        /* chosen position                                      */
    CSO_TR_BV,        /* new value                                            */
    CSO_TR_OLDV,      /* old value                                            */
    CSO_TR_DSEL,      /* evaluated delta of the chosen move (1<<60 if random) */
    CSO_TR_DUNC,      /* delta applied by the commit                          */
    CSO_TR_UNC,       /* uncovered after the step (after a kick if any)       */
    CSO_TR_BEST,      /* best uncovered after the step                        */
    CSO_TR_KICK,      /* 1 if the step ended with a kick                      */
    CSO_TR_DRAWS,     /* rng draws consumed so far, after the step            */
    CSO_TR_SINCE,     /* steps since the last improvement, after the step     */
    CSO_TR_NTIES,     /* candidates tied at the chosen delta                  */
    CSO_TR_W          /* record width                                         */
};

#define CSO_DSEL_RANDOM (1LL << 60)

cso_t  *cso_create(int q, int n, int R, int M, uint64_t seed);
void    cso_destroy(cso_t *s);

/* Initial code.  seed_code holds mseed codewords (row-major, n digits each),
 * mseed may be 0.  mseed > M descends by remove-and-repair; otherwise the
 * remaining M - mseed codewords are placed greedily (greedy=1) or uniformly at
 * random (greedy=0).  Returns 0, or -1 on bad input. */
int     cso_init(cso_t *s, const uint8_t *seed_code, int mseed, int greedy);

/* Run up to `iters` further steps (iters < 0: until covered).  `trace`, if not
 * NULL, receives CSO_TR_W int64 per step.  Returns the number of steps run. */
int64_t cso_run(cso_t *s, int64_t iters, int tabu_len, int64_t stall,
                int kick, int cand, int64_t *trace);

/* order of the two draws in a kick perturbation: 0 = value then position,
 * 1 = position then value */
void    cso_set_kick_order(cso_t *s, int order);
/* --target T: runs stop before a step once the best uncovered count is <= T
 * (T < 0: no target); cso_ttt_it gives the iteration whose move first
 * reached it (-1: not yet; the current iteration if already met when set). */
void    cso_set_target(cso_t *s, int64_t target);
/* --wide: every move of a candidate codeword is a candidate while the
 * reference's uncovered-word list is valid and short; set before cso_init.
 * Returns 0, or -1 when its buffers cannot be allocated. */
int     cso_set_wide(cso_t *s, int on);
/* every move's delta of codeword i: dvec[p*q + v] (v == c_p: CSO_DSEL_RANDOM) */
void    cso_eval_codeword_wide(cso_t *s, int i, int64_t *dvec);
int64_t cso_ttt_it(const cso_t *s);
void    cso_set_threads(cso_t *s, int threads);

int64_t cso_uncovered(const cso_t *s);
int64_t cso_best_uncovered(const cso_t *s);
int64_t cso_iters(const cso_t *s);
int64_t cso_kicks(const cso_t *s);
int64_t cso_space(const cso_t *s);
const uint8_t  *cso_code(const cso_t *s);
const uint8_t  *cso_best(const cso_t *s);
const uint16_t *cso_cnt(const cso_t *s);
void    cso_rng_state(const cso_t *s, uint64_t out[4]);
void    cso_set_rng_state(cso_t *s, const uint64_t in[4]);
int64_t cso_draws(const cso_t *s);

/* Overwrite the current code and rebuild the counters (no rng draws). */
int     cso_set_code(cso_t *s, const uint8_t *code);

/* Delta of one move "codeword i, position p := v" by the per-move sphere
 * walk (uncovered after minus uncovered before), without committing. */
int64_t cso_move_delta(cso_t *s, int i, int p, int v);

/* Shared-sphere evaluation of codeword i towards word u: dvec[p] receives the
 * delta of moving c_p to u_p for every p with c_p != u_p (others untouched).
 * use_list selects the entering-side method (1 = uncovered set, 0 = walks). */
void    cso_eval_codeword(cso_t *s, int i, const uint8_t *u, int64_t *dvec,
                          int use_list);

/* Commit the move; returns the change in the uncovered count. */
int64_t cso_commit(cso_t *s, int i, int p, int v);

/* Add (+1) or remove (-1) the radius-R ball of codeword i's current word. */
void    cso_ball(cso_t *s, int i, int delta);

/* The reference's pick: up to 64 random probes, then a cyclic scan. */
int64_t cso_pick(cso_t *s);

/* Uncovered words of codeword i's ball that nothing else covers. */
int64_t cso_unique_cover(cso_t *s, int i);

/* Write exactly M distinct codewords (duplicates replaced by the smallest
 * unused words) to `out` (M*n digits), from `cw` (nw codewords). */
void    cso_distinct_fill(const cso_t *s, const uint8_t *cw, int nw,
                          uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
