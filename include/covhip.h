/*
 * covhip.h -- C ABI of libcovhip.so, the MI355X (gfx950) engine for the
 * covering-code search hot path: coverage counts over Z_q^n, the shared-sphere
 * tabu move evaluator, the uint8 axis-DP distance transform and the
 * uncovered-set reductions, plus the device-resident search loop built on them.
 *
 * Conventions (those of the reference's ctypes engines, cov/gemm2/dct.cu:299-504
 * and cov/gemm2/chainsolve.cu:370-433):
 *   - every call returns int (or int64 for counts); 0 / a count = success,
 *     negative = error: COV_E* below, or -(100 + hipError_t) from the runtime;
 *   - host buffers are owned by the caller and passed as plain pointers;
 *     device buffers are owned by the cell handle;
 *   - a cell owns one HIP stream; calls on one cell are not thread-safe, calls
 *     on distinct cells may run concurrently (the serving loop relies on it);
 *   - words are indexed MSB-first in base q, w = sum_j d_j q^(n-1-j), as in
 *     cov/engine/covsearch2e.c:916-921 and cov/verify_cov.py:114;
 *   - no exception crosses the ABI; cov_last_error() gives the text.
 * Limits of this build: 2 <= q <= 16, 1 <= n <= 16, 0 <= R <= 8, R < n,
 * 1 <= M <= 65535 (uint16 counters, as cov/engine/covsearch2e.c:1117),
 * q^n <= 4e10 (as :1111-1114).
 */
#ifndef COVHIP_H
#define COVHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COV_OK       0
#define COV_EPARAM  -1   /* bad parameters (as dct_init's -1)               */
#define COV_ESPACE  -2   /* q^n too large (as dct_init's -2)                */
#define COV_ENOMEM  -3   /* device allocation failed                        */
#define COV_ESTATE  -4   /* call not valid in the cell's current state      */
#define COV_ELIMIT  -5   /* input beyond a stated capacity                  */

/* per-step trace record: COV_TR_W int64 per step (same layout as the oracle) */
#define COV_TR_W 16
#define COV_DSEL_RANDOM (1LL << 60)

/* search-loop modes of cov_search_run */
#define COV_MODE_PLAIN 0   /* one launch sequence per step                 */
#define COV_MODE_GRAPH 1   /* steps replayed from a captured hipGraph      */
/* steps per captured graph: graph mode replays whole graphs and runs the
 * remainder of a launch (nsteps mod this) as plain launches */
#define COV_GRAPH_STEPS 40

typedef struct cov_cell cov_cell_t;
typedef struct cov_batch cov_batch_t;

typedef struct {
    int64_t it;               /* iterations done                              */
    int64_t uncovered;        /* current uncovered count                      */
    int64_t best_uncovered;   /* best seen (the RESULT line's uncovered=)     */
    int64_t kicks;
    int64_t since_improve;
    int64_t draws;            /* rng draws consumed so far                    */
    int64_t list_n;           /* uncovered-list length (incl. stale entries)  */
    int32_t done;             /* 1 once the search stopped covered (or on a
                                 fault), 2 once it stopped at its target      */
    int32_t list_valid;
    int32_t fault;            /* 0, or why the search stopped on an error:
                                 1 no candidate move (inconsistent counts),
                                 2 candidate sample beyond the rng ring       */
    int32_t reserved;
} cov_state_t;

/* ------------------------------------------------------------- lifecycle */
int         cov_version(void);
const char *cov_strerror(int code);
/* Device bytes a cell of this shape needs; the serving loop admits by it. */
int64_t     cov_mem_bytes(int q, int n, int R, int M);
/* Replaces dct_init (cov/gemm2/dct.cu:302) with a handle instead of globals. */
int         cov_create(int q, int n, int R, int M, cov_cell_t **out);
/* Replaces dct_free (cov/gemm2/dct.cu:338). */
int         cov_destroy(cov_cell_t *h);
const char *cov_last_error(const cov_cell_t *h);
int         cov_sync(cov_cell_t *h);

/* ------------------------------------------------------ code and coverage */
/* Upload M*n digits (row-major) and rebuild cnt by the recount kernel.
 * Replaces dct_set_code + dct_transform(0,2) (cov/gemm2/dct.cu:351-411) and
 * rebuild() (cov/engine/covsearch2e.c:958-967). */
int     cov_set_code(cov_cell_t *h, const uint8_t *digits);
/* The same with flags: COV_NO_RECOUNT uploads the code only (for a distance
 * transform alone); calls that need the counts then return COV_ESTATE until
 * cov_recount. */
#define COV_NO_RECOUNT 1
int     cov_set_code_flags(cov_cell_t *h, const uint8_t *digits, int flags);
int     cov_get_code(cov_cell_t *h, uint8_t *digits);
int     cov_get_best(cov_cell_t *h, uint8_t *digits, int64_t *best_uncovered);
/* (b) recount: cnt from the current code, LDS-aggregated per word tile. */
int     cov_recount(cov_cell_t *h);
/* (b) +-1 over the radius-R balls of k words given by index.
 * Replaces dct_ball_update (cov/gemm2/dct.cu:413) and ball_apply
 * (cov/engine/covsearch2e.c:370-380). */
int     cov_ball_update(cov_cell_t *h, const int64_t *centers, int k, int delta);
/* (d) #{w : cnt[w] == 0}.  Replaces dct_count_eq(0) (cov/gemm2/dct.cu:440). */
int64_t cov_count_uncovered(cov_cell_t *h);
/* (d) the first min(cap, total) uncovered word indices in ascending order;
 * returns the total.  Replaces gather_uncovered_kernel (cov/search/covsearch.cu:213)
 * with a deterministic order, and verify_pure's witnesses (cov/verify_cov.py:142-151). */
int64_t cov_gather_uncovered(cov_cell_t *h, int64_t *out, int64_t cap);
/* (d) cov_gather_uncovered without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  The block counts, their total and
 * the output staging live in one static device array per GPU (640 KB, held
 * until the process ends) instead of per-call allocations.  Limits: q^n <=
 * 2^26 words and cap <= 65536, else COV_ELIMIT.  Every call on one GPU uses
 * that GPU's array and leaves its contents behind (each call rewrites what
 * it reads); calls that overlap on one GPU, from any cells or threads, race
 * and must be serialised by the caller. */
int64_t cov_gather_uncovered_noalloc(cov_cell_t *h, int64_t *out, int64_t cap);
/* (d) first uncovered word at or after start, cyclically; -1 if none.
 * The scan of pick_uncovered (cov/engine/covsearch2e.c:1003-1015). */
int64_t cov_first_uncovered_from(cov_cell_t *h, int64_t start);
/* Replaces dct_read_cnt (cov/gemm2/dct.cu:498). */
int     cov_read_cnt(cov_cell_t *h, int64_t off, int64_t len, uint16_t *out);

/* ----------------------------------------------------------- evaluator */
/* (a) Shared-sphere evaluation: for each of the k codewords cw[t], the delta in
 * the uncovered count of moving c_p to u_p, for every p with c_p != u_p, into
 * dvec[t*n + p]; entries with c_p == u_p are set to COV_DSEL_RANDOM.
 * Replaces eval_codeword (cov/engine/covsearch2e.c:696-760). */
int     cov_eval_moves(cov_cell_t *h, const uint8_t *u, int k, const int32_t *cw,
                       int64_t *dvec);
/* (a) Every single-coordinate move of each of the k codewords cw[t]:
 * dvec[(t*n + p)*q + v] = the change in the uncovered count of c_p := v, from
 * one enumeration of the codeword's distance-R sphere (leaving side T - H_p)
 * and the entering sphere of each move; COV_DSEL_RANDOM where v == c_p.  The
 * n x q vector of eval_codeword (cov/engine/covsearch2e.c:696-760), which the
 * --wide neighbourhood reads (:1353-1376). */
int     cov_eval_codewords(cov_cell_t *h, int k, const int32_t *cw, int64_t *dvec);
/* Commit the move c_i[p] := v; *dunc receives the change in uncovered.
 * Replaces move_commit (cov/engine/covsearch2e.c:562-590). */
int     cov_commit(cov_cell_t *h, int i, int p, int v, int64_t *dunc);

/* --------------------------------------------------- distance transform */
/* (c) dist[w] = min Hamming distance from w to the code (uint8), by n axis
 * passes of D <- min(D, 1 + min over the fiber).  dist may be NULL (then it
 * stays on the device); hist[d] for d = 0..n (n+1 entries) and the covering
 * radius are returned.  Replaces run_mindist / run_dilate
 * (cov/verify_independent.py:128-185) and covering_radius (cov/verify_cov.py:237). */
int     cov_distance_transform(cov_cell_t *h, uint8_t *dist, int64_t *hist, int *radius);

/* (c) the same transform without the histogram: #{w : dist[w] <= r} and the
 * covering radius (max dist) only; dist stays on the device.  The answer of
 * verify_numpy (uncovered = q^n - within at r = R) and covering_radius. */
int     cov_distance_stats(cov_cell_t *h, int r, int64_t *within, int *radius);
/* (d) after cov_distance_transform (or _stats) of the current code: #{w : dist[w] > r},
 * and the first min(cap, total) such words in ascending order; COV_ESTATE if
 * the code changed since the transform.  With r = R these are the uncovered
 * words and verify_numpy's witnesses (cov/verify_cov.py:169-230). */
int64_t cov_gather_far(cov_cell_t *h, int r, int64_t *out, int64_t cap);

/* --------------------------------------------- count transform and maps */
/* (a22) The distance-count transform: for every word x and d = 0..R,
 * N_d(x) = #{u in S : d_H(x, u) = d}, by n axis steps
 * A_d <- A_d + T_{d-1} - A_{d-1} on (R+1) int32 layers, several axes per
 * pass in LDS tiles.  Replaces dct_transform (cov/gemm2/dct.cu:360-411) with
 * its three sources: COV_CT_CODE the cell's code with multiplicities,
 * COV_CT_UNCOVERED {x : cnt[x] == 0}, COV_CT_PRIVATE {x : cnt[x] == 1}.
 * layers (host, (R+1) x q^n int32, layer-major) receives every N_d; else the
 * map sum_d N_d stays on the device for the map calls below (the gain map
 * for COV_CT_UNCOVERED, the loss map for COV_CT_PRIVATE) and is also copied
 * to map (host, q^n int32) when that is not NULL. */
#define COV_CT_CODE      0
#define COV_CT_UNCOVERED 1
#define COV_CT_PRIVATE   2
int     cov_count_transform(cov_cell_t *h, int source, int32_t *layers, int32_t *map);
/* LDS bytes per count-transform tile (0 = the default, 65536). */
int     cov_set_count_tile(cov_cell_t *h, int lds_bytes);
/* The map of the last sum transform: its maximum (dct_map_max, dct.cu:450);
 * the histogram of its positive values in nbins bins, value v in bin
 * floor((v-1) nbins / vmax) (dct_map_hist, :459); the first min(cap, total)
 * words with a value >= thr, in ascending index order, and their values,
 * returning the total (dct_map_extract, :469; the reference's order is that
 * of its atomics); the values at given words (dct_map_read_at, :484). */
int     cov_map_max(cov_cell_t *h, int32_t *max);
int     cov_map_hist(cov_cell_t *h, int nbins, int32_t vmax, int64_t *hist);
int64_t cov_map_extract(cov_cell_t *h, int32_t thr, int64_t *idx, int32_t *val, int64_t cap);
int     cov_map_read(cov_cell_t *h, const int64_t *idx, int k, int32_t *out);
/* Per centre word, #{x in its radius-R ball : cnt[x] == target}: the exact
 * gain (target 0) or loss (target 1) of a codeword there.  Replaces
 * dct_ball_gather (cov/gemm2/dct.cu:427). */
int     cov_ball_gather(cov_cell_t *h, const int64_t *centers, int k, int target, int32_t *out);
/* Coverage counts of a word list of any length (not the cell's code rows):
 * zero, then +1 over the radius-R ball of each of the m words.  The LNS
 * engine's recount (cov/gemm2/dctlib.py:108-113 with dct_set_code). */
int     cov_load_words(cov_cell_t *h, const int64_t *words, int m);

/* --------------------------------------------------------------- search */
/* Seed the rng (splitmix64 -> xoshiro256++, cov/engine/covsearch2e.c:211-219)
 * and build the initial code: seed_code (mseed words, may be NULL/0), then
 * remove-and-repair if mseed > M, else greedy (greedy=1) or uniform random
 * digits for the rest (:1156-1239).  The descent runs on the device with one
 * wait at its end; it keeps 16 bytes per seed word of device memory, allocated
 * by the first call with that many words and held until cov_destroy
 * (COV_ELIMIT for a descent when the ball has 2^31 words or more). */
int     cov_search_init(cov_cell_t *h, uint64_t seed, const uint8_t *seed_code,
                        int mseed, int greedy);
/* Search knobs: --tabu, --stall, --kick, --cand (:1030-1066). */
int     cov_search_params(cov_cell_t *h, int tenure, int64_t stall, int kick, int cand);
/* --wide (:1068, :1353-1376): while the uncovered-word list is valid and
 * short, every move (p, v != c_p) of a candidate codeword is a candidate,
 * not only the moves towards u; the uncovered list then follows the
 * reference's own rules (rebuilt below 2^19 uncovered words, compacted when
 * it holds twice the uncovered count plus 64, never dropped), since its
 * state decides when the wide neighbourhood applies.  Set before
 * cov_search_init and before the cell joins a batch (allocates M KB). */
int     cov_search_wide(cov_cell_t *h, int on);
/* Run up to nsteps more tabu steps (:1304-1580) on the device.  trace, if not
 * NULL, receives COV_TR_W int64 per step actually run.  Returns the number of
 * steps run (fewer when the code covers). */
int64_t cov_search_run(cov_cell_t *h, int64_t nsteps, int mode, int64_t *trace);
/* Non-blocking form: enqueue the steps and return; cov_sync() waits. */
int     cov_search_launch(cov_cell_t *h, int64_t nsteps, int mode);
int     cov_search_state(cov_cell_t *h, cov_state_t *st);
/* 1 when every launch on the cell's stream has finished, 0 while any runs. */
int     cov_query(cov_cell_t *h);
/* Stop conditions, checked on the device before every step as the reference
 * checks them at the top of its loop (cov/engine/covsearch2e.c:1305-1324):
 * no step begins once best_uncovered <= target (--target; target < 0: none;
 * the state's done then reads 2), nor once `seconds` from now have passed on
 * the device clock (-t / --cpu; seconds == 0: none, seconds < 0: at once).
 * A new target restarts the time-to-target record; a target the best code
 * already meets is met at the current iteration (:1292-1293).  Reset to none
 * by cov_search_init. */
int     cov_search_limits(cov_cell_t *h, int64_t target, double seconds);
/* The time-to-target record (:1536-1542): the iteration whose move first
 * brought best_uncovered to the target (-1 if none yet), and the seconds of
 * device clock since that moment (-1 if none). */
int     cov_search_ttt(cov_cell_t *h, int64_t *ttt_it, double *age_s);

/* -------------------------------------------------- launch shape, timing */
/* Workgroups of the evaluator and commit kernels (the autotuner sets them). */
int     cov_set_launch(cov_cell_t *h, int eval_groups, int commit_groups);
int     cov_get_launch(const cov_cell_t *h, int *eval_groups, int *commit_groups);
/* How a step finds the leaving side of each candidate's delta (the words only
 * the moved codeword covers that the move uncovers).  COV_LEAVE_WALK walks
 * the candidates' distance-R spheres every step; COV_LEAVE_KEPT keeps per
 * codeword counts up to date through the commits (2 more bytes per word, a
 * few atomics per word whose count crosses 1) and recounts only the codewords
 * that moved; COV_LEAVE_AUTO (the default) keeps them when the average
 * coverage M*|ball|/q^n is 4 or more.  The moves are the same either way.  No
 * counterpart in the reference, whose evaluator counts per candidate
 * (cov/engine/covsearch2e.c). */
#define COV_LEAVE_AUTO 0
#define COV_LEAVE_WALK 1
#define COV_LEAVE_KEPT 2
int     cov_set_leave_mode(cov_cell_t *h, int mode);
/* LDS bytes per distance-transform tile (0 = the default, 49152) and threads
 * per transform workgroup (0 = the default, 512; else 256 or 1024); the tile
 * plan groups the n axes into as few HBM passes as tiles of this size allow. */
int     cov_set_transform_tile(cov_cell_t *h, int lds_bytes, int threads);
/* Accumulate per-kernel device time (on = 1): each launch on the cell's
 * stream records a start and a stop event around its own execution
 * (hipExtLaunchKernelGGL), read back in batches with no host sync between
 * launches; search steps are timed in plain mode only.  Read totals in ms and
 * launch counts for kernel kinds 0 = evaluator, 1 = selection and commit,
 * 2 = distance transform pass, 3 = recount, 4 = (unused: the selection runs in
 * the commit), 5 = post, 6 = count transform pass (nkinds <= 7). */
int     cov_timing(cov_cell_t *h, int on);
int     cov_timing_read(cov_cell_t *h, double *ms, int64_t *launches, int nkinds);
/* Diagnostics: microseconds spent since cov_search_init, from the device wall
 * clock, in: 0 select (the commit's start to the end of its selection, in
 * its workgroup 0), 1 post, 2 begin, 3 pick, 4 list upkeep, 5 shuffle,
 * 6 rng refill, 7 kicks, 8 the evaluator's span (previous begin's end ->
 * commit start), 9 the commit's span (selection end -> post start),
 * 10 gap previous begin -> evaluator, 11 (unused) (n <= 12 entries). */
int     cov_search_phases(cov_cell_t *h, double *us, int n);
/* Algorithmic bytes since the last cov_search_init, 2 B per counter read or
 * update (the uint16 cnt): what this engine's evaluator and commit kernels
 * read and write (eval_bytes, commit_bytes), and the reference's move as it
 * runs it, B_move = M n + 2 sum_cw (S_f + E_cw) + 8 S per step, every
 * candidate codeword's full distance-R sphere walked and E_cw its entering
 * side (walks, or the uncovered list), summed over move_steps steps. */
int     cov_search_bytes(cov_cell_t *h, int64_t *eval_bytes, int64_t *commit_bytes,
                         int64_t *move_bytes, int64_t *move_steps);

/* ------------------------------------------------------------ batches */
/* Many cells, each created, configured and initialised on its own (any
 * shapes), advanced by the same launches: every batched step is one
 * evaluator, one selection-and-commit and one post launch; the evaluator's
 * and the commit's workgroups are split among the cells (each cell's share
 * is its cov_set_launch groups, fixed at creation), the post runs one
 * workgroup per cell.  A
 * cell's trajectory in a batch is the one it has alone.  The reference runs
 * one cell per process (cov/engine/covsearch2e.c:1304 is a single-cell loop);
 * this has no counterpart there.  While a batch runs its cells belong to it:
 * no call on a member cell until cov_batch_sync returns, and no member is
 * destroyed before the batch. */
int         cov_batch_create(cov_cell_t *const *cells, int k, cov_batch_t **out);
int         cov_batch_destroy(cov_batch_t *b);
const char *cov_batch_last_error(const cov_batch_t *b);
int         cov_batch_size(const cov_batch_t *b);
/* nsteps more steps of every cell (fewer for a cell that covers); graph mode
 * replays captured batched steps.  launch returns at once, sync waits and
 * reports the first cell that stopped on an error; run = launch + sync. */
int         cov_batch_launch(cov_batch_t *b, int64_t nsteps, int mode);
int         cov_batch_sync(cov_batch_t *b);
/* 1 when the launched steps have finished, 0 while they run (no wait). */
int         cov_batch_query(cov_batch_t *b);
int         cov_batch_run(cov_batch_t *b, int64_t nsteps, int mode);

#ifdef __cplusplus
}
#endif
#endif
