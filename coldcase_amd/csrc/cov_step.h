// cov_step.h -- the device-resident tabu step (kernel a) and its narrow phases.
//
// One search step of cov/engine/covsearch2e.c:1304-1580 is three launches:
//
//   k_eval    (wide)  units of work over the candidate codewords:
//                       leaving side: per (codeword, R-subset of all n
//                         positions) the words of the distance-R sphere with
//                         cnt == 1, read from the ob plane; each count adds to
//                         T and to H_j for the j in the subset (:696-716)
//                       entering side: per (codeword, differing position p,
//                         R-subset avoiding p) the words of c' = c(p:=u_p)'s
//                         sphere with cnt == 0 (zb plane), or per (codeword,
//                         chunk of the uncovered list) the x with d(x,c)=R+1
//                         and x_p = u_p (:717-753)
//                     workgroup 0 computes the Fisher-Yates shuffle of the
//                     candidates meanwhile (the set of codewords to evaluate
//                     does not depend on the order), workgroup 1 refills the
//                     rng ring
//   k_commit  (wide)  every workgroup first selects the move (tabu,
//                     aspiration, reservoir ties) with a parallel prefix
//                     formulation of the sequential scan (:1503-1518), the
//                     same in each; then --cnt[a], ++cnt[b=a+dv] over the
//                     C(n-1,R)(q-1)^R patterns (every word once, no counter
//                     atomics, :562-590); plane bits, block counts and the
//                     uncovered list follow the transitions
//   k_post    (one workgroup)  applies the move, stamps tabu, tracks the best
//                     code, kicks on stagnation, writes the trace record and
//                     begins the next step: pick (:973-1016), distances,
//                     candidates (:1339-1376).
//
// Every decision that consumes randomness reads the ring of xoshiro256++
// draws at a cursor in the reference's order, so the trajectory is the
// reference's bit for bit.
//
// Synchronisation: kernel boundaries order each launch after the one before.
// A device-scope release is an L2 write-back on this part, so the wide kernels'
// workgroups end without one and the narrow phases run as kernels of their
// own.  Only the single move of the C ABI (k_commit_api) still ends with an
// arrival counter: its workgroups release onto it and the last one applies
// the move.  The single-workgroup phases read data that device atomics change
// (planes, block counts) with agent-scope loads.
#pragma once
#include "cov_common.h"
#include "cov_cover.h"

namespace cov {

constexpr int SHUF_PAR = 1024;     // draws handled by the parallel shuffle
constexpr int ATOM_DIRECT = 16;    // commit: lanes with transitions up to which a wave needs no runs
constexpr int EMIT_LDS = 1024;     // begin: slots listed in LDS before their digits are read
constexpr int SEL_PER = 4;         // candidates per lane in the one-wave selection
constexpr int SEL_WAVE = 64 * SEL_PER;   // the most candidates it takes
static_assert(SEL_WAVE == WG, "one staged candidate per thread");

constexpr int MAX_M_WORDS = (65535 + 63) / 64;   // 64-codeword mask words of the largest code

struct NarrowSh {
    WgRed red;
    WgScan scan;
    int64_t buf[WG];
    int64_t v64[8];
    int32_t v32[8];
    uint64_t qm[MAX_M_WORDS];   // begin: codewords at the target distance
    int32_t qfirst[MAX_M_WORDS]; // begin: first slot of each mask word's codewords
    int32_t elist[EMIT_LDS];     // begin: the codeword of each of the first EMIT_LDS slots
};

__device__ __forceinline__ uint32_t zbit_acq(const Cell& C, int64_t w)
{
    return (ld_acq32(&C.zb[w >> 5]) >> (w & 31)) & 1u;
}

// ---------------------------------------------------------- scanning
// first word of the plane-word range [x0, x1) with a set zb bit (masked first
// and last words), or INT64_MAX; all threads
__device__ __forceinline__ int64_t wg_first_bits(const Cell& C, int64_t x0, int64_t x1, uint32_t m0,
                                        uint32_t m1, NarrowSh& sh)
{
    int64_t best = INT64_MAX;
    for (int64_t xb = x0; xb < x1; xb += WG) {
        const int64_t x = xb + threadIdx.x;
        int64_t cand = INT64_MAX;
        if (x < x1) {
            uint32_t z = ld_acq32(&C.zb[x]);
            if (x == x0) z &= m0;
            if (x == x1 - 1) z &= m1;
            if (z) cand = x * 32 + __ffs(z) - 1;
        }
        const int64_t mn = wg_min(cand, sh.red);
        if (mn != INT64_MAX) { best = mn; break; }
    }
    return best;
}

// first uncovered word at or after `start`, cyclically (the scan of
// cov/engine/covsearch2e.c:1003-1015); -1 if none.  All threads.
__device__ __forceinline__ int64_t wg_first_from(const Cell& C, int64_t start, NarrowSh& sh)
{
    const int64_t sw = start >> 5;
    const int64_t sb = start >> BLK_SHIFT;
    const int64_t bw0 = sb << (BLK_SHIFT - 5);
    const int64_t bw1 = (bw0 + (BLK_WORDS / 32) < C.nzw) ? bw0 + (BLK_WORDS / 32) : C.nzw;
    const uint32_t hi_mask = ~0u << (start & 31);
    // 1. rest of the start block
    int64_t r = wg_first_bits(C, sw, bw1, hi_mask, ~0u, sh);
    if (r != INT64_MAX) return r;
    // 2. blocks after, 3. blocks before
    for (int pass = 0; pass < 2; pass++) {
        const int64_t b_lo = pass == 0 ? sb + 1 : 0;
        const int64_t b_hi = pass == 0 ? C.nblk : sb;
        for (int64_t bb = b_lo; bb < b_hi; bb += WG) {
            const int64_t b = bb + threadIdx.x;
            const int64_t cand = (b < b_hi && ld_acq32(&C.blk[b])) ? b : INT64_MAX;
            const int64_t fb = wg_min(cand, sh.red);
            if (fb != INT64_MAX) {
                const int64_t x0 = fb << (BLK_SHIFT - 5);
                const int64_t x1 = (x0 + (BLK_WORDS / 32) < C.nzw) ? x0 + (BLK_WORDS / 32) : C.nzw;
                return wg_first_bits(C, x0, x1, ~0u, ~0u, sh);
            }
        }
    }
    // 4. start of the start block, up to (not including) start
    const uint32_t lo_mask = (start & 31) ? ((1u << (start & 31)) - 1u) : 0u;
    if (sw > bw0 || lo_mask) {
        r = wg_first_bits(C, bw0, sw + 1, ~0u, lo_mask, sh);
        if (r != INT64_MAX && r < start) return r;
    }
    return -1;
}

// make sure `need` draws past the cursor exist (wave 0 generates, with 1024
// to spare); all threads
__device__ __forceinline__ void wg_ensure(const Cell& C, Ctl* ctl, int64_t need)
{
    __syncthreads();
    const int64_t ruse = ctl->ruse;
    const bool short_ = ctl->rgen < ruse + need;
    __syncthreads();
    if (short_) {
        if (threadIdx.x == 0 && ctl->rlow < ruse) ctl->rlow = ruse;
        __syncthreads();
        if (threadIdx.x < 64) wave_ring_fill(C, ctl, ruse + need + 1024, INT64_MAX);
        __syncthreads();
    }
}

// pick_uncovered: 64 probes, then the cyclic scan (:973-1016).  All threads.
// The reads go in three stages rather than one after another: the 65 draws
// (64 probes and the scan's start); then, together, the probes' plane bits,
// the rest of the start block and the counts of the 64 blocks after it; then,
// only when the start block has no uncovered word at or after the start, the
// words of the first non-empty block among those 64.  Past them, the general
// scan takes over.  The word returned is the sequential scan's.
constexpr int PICK_NB = WG - 64 - BLK_WORDS / 32;   // blocks after the start block read at once
static_assert(PICK_NB > 0, "the pick's second round needs threads for the block counts");
__device__ __forceinline__ int64_t wg_pick(const Cell& C, Ctl* ctl, NarrowSh& sh)
{
    wg_ensure(C, ctl, 65);
    const int tid = threadIdx.x, wv = tid >> 6;
    const int64_t pos = ctl->ruse;
    // stage 1: the draws, each reduced mod N where it is read: probe k by
    // lane k of wave 0, the scan's start (draw 64) by every other thread
    const uint64_t rv = ring_at(C, pos + (tid < 64 ? tid : 64));
    const int64_t wdraw = (int64_t)mod_magic(rv, (uint64_t)C.N, C.nmag);
    const int64_t start = tid < 64 ? 0 : wdraw;     // wave 0 reads it after the barrier below
    if (tid == 64) sh.v64[1] = wdraw;
    const int64_t sw = start >> 5, sb = start >> BLK_SHIFT;
    const int64_t bw0 = sb << (BLK_SHIFT - 5);
    const int64_t bw1 = (bw0 + (BLK_WORDS / 32) < C.nzw) ? bw0 + (BLK_WORDS / 32) : C.nzw;
    // stage 2
    int64_t cand_s = INT64_MAX, cand_b = INT64_MAX;
    if (tid < 64) {
        const int64_t w = wdraw;
        const bool hit = zbit_acq(C, w) != 0;
        const uint64_t mask = __ballot(hit);
        const int fh = mask ? (__ffsll((unsigned long long)mask) - 1) : -1;
        if (tid == 0) sh.v32[0] = fh;
        if (fh >= 0 && tid == fh) sh.v64[0] = w;
    } else if (tid < 64 + BLK_WORDS / 32) {
        const int64_t x = sw + (tid - 64);
        if (x < bw1) {
            uint32_t z = ld_acq32(&C.zb[x]);
            if (x == sw) z &= ~0u << (start & 31);
            if (z) cand_s = x * 32 + __ffs(z) - 1;
        }
    } else {
        const int64_t b = sb + 1 + (tid - 64 - BLK_WORDS / 32);
        if (b < C.nblk && ld_acq32(&C.blk[b])) cand_b = b;
    }
    // both candidates grow with the thread index (the plane word x, then the
    // block b, do), so a wave's least one is its first lane's that has one
    cand_s = wave_first64(cand_s != INT64_MAX, cand_s, INT64_MAX);
    cand_b = wave_first64(cand_b != INT64_MAX, cand_b, INT64_MAX);
    if ((tid & 63) == 0) { sh.buf[wv] = cand_s; sh.buf[WAVES + wv] = cand_b; }
    __syncthreads();
    int64_t ms = INT64_MAX, mb = INT64_MAX;
#pragma unroll
    for (int k = 0; k < WAVES; k++) {
        ms = sh.buf[k] < ms ? sh.buf[k] : ms;
        mb = sh.buf[WAVES + k] < mb ? sh.buf[WAVES + k] : mb;
    }
    const int fh = sh.v32[0];
    const int64_t wp = sh.v64[0];
    const int64_t start_all = sh.v64[1];
    __syncthreads();
    // (the cursor is thread 0's: every caller reads it next after a barrier)
    if (fh >= 0) {
        if (tid == 0) ctl->ruse = pos + fh + 1;
        return wp;
    }
    if (tid == 0) ctl->ruse = pos + 65;
    if (ms != INT64_MAX) return ms;
    if (mb != INT64_MAX) {
        // stage 3: the first non-empty block after the start block
        const int64_t x0 = mb << (BLK_SHIFT - 5);
        const int64_t x1 = (x0 + (BLK_WORDS / 32) < C.nzw) ? x0 + (BLK_WORDS / 32) : C.nzw;
        return wg_first_bits(C, x0, x1, ~0u, ~0u, sh);
    }
    return wg_first_from(C, start_all, sh);
}

// ------------------------------------------------------ uncovered list
__device__ __forceinline__ void wg_list_rebuild(const Cell& C, Ctl* ctl, NarrowSh& sh)
{
    if (threadIdx.x == 0) { ctl->ul_n = 0; ctl->ul_ovf = 0; }
    __syncthreads();
    for (int64_t x = threadIdx.x; x < C.nzw; x += WG) {
        const uint32_t z = ld_acq32(&C.zb[x]);
        C.inl[x] = z;
        uint32_t m = z;
        while (m) {
            const int b = __ffs(m) - 1;
            m &= m - 1;
            const int64_t w = x * 32 + b;
            const unsigned long long pos = atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->ul_n), 1ULL);
            if ((int64_t)pos < C.ulcap) { C.ul_idx[pos] = w; C.ul_pk[pos] = pack_index(C, w); }
            else atomicOr(&ctl->ul_ovf, 1);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ctl->ul_valid = ctl->ul_ovf ? 0 : 1;
        if (ctl->ul_n > C.ulcap) ctl->ul_n = C.ulcap;
        ctl->ul_ovf = 0;
    }
    __syncthreads();
}

// drop entries that are covered again (their membership bit is cleared).
// In place and in order, LIST_QK consecutive entries per thread per round:
// every round reads its entries before the scan's barriers and writes the
// kept ones below its own first entry, so no round overwrites an entry that
// is still to be read.
constexpr int LIST_QK = 8;
__device__ __forceinline__ void wg_list_compact(const Cell& C, Ctl* ctl, NarrowSh& sh)
{
    const int64_t n0 = ctl->ul_n;
    int64_t out = 0;
    for (int64_t e0 = 0; e0 < n0; e0 += (int64_t)WG * LIST_QK) {
        const int64_t eb = e0 + (int64_t)threadIdx.x * LIST_QK;
        int64_t idx[LIST_QK];
        uint64_t pk[LIST_QK];
        uint32_t keepm = 0;
        int kept = 0;
#pragma unroll
        for (int k = 0; k < LIST_QK; k++) {
            idx[k] = 0;
            pk[k] = 0;
            if (eb + k < n0) {
                idx[k] = C.ul_idx[eb + k];
                pk[k] = C.ul_pk[eb + k];
                if (zbit_acq(C, idx[k])) {
                    keepm |= 1u << k;
                    kept++;
                } else {
                    bit_clr(C.inl, idx[k]);
                }
            }
        }
        int64_t tot;
        int64_t at = out + wg_incl_sum(kept, sh.scan, &tot) - kept;
#pragma unroll
        for (int k = 0; k < LIST_QK; k++)
            if ((keepm >> k) & 1u) {
                C.ul_idx[at] = idx[k];
                C.ul_pk[at] = pk[k];
                at++;
            }
        out += tot;
    }
    __syncthreads();
    if (threadIdx.x == 0) ctl->ul_n = out;
    __syncthreads();
}

// ---------------------------------------------------------------- begin
// Start step it+1 if the run allows it: pick u, collect the candidate
// codewords (those at distance dtar from u, in index order), reserve the
// shuffle draws and lay out the evaluator's units.  All threads of one WG.
// (spw: the powers as pw_to_lds leaves them)
__device__ __forceinline__ void wg_begin(const Cell& C, Ctl* ctl, NarrowSh& sh, const int64_t* spw)
{
    const int64_t tb0 = clk();
    // one reading of the clock for the whole workgroup, so every thread takes
    // the same branch at the deadline
    __shared__ int64_t now;
    if (threadIdx.x == 0) now = tb0;
    __syncthreads();
    const int done = ctl->done, begun = ctl->begun;
    const int64_t unc = ctl->unc, it = ctl->it, it_stop = ctl->it_stop;
    const int64_t target = ctl->target, best = ctl->best_unc, deadline = ctl->deadline;
    const int64_t t_now = now;
    __syncthreads();
    if (done || begun) return;
    if (unc <= 0) {
        if (threadIdx.x == 0) ctl->done = 1;
        __syncthreads();
        return;
    }
    // the top of the reference's loop (:1305-1324): deadline, iteration
    // budget, target; each ends the run before the step's rng draws
    if (deadline != 0 && t_now >= deadline) return;
    if (it >= it_stop) return;
    if (target >= 0 && best <= target) {
        if (threadIdx.x == 0) ctl->done = 2;
        __syncthreads();
        return;
    }
    if (threadIdx.x == 0) ctl->it = it + 1;

    // The scan below reads every codeword's digits; up to SCAN_ILP * WG
    // codewords they are loaded now, SCAN_ILP per thread in flight (index
    // clamped to the last codeword, so no load is guarded by a branch), and
    // arrive while the list upkeep and the pick run.  The thread keeps them
    // in registers from the scan to the slot emission.
    constexpr int SCAN_ILP = 16;
    const bool one_round = C.M <= SCAN_ILP * WG;
    uint64_t c[SCAN_ILP] = {};
    // (rounds past the last codeword are skipped: the bound is uniform)
    auto load = [&](int i0) {
#pragma unroll
        for (int k = 0; k < SCAN_ILP; k++) {
            if (i0 + k * WG >= C.M) break;
            const int i = i0 + k * WG + threadIdx.x;
            c[k] = C.cpk[i < C.M ? i : C.M - 1];
        }
    };
    if (one_round) load(0);

    // uncovered-list upkeep (speed only: the list never changes a value); an
    // overflow since the last step (inserts beyond capacity) invalidates it
    const int64_t tl0 = clk();
    if (threadIdx.x == 0 && ctl->ul_ovf) {
        ctl->ul_valid = 0;
        ctl->ul_ovf = 0;
        ctl->ul_n = 0;
    }
    __syncthreads();
    {
        const int valid = ctl->ul_valid;
        const int64_t ul_n = ctl->ul_n;
        __syncthreads();
        if (ctl->wide) {
            // the wide neighbourhood's availability depends on the list's
            // own state, so it follows the reference's rules (:1326-1329):
            // rebuilt below 2^19 uncovered words, compacted, never dropped
            if (!valid && unc <= UL_REF_HALF) wg_list_rebuild(C, ctl, sh);
            else if (valid && ul_n > 2 * unc + 64) wg_list_compact(C, ctl, sh);
        } else if (!valid && unc <= list_keep(C) / 2) {
            wg_list_rebuild(C, ctl, sh);
        } else if (valid && unc > list_keep(C)) {
            // too long to be read: drop it rather than keep it up to date
            // (a rebuild rewrites the membership plane)
            if (threadIdx.x == 0) { ctl->ul_valid = 0; ctl->ul_n = 0; }
            __syncthreads();
        } else if (valid && ul_n > 2 * unc + 64) {
            wg_list_compact(C, ctl, sh);
        }
    }

    const int64_t tp0 = clk();
    const int64_t u = wg_pick(C, ctl, sh);
    if (threadIdx.x == 0) { ctl->tph[4] += tp0 - tl0; ctl->tph[3] += clk() - tp0; }
    if (u < 0) {
        if (threadIdx.x == 0) ctl->done = 1;
        __syncthreads();
        return;
    }
    // digits of u, computed by every wave (u is the workgroup's): position j
    // on lane j, then gathered lane by lane into scalar registers
    uint64_t upk = 0;
    {
        const int j = threadIdx.x & 63;
        const int64_t pwj = spw[j < MAXN ? j : 0];
        int64_t d = 0;
        if (j < C.n) {
            if (C.N <= 0xFFFFFFFFLL) d = (int64_t)(((uint32_t)u / (uint32_t)pwj) % (uint32_t)C.q);
            else d = (u / pwj) % C.q;
        }
#pragma unroll
        for (int k = 0; k < MAXN; k++)
            upk |= (uint64_t)__builtin_amdgcn_readlane((int)d, k) << (4 * k);
    }

    // One pass over the codewords, SCAN_ILP loads in flight per thread: the
    // minimum distance to u, and the codewords at distance R+1 as one bit
    // each (a ballot per wave, in LDS).  R+1 is the target distance unless no
    // codeword is that close; then a second pass marks those at the minimum
    // (:1342-1351).
    const int nwm = (C.M + 63) >> 6;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int64_t dmin = INT64_MAX;
    auto mark = [&](int i0, int target, bool with_min) {
#pragma unroll
        for (int k = 0; k < SCAN_ILP; k++) {
            if (i0 + k * WG >= C.M) break;
            const int i = i0 + k * WG + threadIdx.x;
            const int64_t d = nib_dist(c[k], upk);
            if (with_min && i < C.M && d < dmin) dmin = d;
            const uint64_t m = __ballot(i < C.M && d == target);
            const int wi = ((i0 + k * WG) >> 6) + wv;
            if (lane == 0 && wi < nwm) sh.qm[wi] = m;
        }
    };
    // codes above SCAN_DEEP codewords: SCAN_BIG loads in flight per thread
    // per round (the scan is a chain of load round trips on one workgroup);
    // below, rounds of SCAN_ILP cost fewer instructions
    constexpr int SCAN_BIG = 32, SCAN_DEEP = 16384;
    auto mark_pass = [&](int target, bool with_min) {
        if (one_round) {
            mark(0, target, with_min);
            return;
        }
        if (C.M <= SCAN_DEEP) {
            for (int i0 = 0; i0 < C.M; i0 += SCAN_ILP * WG) {
                load(i0);
                mark(i0, target, with_min);
            }
            return;
        }
        for (int i0 = 0; i0 < C.M; i0 += SCAN_BIG * WG) {
            uint64_t cb[SCAN_BIG];
#pragma unroll
            for (int k = 0; k < SCAN_BIG; k++) {
                const int i = i0 + k * WG + threadIdx.x;
                cb[k] = i < C.M ? C.cpk[i] : 0;
            }
#pragma unroll
            for (int k = 0; k < SCAN_BIG; k++) {
                const int i = i0 + k * WG + threadIdx.x;
                const int64_t d = nib_dist(cb[k], upk);
                if (with_min && i < C.M && d < dmin) dmin = d;
                const uint64_t m = __ballot(i < C.M && d == target);
                const int wi = ((i0 + k * WG) >> 6) + wv;
                if (lane == 0 && wi < nwm) sh.qm[wi] = m;
            }
        }
    };
    mark_pass(C.R + 1, true);
    __syncthreads();
    // a scan of the mask words' counts gives each word its first slot: for
    // one-round codes (at most 64 words) every wave scans them itself, word
    // j on lane j, into qf (no barrier); else a workgroup scan into LDS
    constexpr int QPT = (MAX_M_WORDS + WG - 1) / WG;   // mask words per thread
    static_assert(SCAN_ILP * WG <= 64 * 64, "one-round codes have at most 64 mask words");
    int64_t nq;
    int32_t qf = 0;
    auto slots = [&]() {
        if (one_round) {
            const int32_t cj = lane < nwm ? (int32_t)__popcll(sh.qm[lane]) : 0;
            int32_t inc = cj;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int32_t y = __shfl_up(inc, o, 64);
                if (lane >= o) inc += y;
            }
            qf = inc - cj;
            nq = __builtin_amdgcn_readlane(inc, 63);
            return;
        }
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < QPT; k++) {
            const int j = QPT * threadIdx.x + k;
            if (j < nwm) cnt += __popcll(sh.qm[j]);
        }
        int run = (int)(wg_incl_sum(cnt, sh.scan, &nq) - cnt);
#pragma unroll
        for (int k = 0; k < QPT; k++) {
            const int j = QPT * threadIdx.x + k;
            if (j < nwm) {
                sh.qfirst[j] = run;
                run += __popcll(sh.qm[j]);
            }
        }
        __syncthreads();
    };
    slots();
    // u is uncovered, so no codeword is nearer than R+1: with one at R+1 the
    // target is R+1, else the minimum distance (its own pass)
    int dtar = C.R + 1;
    if (nq == 0) {
        dmin = wg_min(dmin, sh.red);
        dtar = (dmin <= C.R + 1) ? C.R + 1 : (int)dmin;
        if (dtar != C.R + 1) {
            mark_pass(dtar, false);
            __syncthreads();
            slots();
        }
    }
    // codeword i takes slot first(i / 64) + its rank in its mask word: index
    // order.  A slot holds the codeword and its packed digits; its differing
    // positions are those where the digits differ from u's (slot_diff).
    auto emit = [&](int i, uint64_t mine) {
        const int64_t slot = sh.qfirst[i >> 6] + __popcll(sh.qm[i >> 6] & ((1ull << (i & 63)) - 1));
        C.slot_cw[slot] = i;
        C.slot_cpk[slot] = mine;
    };
    if (one_round) {
        // each wave, round by round, its mask word: a marked lane writes its
        // codeword at its slot (the word's first slot plus the lane's rank
        // among the word's marked lanes)
        const int wvu = __builtin_amdgcn_readfirstlane(wv);
#pragma unroll
        for (int k = 0; k < SCAN_ILP; k++) {
            if (k * WG >= C.M) break;
            const int wi = k * WAVES + wvu;         // codewords k WG + 64 wv + lane
            if (wi < nwm) {
                const uint64_t m = sh.qm[wi];
                const int32_t f = __builtin_amdgcn_readlane(qf, wi);
                if ((m >> lane) & 1) {
                    const int64_t slot = f + __popcll(m & ((1ull << lane) - 1));
                    C.slot_cw[slot] = k * WG + 64 * wvu + lane;
                    C.slot_cpk[slot] = c[k];
                }
            }
        }
    } else {
        // each thread lists the set bits of the QPT mask words it counted at
        // their slots (in LDS, past EMIT_LDS in slot_cw itself); then one slot
        // per thread reads its codeword's digits, the loads issued together
#pragma unroll
        for (int k = 0; k < QPT; k++) {
            const int j = QPT * threadIdx.x + k;
            if (j < nwm) {
                uint64_t m = sh.qm[j];
                int r = sh.qfirst[j];
                while (m) {
                    const int i = 64 * j + __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    if (r < EMIT_LDS) sh.elist[r] = i;
                    else C.slot_cw[r] = i;
                    r++;
                }
            }
        }
        __syncthreads();
        for (int64_t sl = threadIdx.x; sl < nq; sl += WG) {
            const int i = sl < EMIT_LDS ? sh.elist[sl] : C.slot_cw[sl];
            emit(i, C.cpk[i]);
        }
    }
    // list decision and unit counts; with --wide every move of the candidate
    // codewords is a candidate while the list is used (:1353-1376)
    const int use_list = ctl->ul_valid &&
                         (ctl->ul_n * C.n < (int64_t)(C.R + 1) * C.E * C.P);
    const int wnow = ctl->wide && use_list;
    const int per = wnow ? C.n * C.q1 : dtar;
    const int64_t ncand = nq * per;
    const int64_t take = ncand < ctl->cand ? ncand : ctl->cand;
    const int64_t nchunk = (ctl->ul_n + LIST_CHUNK - 1) / LIST_CHUNK;
    // the shuffle's draws (:1384-1389) are reserved now; the evaluator's
    // helper wave computes the permutation while the units run
    // (and, for the one-wave selection, the take + 1 draws after them, which
    // the evaluator stages for it)
    if (take <= RB - 4096) wg_ensure(C, ctl, take <= SEL_WAVE ? 2 * take + 2 : take);
    if (threadIdx.x == 0) {
        if (ncand == 0) { ctl->fault = 1; ctl->done = 1; }
        if (take > RB - 4096) { ctl->fault = 2; ctl->done = 1; }
        ctl->u = u;
        ctl->upk = upk;
        ctl->dtar = dtar;
        ctl->nq = (int32_t)nq;
        ctl->ncand = (int32_t)ncand;
        ctl->take = (int32_t)take;
        ctl->use_list = use_list;
        ctl->wnow = wnow;
        ctl->per = per;
        ctl->nchunk = use_list ? nchunk : 0;
        // the leaving side: the moved codewords' spheres when the counts are
        // kept (th), else those of every codeword at distance R+1 from u
        const int64_t nlv = (ctl->th_on ? ctl->nstale : nq) * (int64_t)C.F;
        ctl->n_leave = nlv;
        ctl->n_enter = use_list ? nq * nchunk : nq * dtar * C.E;
        ctl->sh_pos = ctl->ruse;
        ctl->rlow = ctl->ruse;
        ctl->ruse += take;
        const int64_t enter = use_list ? nq * ctl->ul_n * C.n : 2 * nq * dtar * C.E * C.P;
        ctl->eval_bytes += (int64_t)C.M * C.n + 2 * nlv * C.P + enter;
        ctl->move_bytes += (int64_t)C.M * C.n + 2 * nq * C.F * C.P + enter + 8 * (int64_t)C.E * C.P;
        ctl->move_steps += 1;
        ctl->selected = 0;
        ctl->begun = (ncand > 0 && take <= RB - 4096) ? 1 : 0;
        // the evaluator's ring refill: a chunk when fewer than 4096 draws are
        // ahead and the chunk overwrites no draw still to be read
        const int64_t g0 = ctl->rgen;
        ctl->fill_on = g0 < ctl->ruse + 4096 && ctl->budget >= RNG_CHUNK && g0 + RNG_CHUNK <= ctl->rlow + RB;
        ctl->fill_g0 = g0;
        for (int k = 0; k < 4; k++) ctl->fill_s[k] = ctl->rst[k];
        ctl->tph[2] += clk() - tb0;
    }
    __syncthreads();
}

// ------------------------------------------------------- evaluator units
// Set bits of `plane` over the P words of the sphere of (base, pk) on subset
// sub, summed over the wave.  K loads per lane are issued before the first is
// used, so a sphere of up to 64 K words costs one memory round trip: the
// loads are unconditional (a pattern number past P is clamped to P - 1, an
// in-bounds word whose bit is then masked), since a load guarded by its own
// branch is waited for inside that branch, one round trip per load.
template <int K>
__device__ __forceinline__ int64_t sphere_count_k(const Cell& C, const uint32_t* __restrict__ plane, int64_t base,
                                                  uint64_t pk, uint32_t sub, const WalkTab& tb, bool tab)
{
    const int lane = threadIdx.x & 63;
    const uint32_t last = (uint32_t)(C.P - 1);
    int64_t c = 0;
    for (int64_t t0 = 0; t0 < C.P; t0 += 64 * K) {
        uint32_t v[K], sh[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int64_t t = t0 + k * 64 + lane;
            const uint32_t tc = t < C.P ? (uint32_t)t : last;
            const int64_t w = tab ? walk_word(C, tb, base, tc) : pattern_word(C, base, pk, sub, C.R, tc, nullptr);
            v[k] = plane[w >> 5];
            sh[k] = (uint32_t)(w & 31);
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const bool in = t0 + k * 64 + lane < C.P;
            c += __popcll(__ballot(in && ((v[k] >> sh[k]) & 1u)));
        }
    }
    return c;
}
__device__ __forceinline__ int64_t sphere_count(const Cell& C, const uint32_t* __restrict__ plane, int64_t base,
                                                uint64_t pk, uint32_t sub, WalkTab& tb)
{
    // the unit's subset, centre digits and word are the wave's: made uniform,
    // the table build reads the cell's powers q^(n-1-j) with scalar loads
    // (a per-lane index into the kernel argument is a vector memory load)
    sub = __builtin_amdgcn_readfirstlane(sub);
    pk = uniform64(pk);
    base = (int64_t)uniform64((uint64_t)base);
    const bool tab = C.wok != 0;
    if (tab) walk_tables(C, pk, sub, tb);
    const int64_t c = C.P > 256 ? sphere_count_k<16>(C, plane, base, pk, sub, tb, tab)
                                : sphere_count_k<4>(C, plane, base, pk, sub, tb, tab);
    if (tab) walk_release();
    return c;
}

// leaving side: words of the distance-R sphere of codeword slot on subset f;
// four independent gathers in flight per lane
__device__ __forceinline__ void unit_leave(const Cell& C, int slot, int f, WalkTab& tb)
{
    const int lane = threadIdx.x & 63;
    const int i = C.slot_cw[slot];
    const int64_t base = C.cidx[i];
    const uint64_t pk = C.cpk[i];
    const uint32_t sub = C.fsub[f];
    const int64_t c1 = sphere_count(C, C.ob, base, pk, sub, tb);
    if (lane == 0 && c1) {
        int32_t* a = C.acc + (int64_t)slot * ACCW;
        atomicAdd(&a[0], (int)c1);
        for (int k = 0; k < C.R; k++) atomicAdd(&a[1 + ((sub >> (4 * k)) & 15)], (int)c1);
    }
}

// the leaving counts of moved codeword i (now at word idx, digits pk),
// subset f (its th row starts at zero)
__device__ __forceinline__ void unit_stale(const Cell& C, int i, int64_t idx, uint64_t pk, int f, WalkTab& tb)
{
    const int lane = threadIdx.x & 63;
    const uint32_t sub = C.fsub[f];
    const int64_t c1 = sphere_count(C, C.ob, idx, pk, sub, tb);
    if (lane == 0 && c1) {
        int32_t* t = C.th + (int64_t)i * THW;
        atomicAdd(&t[0], (int)c1);
        for (int k = 0; k < C.R; k++) atomicAdd(&t[1 + ((sub >> (4 * k)) & 15)], (int)c1);
    }
}

// entering side by walks: words of the sphere of c(p:=u_p) on subset e
// avoiding p, with cnt == 0
__device__ __forceinline__ void unit_walk(const Cell& C, int slot, int p, int up, int e, WalkTab& tb)
{
    const int lane = threadIdx.x & 63;
    const int i = C.slot_cw[slot];
    const uint64_t pk0 = C.cpk[i];
    const int cp = nib_get(pk0, p);
    const int64_t base = C.cidx[i] + (int64_t)(up - cp) * C.pw[p];
    const uint64_t pk = nib_put(pk0, p, up);
    const uint32_t sub = C.psub[(int64_t)p * C.E + e];
    const int64_t c0 = sphere_count(C, C.zb, base, pk, sub, tb);
    if (lane == 0 && c0) atomicAdd(&C.acc[(int64_t)slot * ACCW + 1 + MAXN + p], (int)c0);
}

// entering side from the uncovered list: entries e0..e1 against codeword slot
// --wide form: for every uncovered x at distance R+1 from the slot's
// codeword c, one count at (j, x_j) for each position j where x and c
// differ: the entering count of every move (j, v) of c (:720-734).  The
// counters (position j, value v) are spread over the lanes, lane l owning
// j * MAXQ + v = l, l + 64, l + 128, l + 192, and each takes the ballot
// counts of its pair.
__device__ __forceinline__ void unit_list_wide(const Cell& C, int slot, int64_t e0, int64_t e1, int64_t ul_n)
{
    const int lane = threadIdx.x & 63;
    const uint64_t cpk = C.cpk[C.slot_cw[slot]];
    // four named counters, not an array: a runtime index would put it in
    // private memory
    int32_t o0 = 0, o1 = 0, o2 = 0, o3 = 0;
    for (int64_t x0 = e0; x0 < e1; x0 += 64) {
        const int64_t x = x0 + lane;
        uint64_t ne = 0, xpk = 0;
        if (x < e1 && x < ul_n) {
            const int64_t xi = C.ul_idx[x];
            xpk = C.ul_pk[x];
            if (bit_of(C.zb, xi) && nib_dist(xpk, cpk) == C.R + 1) ne = nib_ne(xpk, cpk);
        }
        for (int j = 0; j < C.n; j++) {
            const bool bj = (ne >> (4 * j)) & 1;
            if (!__any(bj)) continue;
            const int xj = nib_get(xpk, j);
            for (int v = 0; v < C.q; v++) {
                const int c = __popcll(__ballot(bj && xj == v));
                const int id = j * MAXQ + v;
                if (c && lane == (id & 63)) {
                    const int kq = id >> 6;
                    o0 += kq == 0 ? c : 0;
                    o1 += kq == 1 ? c : 0;
                    o2 += kq == 2 ? c : 0;
                    o3 += kq == 3 ? c : 0;
                }
            }
        }
    }
    int32_t* a = C.accw + (int64_t)slot * WIDE_W;
    if (o0) atomicAdd(&a[lane], o0);
    if (o1) atomicAdd(&a[lane + 64], o1);
    if (o2) atomicAdd(&a[lane + 128], o2);
    if (o3) atomicAdd(&a[lane + 192], o3);
}

__device__ __forceinline__ void unit_list(const Cell& C, int slot, int64_t e0, int64_t e1,
                                          uint64_t upk, int64_t ul_n)
{
    const int lane = threadIdx.x & 63;
    const int i = C.slot_cw[slot];
    const uint64_t cpk = C.cpk[i];
    const uint64_t diff_cu = nib_ne(cpk, upk);      // positions p with c_p != u_p
    int32_t cnt[MAXN];
#pragma unroll
    for (int p = 0; p < MAXN; p++) cnt[p] = 0;
    for (int64_t x0 = e0; x0 < e1; x0 += 64) {
        const int64_t x = x0 + lane;
        uint64_t hitmask = 0;
        if (x < e1 && x < ul_n) {
            const int64_t xi = C.ul_idx[x];
            const uint64_t xpk = C.ul_pk[x];
            if (bit_of(C.zb, xi) && nib_dist(xpk, cpk) == C.R + 1)
                hitmask = ~nib_ne(xpk, upk) & diff_cu & NIB1;   // x_p == u_p != c_p
        }
#pragma unroll
        for (int p = 0; p < MAXN; p++) {
            if (p < C.n && ((diff_cu >> (4 * p)) & 1))
                cnt[p] += __popcll(__ballot((hitmask >> (4 * p)) & 1));
        }
    }
    if (lane == 0) {
        int32_t* a = C.acc + (int64_t)slot * ACCW + 1 + MAXN;
#pragma unroll
        for (int p = 0; p < MAXN; p++)
            if (cnt[p]) atomicAdd(&a[p], cnt[p]);
    }
}

// quotient and remainder, in 32-bit arithmetic when both fit (the unit
// numbers of every workload do; the 64-bit division is a long routine)
__device__ __forceinline__ int64_t div_small(int64_t a, int64_t b, int64_t* r)
{
    if (((uint64_t)a | (uint64_t)b) >> 32 == 0) {
        const uint32_t q = (uint32_t)a / (uint32_t)b;
        *r = (int64_t)((uint32_t)a - q * (uint32_t)b);
        return q;
    }
    const int64_t q = a / b;
    *r = a - q * b;
    return q;
}

// run this wave's share of the step's evaluator units
__device__ __forceinline__ void eval_units(const Cell& C, const Ctl* ctl, int64_t wid, int64_t nw, WalkTab& tb)
{
    const int64_t nl = ctl->n_leave, ne = ctl->n_enter;
    const int use_list = ctl->use_list, dtar = ctl->dtar, wnow = ctl->wnow;
    const int64_t nchunk = ctl->nchunk, ul_n = ctl->ul_n;
    const uint64_t upk = ctl->upk;
    const int64_t per_slot_walk = (int64_t)dtar * C.E;
    const int th_on = ctl->th_on;
    for (int64_t g = wid; g < nl + ne; g += nw) {
        if (g < nl) {
            int64_t f;
            const int slot = (int)div_small(g, C.F, &f);
            if (th_on)
                unit_stale(C, ctl->stale[slot], ctl->stale_idx[slot], ctl->stale_pk[slot], (int)f, tb);
            else unit_leave(C, slot, (int)f, tb);
        } else if (use_list) {
            int64_t ch;
            const int slot = (int)div_small(g - nl, nchunk, &ch);
            const int64_t e0 = ch * LIST_CHUNK;
            const int64_t e1 = (e0 + LIST_CHUNK < ul_n) ? e0 + LIST_CHUNK : ul_n;
            if (wnow) unit_list_wide(C, slot, e0, e1, ul_n);
            else unit_list(C, slot, e0, e1, upk, ul_n);
        } else {
            int64_t r, e64;
            const int slot = (int)div_small(g - nl, per_slot_walk, &r);
            const int t = (int)div_small(r, C.E, &e64);
            const int e = (int)e64;
            const int p = slot_diff(C.slot_cpk[slot], upk, t);
            unit_walk(C, slot, p, nib_get(upk, p), e, tb);
        }
    }
}

// ------------------------------------------------------------- shuffle
struct ShufSh {
    uint32_t j[SHUF_PAR + 4] __attribute__((aligned(16)));   // swap partner of each draw
    int32_t g[SHUF_PAR];
    int32_t f[SHUF_PAR];
    int32_t off[SHUF_PAR + 1];   // draws per partner value, then where each value's draws start
    int32_t item[SHUF_PAR];      // the draws grouped by partner value
    int32_t wsum[WG / 64];
};

// Partial Fisher-Yates over the ncand candidates (original index k =
// slot*dtar + t): `take` draws from ring position sh_pos, the s-th swapping
// slot s with j_s = s + r_s mod (ncand - s) (:1383-1389).  Writes C.shuf.
//
// Parallel form (take <= SHUF_PAR).  Slot s is final after swap s and holds
// what slot j_s held just before it.  A slot p >= s is written before time s
// only by swaps t < s with j_t = p (its own swap comes at time p >= s), and
// such a swap puts there what slot t held at time t.  So with
//   g(s) = last t < s with j_t = j_s,   f(t) = last t' < t with j_t' = t,
// out[s] = j_s if g(s) < 0, else root(g(s)), where root(t) = t if f(t) < 0,
// else root(f(t)).  All threads of one workgroup; every thread then follows
// its own chains.  g and f are both "the last earlier draw with a given
// partner value": with at most SHUF_PAR candidates the draws are grouped by
// partner value (a count, a prefix sum and a scatter in LDS), and each draw
// looks through its group, a few entries; with more candidates the (draw,
// chunk of 32 earlier draws) pairs are scanned with LDS maxima.
// Returns the calling thread's own entry C.shuf[tid] when the parallel
// forms ran (tid < take), else ~0u.
__device__ __forceinline__ uint32_t wg_shuffle(const Cell& C, const Ctl* ctl, ShufSh& sh)
{
    uint32_t mine = ~0u;
    const int tid = threadIdx.x;
    const int64_t ncand = ctl->ncand, take = ctl->take, pos0 = ctl->sh_pos;
    if (ncand <= SHUF_PAR) {
        const int nc = (int)ncand, tk = (int)take;
        for (int v = tid; v <= nc; v += WG) sh.off[v] = 0;
        for (int s = tid; s < tk; s += WG)
            sh.j[s] = (uint32_t)(s + (int)mod_u16(ring_at(C, pos0 + s), (uint32_t)(nc - s)));
        __syncthreads();
        for (int s = tid; s < tk; s += WG) atomicAdd(&sh.off[sh.j[s]], 1);
        __syncthreads();
        // exclusive prefix sum of the counts: thread tid owns values
        // 4 tid .. 4 tid + 3 (SHUF_PAR = 4 WG)
        static_assert(SHUF_PAR == 4 * WG, "four values per thread");
        const int lane = tid & 63, wv = tid >> 6;
        int c4[4], tot = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int v = 4 * tid + k;
            c4[k] = v < nc ? sh.off[v] : 0;
            tot += c4[k];
        }
        int incl = tot;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        if (lane == 63) sh.wsum[wv] = incl;
        __syncthreads();
        int base = incl - tot;
        for (int w = 0; w < wv; w++) base += sh.wsum[w];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int v = 4 * tid + k;
            if (v < nc) {
                sh.off[v] = base;
                sh.g[v] = base;        // the scatter's cursor for value v
            }
            base += c4[k];
        }
        if (tid == 0) sh.off[nc] = tk;
        __syncthreads();
        for (int s = tid; s < tk; s += WG) sh.item[atomicAdd(&sh.g[sh.j[s]], 1)] = s;
        __syncthreads();
        // the last earlier draw with partner j_s, and with partner s
        int gs[SHUF_PAR / WG], fs[SHUF_PAR / WG];
#pragma unroll
        for (int k = 0; k < SHUF_PAR / WG; k++) {
            const int s = k * WG + tid;
            gs[k] = -1;
            fs[k] = -1;
            if (s < tk) {
                const int v = (int)sh.j[s];
                for (int x = sh.off[v]; x < sh.off[v + 1]; x++) {
                    const int t = sh.item[x];
                    if (t < s && t > gs[k]) gs[k] = t;
                }
                for (int x = sh.off[s]; x < sh.off[s + 1]; x++) {
                    const int t = sh.item[x];
                    if (t < s && t > fs[k]) fs[k] = t;
                }
            }
        }
        // (no barrier: the searches read j, off and item, not g or f; the
        // scatter's last use of g as a cursor was before the barrier above)
#pragma unroll
        for (int k = 0; k < SHUF_PAR / WG; k++) {
            const int s = k * WG + tid;
            if (s < tk) {
                sh.g[s] = gs[k];
                sh.f[s] = fs[k];
            }
        }
        __syncthreads();
        for (int s = tid; s < tk; s += WG) {
            int t = sh.g[s];
            uint32_t out;
            if (t < 0) {
                out = sh.j[s];
            } else {
                while (sh.f[t] >= 0) t = sh.f[t];
                out = (uint32_t)t;
            }
            C.shuf[s] = out;
            if (s == tid) mine = out;
        }
        return mine;
    }
    if (take <= SHUF_PAR) {
        for (int64_t s = tid; s < take; s += WG) {
            sh.j[s] = (uint32_t)(s + (int64_t)mod_any(ring_at(C, pos0 + s), (uint64_t)(ncand - s)));
            sh.g[s] = -1;
            sh.f[s] = -1;
        }
        if (tid < 4 && take + tid < SHUF_PAR + 4) sh.j[take + tid] = 0xFFFFFFFFu;
        __syncthreads();
        // every (draw s, chunk of 32 earlier draws) pair in parallel
        const int64_t nchunk = (take + 31) / 32;
        for (int64_t x = tid; x < take * nchunk; x += WG) {
            const int64_t s = x % take, c = x / take;
            const int64_t t0 = c * 32;
            if (t0 >= s) continue;
            const int64_t t1 = (t0 + 32 < s) ? t0 + 32 : s;
            const uint32_t js = sh.j[s];
            int32_t g = -1, f = -1;
            for (int64_t t4 = t0; t4 < t1; t4 += 4) {
                const uint4 q4 = *reinterpret_cast<const uint4*>(&sh.j[t4]);
                const uint32_t v[4] = {q4.x, q4.y, q4.z, q4.w};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int32_t t = (int32_t)(t4 + k);
                    if (t < t1) {
                        if (v[k] == js) g = t;
                        if (v[k] == (uint32_t)s) f = t;
                    }
                }
            }
            if (g >= 0) atomicMax(&sh.g[s], g);
            if (f >= 0) atomicMax(&sh.f[s], f);
        }
        __syncthreads();
        for (int64_t s = tid; s < take; s += WG) {
            int32_t t = sh.g[s];
            uint32_t out;
            if (t < 0) {
                out = sh.j[s];
            } else {
                while (sh.f[t] >= 0) t = sh.f[t];
                out = (uint32_t)t;
            }
            C.shuf[s] = out;
            if (s == tid) mine = out;
        }
        return mine;
    }
    // serial path for large samples (thread 0, the candidate array in HBM)
    uint32_t* a = C.shuf;
    for (int64_t k = tid; k < ncand; k += WG) a[k] = (uint32_t)k;
    __syncthreads();
    for (int64_t s0 = 0; s0 < take; s0 += WG) {
        const int64_t s = s0 + tid;
        if (s < take) sh.j[tid] = (uint32_t)(s + (int64_t)(ring_at(C, pos0 + s) % (uint64_t)(ncand - s)));
        __syncthreads();
        if (tid == 0) {
            const int64_t lim = (take - s0 < WG) ? take - s0 : WG;
            for (int64_t t = 0; t < lim; t++) {
                const int64_t ss = s0 + t;
                const uint32_t j = sh.j[t];
                const uint32_t x = a[ss], y = a[j];
                a[j] = x;
                a[ss] = y;
            }
        }
        __syncthreads();
    }
    return mine;
}

// Candidate k of the unshuffled order: its slot (candidate codeword),
// position and new value.  Standard: k = slot * dtar + t, the t-th position
// where the codeword differs from u, new value u_p.  --wide: k = slot * n(q-1)
// + p(q-1) + t, the t-th value other than c_p (:1359-1376 order).
__device__ __forceinline__ void cand_move(const Cell& C, const Ctl* ctl, uint32_t k, int* pslot, int* pp, int* pv)
{
    const int per = ctl->per;
    const int slot = (int)(k / (uint32_t)per);
    const int r = (int)(k - (uint32_t)slot * (uint32_t)per);
    int p, v;
    if (ctl->wnow) {
        p = r / C.q1;
        const int t = r - p * C.q1;
        const int cp = nib_get(C.cpk[C.slot_cw[slot]], p);
        v = t + (t >= cp);
    } else {
        p = slot_diff(C.slot_cpk[slot], ctl->upk, r);
        v = nib_get(ctl->upk, p);
    }
    *pslot = slot;
    *pp = p;
    *pv = v;
}

// the entering count of move (p, v) of the slot's codeword
__device__ __forceinline__ int32_t cand_enter(const Cell& C, const Ctl* ctl, int slot, int p, int v)
{
    return ctl->wnow ? C.accw[(int64_t)slot * WIDE_W + p * MAXQ + v] : C.acc[(int64_t)slot * ACCW + 1 + MAXN + p];
}

// The one-wave selection's reads that do not depend on the evaluator's
// counts, made by the shuffle's workgroup once the order is known: for each
// of the first take <= SEL_WAVE candidates in shuffled order, its codeword,
// slot, position, new and old digit, tabu stamp and the codeword's word
// index; and the take + 1 ring draws from the selection's cursor (begin made
// sure they exist).  Nothing the evaluator runs changes them.  All threads of
// the group.
// The first STAGE_SLOTS slots' records, copied into LDS by the shuffle's
// workgroup at its start (loads issued with the control block's), so that the
// staging below reads no memory for them.
constexpr int STAGE_SLOTS = 128;
struct StageSh {
    int32_t cw[STAGE_SLOTS];
    uint64_t cpk[STAGE_SLOTS];
};
struct StageReg {
    int32_t cw;
    uint64_t cpk;
};
__device__ __forceinline__ StageReg stage_load(const Cell& C)
{
    StageReg r{0, 0};
    if (threadIdx.x < STAGE_SLOTS) {
        const int64_t s = threadIdx.x < C.M ? threadIdx.x : C.M - 1;
        r.cw = C.slot_cw[s];
        r.cpk = C.slot_cpk[s];
    }
    return r;
}
__device__ __forceinline__ void stage_to_lds(const StageReg& r, StageSh& st)
{
    const int t = threadIdx.x;
    if (t < STAGE_SLOTS) {
        st.cw[t] = r.cw;
        st.cpk[t] = r.cpk;
    }
}
__device__ __forceinline__ void wg_stage(const Cell& C, const Ctl* ctl, uint32_t mine, const StageSh& st)
{
    const int take = ctl->take;
    if (take > SEL_WAVE) return;
    // take <= SEL_WAVE = WG <= SHUF_PAR: the shuffle ran a parallel form, in
    // which this thread made entry s = tid itself and returned it
    static_assert(SEL_WAVE <= WG && SEL_WAVE <= SHUF_PAR, "the staged entries come from the shuffle's own threads");
    const int s = threadIdx.x;
    if (s < take) {
        // cand_move's candidate, from the slot records
        const int per = ctl->per;
        const int slot = (int)(mine / (uint32_t)per);
        const int r = (int)(mine - (uint32_t)slot * (uint32_t)per);
        const bool in = slot < STAGE_SLOTS;
        const int i = in ? st.cw[slot] : C.slot_cw[slot];
        const uint64_t cpk = in ? st.cpk[slot] : C.slot_cpk[slot];
        // the word index from the digits (Horner's rule: position 0 is the
        // most significant digit)
        int64_t idx = 0;
        for (int j = 0; j < C.n; j++) idx = idx * C.q + nib_get(cpk, j);
        int p, v;
        if (ctl->wnow) {
            p = r / C.q1;
            const int t = r - p * C.q1;
            const int cp = nib_get(cpk, p);
            v = t + (t >= cp);
        } else {
            p = slot_diff(cpk, ctl->upk, r);
            v = nib_get(ctl->upk, p);
        }
        const int cold = nib_get(cpk, p);
        const int32_t stamp = C.tabu[((int64_t)i * C.n + p) * C.q + v];
        reinterpret_cast<int4*>(C.cst)[s] = make_int4(i, slot, p | (v << 8) | (cold << 16), stamp);
        C.cbs[s] = idx;
        C.cps[s] = cpk;
    }
    if (s <= take) C.csr[s] = ring_at(C, ctl->ruse + s);
    if (s == 0 && take == SEL_WAVE) C.csr[SEL_WAVE] = ring_at(C, ctl->ruse + SEL_WAVE);
}

// The one-wave selection's inputs for lane `lane` of the commit's wave 0,
// loaded before the control block is known: the staged entries and draws of
// candidates SEL_PER lane .. + SEL_PER - 1 (entries past `take` are stale and
// unused; their indices are clamped so every load stays in bounds), then, from
// those, both forms of each count (the accumulators and the kept counts, the
// narrow and the --wide entering counts): the control block says which to
// use.  The kept counts (th) do not change during the commit: its events go
// to dth.
struct SelPre {
    int4 st[SEL_PER];
    int64_t cb[SEL_PER];
    uint64_t cp[SEL_PER];
    uint64_t rv[SEL_PER];
    uint64_t rv_last;
    int32_t a0a[SEL_PER], apa[SEL_PER], ena[SEL_PER], a0t[SEL_PER], apt[SEL_PER], enw[SEL_PER];
};
__device__ __forceinline__ void sel_preload(const Cell& C, int lane, SelPre& pr)
{
#pragma unroll
    for (int k = 0; k < SEL_PER; k++) {
        const int s = SEL_PER * lane + k;
        pr.st[k] = reinterpret_cast<const int4*>(C.cst)[s];
        pr.cb[k] = C.cbs[s];
        pr.cp[k] = C.cps[s];
        pr.rv[k] = C.csr[s];
    }
    pr.rv_last = C.csr[SEL_WAVE];
#pragma unroll
    for (int k = 0; k < SEL_PER; k++) {
        const uint32_t ci = (uint32_t)pr.st[k].x < (uint32_t)C.M ? (uint32_t)pr.st[k].x : 0u;
        const uint32_t slot = (uint32_t)pr.st[k].y < (uint32_t)C.M ? (uint32_t)pr.st[k].y : 0u;
        const uint32_t p0 = (uint32_t)pr.st[k].z & 255u, v0 = ((uint32_t)pr.st[k].z >> 8) & 255u;
        const uint32_t p = p0 < (uint32_t)C.n ? p0 : 0u, v = v0 < (uint32_t)C.q ? v0 : 0u;
        const int32_t* a = C.acc + (int64_t)slot * ACCW;
        const int32_t* t = C.th + (int64_t)ci * THW;
        pr.a0a[k] = a[0];
        pr.apa[k] = a[1 + p];
        pr.ena[k] = a[1 + MAXN + p];
        pr.a0t[k] = t[0];
        pr.apt[k] = t[1 + p];
        pr.enw[k] = C.accw ? C.accw[(int64_t)slot * WIDE_W + p * MAXQ + v] : 0;
    }
}

// ------------------------------------------------------------- select
struct SelSh {
    WgRed red;
    WgScan scan;
    int64_t inc[WG];
    uint64_t rv[SEL_WAVE + 1];  // one-wave selection: ring draws r0 .. r0 + take
    int64_t pick[7];            // the selected candidate: i, p, v, old value, word index, delta, digits
};

// inclusive scans over the 64 lanes of a wave (lane order)
__device__ __forceinline__ int64_t wave_incl_min(int64_t x)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(x, o, 64);
        if (lane >= o && y < x) x = y;
    }
    return x;
}
__device__ __forceinline__ int64_t wave_incl_max(int64_t x)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(x, o, 64);
        if (lane >= o && y > x) x = y;
    }
    return x;
}
__device__ __forceinline__ int64_t wave_incl_add(int64_t x)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}
// the value of the lane below (the exclusive form of a scan); `first` on lane 0
__device__ __forceinline__ int64_t wave_excl(int64_t incl, int64_t first)
{
    const int64_t y = __shfl_up(incl, 1, 64);
    return (threadIdx.x & 63) == 0 ? first : y;
}

__device__ __forceinline__ int64_t cand_delta(const Cell& C, const Ctl* ctl, uint32_t k,
                                              int* pi, int* pp, int* pv)
{
    int slot, p, v;
    cand_move(C, ctl, k, &slot, &p, &v);
    const int32_t* a = C.acc + (int64_t)slot * ACCW;
    const int i = C.slot_cw[slot];
    *pi = i;
    *pp = p;
    *pv = v;
    const int32_t* lv = ctl->th_on ? C.th + (int64_t)i * THW : a;
    return (int64_t)(lv[0] - lv[1 + p]) - (int64_t)cand_enter(C, ctl, slot, p, v);
}

// The selection below for take <= SEL_WAVE candidates, on wave 0 alone: lane l
// holds candidates SEL_PER*l .. SEL_PER*l + SEL_PER-1 of the shuffled order,
// so the three prefix scans are a few register steps within the lane plus a
// wave scan, with no workgroup barrier; the ring draws r0 .. r0 + take sit in
// LDS, and the lane holding the selected candidate publishes it.  Same
// results as the workgroup form.  Called by all threads.
__device__ __forceinline__ void wave_select(const Cell& C, Ctl* ctl, SelSh& sh, const SelPre& pr)
{
    const int tid = threadIdx.x;
    if (tid < 64) {
        const int lane = tid;
        const int64_t take = ctl->take, it = ctl->it, unc = ctl->unc, best = ctl->best_unc;
        const int64_t r0 = ctl->ruse;
        int ci[SEL_PER], cp[SEL_PER], cv[SEL_PER], cold[SEL_PER];
        int64_t cd[SEL_PER], cbase[SEL_PER];
        uint64_t cpks[SEL_PER];
        bool valid[SEL_PER];
        // the loads were made by sel_preload: pick the forms the control
        // block names
        const bool th_on = ctl->th_on != 0, wnow = ctl->wnow != 0;
        int32_t a0[SEL_PER], ap[SEL_PER], an[SEL_PER], stamp[SEL_PER];
#pragma unroll
        for (int k = 0; k < SEL_PER; k++) {
            ci[k] = pr.st[k].x;
            cp[k] = pr.st[k].z & 255;
            cv[k] = (pr.st[k].z >> 8) & 255;
            cold[k] = pr.st[k].z >> 16;
            stamp[k] = pr.st[k].w;
            cbase[k] = pr.cb[k];
            cpks[k] = pr.cp[k];
            a0[k] = th_on ? pr.a0t[k] : pr.a0a[k];
            ap[k] = th_on ? pr.apt[k] : pr.apa[k];
            an[k] = wnow ? pr.enw[k] : pr.ena[k];
        }
#pragma unroll
        for (int k = 0; k < SEL_PER; k++) {
            const int64_t s = (int64_t)SEL_PER * lane + k;
            cd[k] = INT64_MAX;
            valid[k] = false;
            if (s < take) {
                cd[k] = (int64_t)(a0[k] - ap[k]) - (int64_t)an[k];
                valid[k] = !(stamp[k] > it && unc + cd[k] >= best);
            }
        }
#pragma unroll
        for (int k = 0; k < SEL_PER; k++) {
            const int64_t s = (int64_t)SEL_PER * lane + k;
            if (s <= take) sh.rv[s] = pr.rv[k];
        }
        if (lane == 0 && take == SEL_WAVE) sh.rv[SEL_WAVE] = pr.rv_last;
        // The scans below run on 32-bit values (a delta is a difference of
        // sphere counts; T counts at most SEL_WAVE ties), one lane shuffle per
        // step; the prefix and total maxima use the order of T instead.
        // exclusive prefix minimum of the valid deltas
        int32_t x[SEL_PER], lmin = INT32_MAX;
#pragma unroll
        for (int k = 0; k < SEL_PER; k++) {
            x[k] = valid[k] ? (int32_t)cd[k] : INT32_MAX;
            lmin = x[k] < lmin ? x[k] : lmin;
        }
        int32_t im = lmin;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t y = __shfl_up(im, o, 64);
            if (lane >= o && y < im) im = y;
        }
        int32_t pm = __shfl_up(im, 1, 64);
        if (lane == 0) pm = INT32_MAX;
        bool nb[SEL_PER], tie[SEL_PER];
        int32_t nt = 0;
#pragma unroll
        for (int k = 0; k < SEL_PER; k++) {
            nb[k] = valid[k] && x[k] < pm;
            tie[k] = valid[k] && x[k] == pm;
            nt += tie[k] ? 1 : 0;
            pm = x[k] < pm ? x[k] : pm;
        }
        // T: ties up to each candidate; Tb: the last restart's T (at least 0)
        int32_t tincl = nt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t y = __shfl_up(tincl, o, 64);
            if (lane >= o) tincl += y;
        }
        int32_t T = tincl - nt;
        int32_t Tk[SEL_PER], mk = -1;
#pragma unroll
        for (int k = 0; k < SEL_PER; k++) {
            T += tie[k] ? 1 : 0;
            Tk[k] = T;
            if (nb[k]) mk = T;
        }
        // T never decreases in the candidates' order, so the lanes' mk (where
        // set) do not either: the exclusive prefix maximum is the mk of the
        // nearest lane below with one, the maximum that of the last such lane
        const uint64_t mkm = __ballot(mk >= 0);
        const uint64_t below = mkm & ((1ull << lane) - 1);
        const int32_t mkb = __shfl(mk, below ? 63 - __clzll((long long)below) : 0, 64);
        int32_t tb = below ? mkb : -1;
        // the draws written by other lanes of the wave are read below
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        int32_t wl = -1;
#pragma unroll
        for (int k = 0; k < SEL_PER; k++) {
            if (nb[k]) tb = Tk[k];
            const int32_t Tb = tb < 0 ? 0 : tb;
            bool ev = nb[k];
            if (tie[k]) ev = mod_any(sh.rv[Tk[k] - 1], (uint64_t)(1 + Tk[k] - Tb)) == 0;
            if (ev) wl = SEL_PER * lane + k;
        }
        // (the candidate indices grow with the lane: the largest is the last
        // lane's that has one)
        const uint64_t wlm = __ballot(wl >= 0);
        const int64_t winner = wlm ? __builtin_amdgcn_readlane(wl, 63 - __clzll((long long)wlm)) : -1;
        const int64_t Tall = __builtin_amdgcn_readlane(tincl, 63);
        const int64_t Tball = mkm ? __builtin_amdgcn_readlane(mk, 63 - __clzll((long long)mkm)) : 0;
        const int64_t ssel = winner >= 0 ? winner : (int64_t)mod_any(sh.rv[0], (uint64_t)take);
#pragma unroll
        for (int k = 0; k < SEL_PER; k++) {
            if ((int64_t)SEL_PER * lane + k == ssel) {
                sh.pick[0] = ci[k]; sh.pick[1] = cp[k]; sh.pick[2] = cv[k]; sh.pick[3] = cold[k];
                sh.pick[4] = cbase[k]; sh.pick[5] = cd[k]; sh.pick[6] = (int64_t)cpks[k];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane == 0) {
            const int64_t draws = winner >= 0 ? Tall : 1;
            const uint64_t rs = sh.rv[draws];
            ctl->ruse = r0 + draws + 1;
            ctl->rlow = ctl->ruse;
            const int bi = (int)sh.pick[0], bp = (int)sh.pick[1], bv = (int)sh.pick[2], oldv = (int)sh.pick[3];
            ctl->bi = bi; ctl->bp = bp; ctl->bv = bv; ctl->oldv = oldv;
            ctl->dsel = winner >= 0 ? sh.pick[5] : DSEL_RANDOM;
            ctl->nties = winner >= 0 ? (int)(1 + Tall - Tball) : 0;
            ctl->stamp = (int32_t)(it + ctl->tenure + (int64_t)(rs % 4));
            ctl->base = sh.pick[4];
            ctl->bpk = (uint64_t)sh.pick[6];
            ctl->dv = (int64_t)(bv - oldv) * C.pw[bp];
            ctl->ncw = ctl->nq;
            ctl->selected = 1;
        }
    }
}

// Selection of :1503-1518 in parallel.  In shuffled order, a valid candidate
// (not tabu, or aspirating) with d below the running minimum restarts the tie
// count; one equal to it draws rmod(nties).  The running minimum before s is
// an exclusive prefix minimum; T(s) counts ties up to s and every tie's draw
// is the T(s)-th of the step; nties(s) = 1 + T(s) - T(last restart <= s).
// The winner is the last restart or successful tie.  All threads of one WG.
__device__ __forceinline__ void wg_select(const Cell& C, Ctl* ctl, SelSh& sh, const SelPre& pr)
{
    const int tid = threadIdx.x;
    // up to SEL_WAVE candidates: wave 0 alone, from the evaluator's staged
    // entries and draws (begin reserved the draws; the ring is not read here)
    if (ctl->take <= SEL_WAVE) {
        wave_select(C, ctl, sh, pr);
        return;
    }
    if (tid == 0) ctl->rlow = ctl->ruse;
    __syncthreads();
    const bool short_ = ctl->rgen < ctl->ruse + ctl->take + 2;
    __syncthreads();
    if (short_ && tid < 64) wave_ring_fill(C, ctl, ctl->ruse + ctl->take + 2 + 1024, INT64_MAX);
    __syncthreads();
    const int64_t take = ctl->take, it = ctl->it, unc = ctl->unc, best = ctl->best_unc;
    const int64_t r0 = ctl->ruse;
    const int n = C.n, q = C.q;
    {
    int64_t carry_pm = INT64_MAX, carry_T = 0, carry_Tb = 0, winner = -1;
    for (int64_t s0 = 0; s0 < take; s0 += WG) {
        const int64_t s = s0 + tid;
        bool valid = false;
        int64_t d = INT64_MAX;
        if (s < take) {
            int i, p, v;
            d = cand_delta(C, ctl, C.shuf[s], &i, &p, &v);
            const bool is_tabu = C.tabu[((int64_t)i * n + p) * q + v] > it;
            valid = !(is_tabu && unc + d >= best);
        }
        const int64_t x = valid ? d : INT64_MAX;
        const int64_t incl = wg_incl_min(x, sh.scan);
        sh.inc[tid] = incl;
        __syncthreads();
        int64_t excl = tid ? sh.inc[tid - 1] : INT64_MAX;
        const int64_t chunk_min = sh.inc[WG - 1];
        __syncthreads();
        const int64_t pm = excl < carry_pm ? excl : carry_pm;
        const bool nb = valid && d < pm;
        const bool tie = valid && d == pm;
        int64_t tot;
        const int64_t T = carry_T + wg_incl_sum(tie ? 1 : 0, sh.scan, &tot);
        int64_t Tb = wg_incl_max(nb ? T : -1, sh.scan);
        if (Tb < carry_Tb) Tb = carry_Tb;
        bool ev = nb;
        if (tie) {
            const uint64_t rv = ring_at(C, r0 + T - 1);
            ev = mod_any(rv, (uint64_t)(1 + T - Tb)) == 0;
        }
        const int64_t wch = wg_max(ev ? s : -1, sh.red);
        if (wch > winner) winner = wch;
        // carries: last thread's values
        if (tid == WG - 1) { sh.inc[0] = T; sh.inc[1] = Tb; }
        __syncthreads();
        carry_T = sh.inc[0];
        carry_Tb = sh.inc[1];
        if (chunk_min < carry_pm) carry_pm = chunk_min;
        __syncthreads();
    }
    if (tid == 0) {
        int64_t draws, dsel;
        int nties;
        int64_t s;
        if (winner < 0) {
            s = (int64_t)mod_any(ring_at(C, r0), (uint64_t)take);
            draws = 1;
            dsel = DSEL_RANDOM;
            nties = 0;
        } else {
            s = winner;
            draws = carry_T;
            nties = (int)(1 + carry_T - carry_Tb);
            int i_, p_, v_;
            dsel = cand_delta(C, ctl, C.shuf[s], &i_, &p_, &v_);
        }
        const uint64_t rs = ring_at(C, r0 + draws);
        draws += 1;
        ctl->ruse = r0 + draws;
        ctl->rlow = ctl->ruse;
        int bi, bp, bv;
        cand_delta(C, ctl, C.shuf[s], &bi, &bp, &bv);
        const int oldv = C.code[(int64_t)bi * n + bp];
        ctl->bi = bi; ctl->bp = bp; ctl->bv = bv; ctl->oldv = oldv;
        ctl->dsel = dsel;
        ctl->nties = nties;
        ctl->stamp = (int32_t)(it + ctl->tenure + (int64_t)(rs % 4));
        ctl->base = C.cidx[bi];
        ctl->bpk = C.cpk[bi];
        ctl->dv = (int64_t)(bv - oldv) * C.pw[bp];
        ctl->ncw = ctl->nq;
        ctl->selected = 1;
    }
    }
    __syncthreads();
}

// After the step's commit (one workgroup, before the trace): the evaluator's
// accumulators start the next step at zero, and the distinct codewords among
// the first `take` candidates (fewer than nq only when --cand cuts the
// sample).  Here rather than in the selection, which every commit workgroup
// computes from the accumulators.  All threads; ends without a barrier (its
// one control-block write is thread 0's).
__device__ __forceinline__ void wg_sel_after(const Cell& C, Ctl* ctl, WgRed& red)
{
    const int tid = threadIdx.x;
    const int64_t take = ctl->take;
    for (int64_t x = tid; x < (int64_t)ctl->nq * ACCW; x += WG) C.acc[x] = 0;
    if (ctl->wnow)
        for (int64_t x = tid; x < (int64_t)ctl->nq * WIDE_W; x += WG) C.accw[x] = 0;
    if (take < ctl->ncand) {
        int64_t cntd = 0;
        for (int64_t s0 = 0; s0 < take; s0 += WG) {
            const int64_t s = s0 + tid;
            int first = 0;
            if (s < take) {
                const uint32_t slot = C.shuf[s] / (uint32_t)ctl->per;
                first = 1;
                for (int64_t s2 = 0; s2 < s; s2++)
                    if (C.shuf[s2] / (uint32_t)ctl->per == slot) { first = 0; break; }
            }
            cntd += first;
        }
        cntd = wg_sum(cntd, red);
        if (tid == 0) ctl->ncw = (int32_t)cntd;
    }
}

// Arrival of the nblk workgroups of one kernel, in two levels, so that the
// arrivals spread over ARR_G lines instead of queueing on one: workgroup bid
// releases onto sub-counter bid % G (G = min(nblk, ARR_G)); the
// workgroup that completes a sub-counter acquires it, clears it for the next
// launch and releases onto `top`; the one that completes `top` (G arrivals)
// is the last, acquires, and returns true.  Every workgroup's writes before
// the call happen before the last one's reads after it.  All threads.
__device__ __forceinline__ bool wg_arrive_last(int32_t* sub, int32_t* top, int* flag, int bid, int nblk)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const int G = nblk < ARR_G ? nblk : ARR_G;
        const int g = bid % G;
        const int gsize = nblk / G + (g < nblk % G ? 1 : 0);
        int last = 0;
        int32_t* sg = sub + (int64_t)g * ARR_STRIDE;
        const int prev = __hip_atomic_fetch_add(sg, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == gsize - 1) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(sg, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int pt = __hip_atomic_fetch_add(top, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            last = (pt == G - 1);
        }
        *flag = last;
    }
    __syncthreads();
    const bool last = *flag != 0;
    if (last) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __syncthreads();
    }
    return last;
}

// --------------------------------------------------------------- kernels
// The evaluator of one cell on nblk workgroups; this one is number bid.
__device__ __forceinline__ void eval_body(const Cell& C, int bid, int nblk)
{
    // the control block is read from an LDS copy, made with one round trip
    // for all its fields (the units' and the shuffle's inputs alike), and
    // written in global memory; the phase timers are added without reading
    Ctl* ctl = C.ctl;
    __shared__ Ctl lc;
    __shared__ ShufSh shs;
    __shared__ WalkTab wtab[WAVES];
    const int64_t th0 = clk();
    desc_touch(C);
    // this workgroup's slice of the kept-count changes since the last
    // evaluator, loaded with the control block and folded into th (their
    // rows are not being recounted: post cleared those changes)
    const int64_t nth = (int64_t)C.M * THW;
    const int64_t xf = (int64_t)bid * WG + threadIdx.x, sf = (int64_t)nblk * WG;
    int32_t dv0 = 0;
    if (xf < nth) dv0 = C.dth[xf];
    // workgroup 0: the slot records its staging reads, with the control block
    __shared__ StageSh sts;
    StageReg srg{0, 0};
    if (bid == 0) srg = stage_load(C);
    ctl_to_lds(&lc, ctl);
    if (!lc.begun || lc.done || lc.selected) return;
    if (bid == 0) stage_to_lds(srg, sts);   // read after the shuffle's barriers
    if (lc.th_on) {
        if (dv0) { atomicAdd(&C.th[xf], dv0); C.dth[xf] = 0; }
        for (int64_t x = xf + sf; x < nth; x += sf) {
            const int32_t v = C.dth[x];
            if (v) { atomicAdd(&C.th[x], v); C.dth[x] = 0; }
        }
    }
    const int wv = threadIdx.x >> 6;
    if (bid == 0 && threadIdx.x == 0 && lc.t_cme) tph_add(ctl, 10, th0 - lc.t_cme);
    // workgroup 0: the shuffle.  The rng refill, keeping ~4096 draws ahead:
    // on three or more groups a chunk of RNG_CHUNK draws in parts by the
    // waves of workgroups 1 .. nfw (when begin decided on one), else wave 0
    // of workgroup 1 (or of 0 on a one-group grid, after the shuffle), one
    // draw after another.  The other workgroups run the units.
    const int nfw = nblk >= 3 ? (nblk - 2 < RNG_WGS ? nblk - 2 : RNG_WGS) : 0;
    const int rng_wg = nblk > 1 ? 1 : 0;
    if (bid == 0) {
        const uint32_t mine = wg_shuffle(C, &lc, shs);
        wg_stage(C, &lc, mine, sts);
        if (threadIdx.x == 0) tph_add(ctl, 5, clk() - th0);
    }
    if (nfw > 0 && bid >= 1 && bid <= nfw) {
        if (lc.fill_on) {
            ring_fill_parts(C, &lc, ctl, (bid - 1) * WAVES + wv, nfw * WAVES);
        } else if (bid == 1 && wv == 0) {
            if (lc.rgen < lc.ruse + 4096) wave_ring_fill(C, ctl, lc.ruse + 4096, lc.budget);
        }
        if (bid == 1 && threadIdx.x == 0) tph_add(ctl, 6, clk() - th0);
    } else if (nfw == 0 && bid == rng_wg && wv == 0) {
        if (lc.rgen < lc.ruse + 4096) wave_ring_fill(C, ctl, lc.ruse + 4096, lc.budget);
        if (threadIdx.x == 0) tph_add(ctl, 6, clk() - th0);
    }
    if (nfw > 0) {
        // units over the waves of workgroups past the filling ones
        // (wave-major numbering: the units go to wave 0 of every workgroup
        // first, spreading them over the CUs: a unit's scattered loads are
        // bound by its CU's rate of cache-line requests)
        if (bid > nfw)
            eval_units(C, &lc, (int64_t)(bid - 1 - nfw) + (int64_t)wv * (nblk - 1 - nfw),
                       (int64_t)(nblk - 1 - nfw) * WAVES, wtab[wv]);
    } else if (bid != 0) {
        // units over all waves of workgroup 1, except the rng wave
        const int64_t wid = (int64_t)(bid - 1) * WAVES + wv - 1;
        const int64_t nw = (int64_t)(nblk - 1) * WAVES - 1;
        if (!(bid == 1 && wv == 0)) eval_units(C, &lc, wid, nw, wtab[wv]);
    } else if (nblk == 1) {
        eval_units(C, &lc, wv, WAVES, wtab[wv]);
    }
}

__global__ void __launch_bounds__(WG, 4) k_eval(const Cell C)
{
    eval_body(C, blockIdx.x, gridDim.x);
}

// kick: teleport `kick` codewords near uncovered words (:1553-1578).  All threads.
__device__ __forceinline__ void wg_ball(const Cell& C, Ctl* ctl, int i, int delta, NarrowSh& sh)
{
    const int64_t base = C.cidx[i];
    const uint64_t pk = C.cpk[i];
    const bool list_on = ctl->ul_valid != 0;
    const int tr = ctl->th_on ? i : -1;
    int64_t d = 0;
    for (int64_t g = threadIdx.x; g < C.V; g += WG) {
        uint64_t wpk;
        const int64_t w = ball_word_at(C, base, pk, g, &wpk);
        d += ball_word(C, ctl, w, wpk, delta, list_on, tr);
    }
    d = wg_sum(d, sh.red);
    if (threadIdx.x == 0) ctl->unc += d;
    __syncthreads();
}

// codeword i moved to word idx, digits pk: its leaving counts are recounted
// by the next evaluator (thread 0; the list holds the move and the kicks,
// STALE_MAX at most; a codeword moved twice keeps its last place)
__device__ __forceinline__ void stale_add(Ctl* ctl, int i, int64_t idx, uint64_t pk)
{
    int k = 0;
    while (k < ctl->nstale && ctl->stale[k] != i) k++;
    if (k == ctl->nstale) {
        if (k == STALE_MAX) return;
        ctl->nstale = k + 1;
    }
    ctl->stale[k] = i;
    ctl->stale_idx[k] = idx;
    ctl->stale_pk[k] = pk;
}

// codeword i changed: log it for the next best copy (thread 0)
__device__ __forceinline__ void log_dirty(const Cell& C, Ctl* ctl, int i)
{
    const int32_t k = ctl->ndirty;
    if (k < DIRTY_CAP) C.dirty[k] = i;
    ctl->ndirty = k + 1;
}

__device__ __forceinline__ void wg_kick(const Cell& C, Ctl* ctl, NarrowSh& sh)
{
    const int n = C.n, q = C.q;
    for (int t = 0; t < ctl->kick; t++) {
        wg_ensure(C, ctl, 1);
        if (threadIdx.x == 0) {
            sh.v32[1] = (int)(ring_at(C, ctl->ruse) % (uint64_t)C.M);
            ctl->ruse += 1;
        }
        __syncthreads();
        const int i = sh.v32[1];
        __syncthreads();
        wg_ball(C, ctl, i, -1, sh);
        const int64_t w = wg_pick(C, ctl, sh);
        wg_ensure(C, ctl, 2 * n + 2);
        if (threadIdx.x == 0) {
            uint64_t pk;
            int64_t pos = ctl->ruse;
            if (w >= 0) {
                pk = pack_index(C, w);
                for (int t2 = 0; t2 < C.R; t2++) {
                    // gcc evaluates the assigned value's draw before the index's
                    const int val = (int)(ring_at(C, pos) % (uint64_t)q);
                    const int at = (int)(ring_at(C, pos + 1) % (uint64_t)n);
                    pos += 2;
                    pk = nib_put(pk, at, val);
                }
            } else {
                pk = 0;
                for (int j = 0; j < n; j++) pk = nib_put(pk, j, (int)(ring_at(C, pos++) % (uint64_t)q));
            }
            ctl->ruse = pos;
            int64_t idx = 0;
            for (int j = 0; j < n; j++) {
                const int v = nib_get(pk, j);
                C.code[(int64_t)i * n + j] = (uint8_t)v;
                idx += (int64_t)v * C.pw[j];
            }
            C.cpk[i] = pk;
            C.cidx[i] = idx;
            log_dirty(C, ctl, i);
            if (ctl->th_on) stale_add(ctl, i, idx, pk);
        }
        __syncthreads();
        wg_ball(C, ctl, i, +1, sh);
    }
    for (int64_t x = threadIdx.x; x < (int64_t)C.M * n * q; x += WG) C.tabu[x] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        ctl->nkicks += 1;
        ctl->since = 0;
    }
    __syncthreads();
}

// apply the committed move and close the step.  All threads of one WG.
__device__ __forceinline__ void wg_post(const Cell& C, Ctl* ctl, NarrowSh& sh)
{
    const int64_t tq0 = clk();
    const int n = C.n, q = C.q;
    if (threadIdx.x == 0) {
        const int64_t dunc = ctl->dunc;
        ctl->dunc = 0;
        ctl->unc += dunc;
        const int bi = ctl->bi, bp = ctl->bp, bv = ctl->bv, oldv = ctl->oldv;
        // (bpk and base are codeword bi's digits and word index before the move)
        const uint64_t npk = nib_put(ctl->bpk, bp, bv);
        const int64_t nidx = ctl->base + ctl->dv;
        C.code[(int64_t)bi * n + bp] = (uint8_t)bv;
        C.cpk[bi] = npk;
        C.cidx[bi] = nidx;
        C.tabu[((int64_t)bi * n + bp) * q + oldv] = ctl->stamp;
        log_dirty(C, ctl, bi);
        ctl->nstale = 0;
        if (ctl->th_on) stale_add(ctl, bi, nidx, npk);
        if (ctl->ul_ovf) { ctl->ul_valid = 0; ctl->ul_ovf = 0; ctl->ul_n = 0; }
        ctl->commit_bytes += (ctl->th_on ? 16 : 8) * (int64_t)C.E * C.P;
        sh.v64[3] = dunc;
        sh.v32[2] = 0;
        if (ctl->unc < ctl->best_unc) {
            ctl->best_unc = ctl->unc;
            ctl->since = 0;
            sh.v32[2] = 1;
            if (ctl->best_unc == 0) ctl->done = 1;
            if (ctl->target >= 0 && ctl->ttt_it < 0 && ctl->best_unc <= ctl->target) {
                ctl->ttt_it = ctl->it;
                ctl->t_ttt = clk();
            }
        } else if (++ctl->since > ctl->stall) {
            sh.v32[2] = 2;
        }
    }
    __syncthreads();
    const int what = sh.v32[2];
    const int64_t dunc = sh.v64[3];
    __syncthreads();
    if (what == 1) {
        // best := code: the codewords changed since the last copy, or all of
        // them when more changed than the log holds
        const int32_t nd = ctl->ndirty;
        if (nd <= DIRTY_CAP) {
            for (int32_t x = threadIdx.x; x < nd * n; x += WG) {
                const int32_t e = x / n;
                const int64_t at = (int64_t)C.dirty[e] * n + (x - e * n);
                C.best[at] = C.code[at];
            }
        } else {
            const int64_t nb = (int64_t)C.M * n, n16 = nb / 16;
            const uint4* src = reinterpret_cast<const uint4*>(C.code);
            uint4* dst = reinterpret_cast<uint4*>(C.best);
            for (int64_t x = threadIdx.x; x < n16; x += WG) dst[x] = src[x];
            for (int64_t x = n16 * 16 + threadIdx.x; x < nb; x += WG) C.best[x] = C.code[x];
        }
        __syncthreads();
        if (threadIdx.x == 0) ctl->ndirty = 0;
    } else if (what == 2) {
        const int64_t tk0 = clk();
        wg_kick(C, ctl, sh);
        if (threadIdx.x == 0) ctl->tph[7] += clk() - tk0;
    }
    // the moved codewords' leaving counts restart at zero, after every update
    // of this step has been applied to them
    if (ctl->th_on) {
        __syncthreads();
        const int ns = ctl->nstale;
        for (int x = threadIdx.x; x < ns * THW; x += WG) {
            const int64_t at = (int64_t)ctl->stale[x / THW] * THW + x % THW;
            C.th[at] = 0;
            C.dth[at] = 0;
        }
    }
    if (threadIdx.x == 0) {
        if (ctl->trace && ctl->tpos < ctl->tcap) {
            int64_t* tr = ctl->trace + ctl->tpos * TRW;
            tr[0] = ctl->it;
            tr[1] = ctl->u;
            tr[2] = ctl->ncand;
            tr[3] = ctl->ncw;
            tr[4] = ctl->bi;
            tr[5] = ctl->bp;
            tr[6] = ctl->bv;
            tr[7] = ctl->oldv;
            tr[8] = ctl->dsel;
            tr[9] = dunc;
            tr[10] = ctl->unc;
            tr[11] = ctl->best_unc;
            tr[12] = what == 2;
            tr[13] = ctl->ruse;
            tr[14] = ctl->since;
            tr[15] = ctl->nties;
        }
        ctl->tpos += 1;
        ctl->begun = 0;
        ctl->selected = 0;
        ctl->tph[1] += clk() - tq0;
    }
    __syncthreads();
}

// offsets of the commit's words around the moving codeword: T[j * 16 + k]
// = (v - c_j) q^(n-1-j) for the k-th value v != c_j of position j
__host__ __device__ __forceinline__ int64_t commit_offset(const Cell& C, uint64_t pk, int j, int k)
{
    const int c = nib_get(pk, j);
    return (int64_t)(k + (k >= c) - c) * C.pw[j];
}
// The powers q^(n-1-j) in LDS (spw), copied at a kernel's start by threads
// j < MAXN, before a barrier: their reads by position index are then LDS
// reads, where an index per lane into the descriptor would be a vector memory
// load issued only when needed.
__device__ __forceinline__ void pw_to_lds(const Cell& C, int64_t* spw)
{
    if (threadIdx.x < MAXN) spw[threadIdx.x] = C.pw[threadIdx.x];
}
// (all threads; spw as pw_to_lds leaves it, after a barrier)
__device__ __forceinline__ void commit_table(const Cell& C, const int64_t* spw, uint64_t pk, int64_t* T)
{
    for (int x = threadIdx.x; x < C.n * C.q1; x += WG) {
        const int j = (int)div_q1(C, (uint32_t)x), k = x - j * C.q1;
        const int c = nib_get(pk, j);
        T[j * 16 + k] = (int64_t)(k + (k >= c) - c) * spw[j];
    }
    __syncthreads();
}
// the word of pattern t on subset sub around base (pattern_word's word,
// through the table)
__host__ __device__ __forceinline__ int64_t commit_word(const Cell& C, const int64_t* T, int64_t base, uint32_t sub,
                                                       uint32_t t)
{
    int64_t a = base;
    for (int i = C.R - 1; i >= 0; i--) {
        const uint32_t tq = div_q1(C, t);
        const int k = (int)(t - tq * (uint32_t)C.q1);
        t = tq;
        a += T[((sub >> (4 * i)) & 15) * 16 + k];
    }
    return a;
}

// The subsets avoiding position bp, in their table order, are those avoiding
// the last position (the subsets of 0 .. n-2) with every position >= bp moved
// up by one: the table builds both from the same index combinations.  So a
// unit's subset needs bp only after its load (sub_base: the subset of unit g
// avoiding n-1, 0 past the end), and the load can precede the selection.
__device__ __forceinline__ uint32_t sub_base(const Cell& C, int64_t g)
{
    if (g >= (int64_t)C.E * C.P) return 0u;
    int64_t t64;
    const int64_t e = div_small(g, C.P, &t64);
    return C.psub[(int64_t)(C.n - 1) * C.E + e];
}
__host__ __device__ __forceinline__ uint32_t sub_avoiding(uint32_t base, int R, int bp)
{
    uint32_t out = 0;
    for (int i = 0; i < R; i++) {
        const uint32_t v = (base >> (4 * i)) & 15u;
        out |= (v + (v >= (uint32_t)bp ? 1u : 0u)) << (4 * i);
    }
    return out;
}

// Runs of contiguous lanes with equal keys (key < 0: no operation) combine
// their values; the last lane of each run gets the run's total (segmented
// inclusive scan).  Neighbouring patterns of a sphere walk are neighbouring
// words, so a wave's plane-word and block updates come in long runs, and one
// atomic per run replaces one per lane where a wave has many transitions.
// All 64 lanes.
__device__ __forceinline__ bool run_tail(int64_t key, bool* head)
{
    const int lane = threadIdx.x & 63;
    const int64_t kp = __shfl_up(key, 1, 64);
    const int64_t kn = __shfl_down(key, 1, 64);
    *head = lane == 0 || kp != key;
    return lane == 63 || kn != key;
}
template <int K>
__device__ __forceinline__ void run_or(bool head, uint32_t (&v)[K])
{
    const int lane = threadIdx.x & 63;
    int f = head ? 1 : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y[K];
#pragma unroll
        for (int k = 0; k < K; k++) y[k] = __shfl_up(v[k], o, 64);
        const int fy = __shfl_up(f, o, 64);
        if (lane >= o && !f) {
#pragma unroll
            for (int k = 0; k < K; k++) v[k] |= y[k];
        }
        if (lane >= o) f |= fy;
    }
}
__device__ __forceinline__ int run_add(bool head, int v)
{
    const int lane = threadIdx.x & 63;
    int f = head ? 1 : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(v, o, 64);
        const int fy = __shfl_up(f, o, 64);
        if (lane >= o && !f) v += y;
        if (lane >= o) f |= fy;
    }
    return v;
}

// the commit units: every pattern of the subsets avoiding bp, flattened; a
// wave takes 64 consecutive patterns per round (all lanes take part in every
// round, those past the end without an operation).  sub0: sub_base of the
// thread's first unit (gt), loaded by the caller.
__device__ __forceinline__ int64_t commit_units(const Cell& C, const Ctl* ctl, int64_t gt, int64_t nt,
                                                const int64_t* T, bool th, uint32_t sub0)
{
    const int lane = threadIdx.x & 63;
    const int bi = ctl->bi, bp = ctl->bp, bv = ctl->bv;
    const int64_t base = ctl->base, dv = ctl->dv;
    const uint64_t pk = ctl->bpk;
    const bool list_on = ctl->ul_valid != 0;
    const int64_t S = (int64_t)C.E * C.P;
    int64_t d = 0;
    for (int64_t g0 = gt - lane; g0 < S; g0 += nt) {
        const int64_t g = g0 + lane;
        const bool act = g < S;
        int64_t a = -1, b = -1;
        uint32_t za = 0, oa_set = 0, oa_clr = 0, zb_clr = 0, ob_set = 0, ob_clr = 0;
        int da = 0, db = 0;
        if (act) {
            int64_t t64;
            const int64_t e = div_small(g, C.P, &t64);
            const uint32_t t = (uint32_t)t64;
            const uint32_t sb = g == gt ? sub0 : C.psub[(int64_t)(C.n - 1) * C.E + e];
            const uint32_t sub = sub_avoiding(sb, C.R, bp);
            a = commit_word(C, T, base, sub, t);
            b = a + dv;
            // the counts and (when kept) the coverer XORs read together,
            // before any store: the arrays may alias as far as the compiler
            // knows, so a read after a store would wait for its own round trip
            const uint16_t ca = cnt_get(C, a), cb = cnt_get(C, b);
            const uint32_t xa = th ? xor_get(C, a) : 0u, xb = th ? xor_get(C, b) : 0u;
            cnt_put(C, a, (uint16_t)(ca - 1));
            cnt_put(C, b, (uint16_t)(cb + 1));
            if (th) {
                // bi leaves a's coverers and joins b's; a word left with one
                // coverer counts for it, b's former only coverer loses b
                xor_put(C, a, xa ^ (uint32_t)bi);
                xor_put(C, b, xb ^ (uint32_t)bi);
                if (ca == 2 || cb == 1) {
                    const uint64_t apk = pattern_digits(C, pk, sub, C.R, t);
                    if (ca == 2) th_event(C, (int)(xa ^ (uint32_t)bi), apk, +1);
                    if (cb == 1) th_event(C, (int)xb, nib_put(apk, bp, bv), -1);
                }
            }
            const uint32_t ma = 1u << (a & 31), mb = 1u << (b & 31);
            if (ca == 1) {
                za = ma; oa_clr = ma; da = 1;
                if (list_on) list_insert(C, C.ctl, a, pattern_digits(C, pk, sub, C.R, t));
                d++;
            } else if (ca == 2) {
                oa_set = ma;
            }
            if (cb == 0) {
                zb_clr = mb; ob_set = mb; db = -1;
                d--;
            } else if (cb == 1) {
                ob_clr = mb;
            }
        }
        // the planes and the uncovered-count blocks.  Each word is one unit's,
        // so the bits are distinct and the operations commute: a wave with few
        // transitions (most words cross no threshold on well covered cells)
        // issues an atomic per transition; one with many (sparsely covered
        // cells, where most words do) first combines runs of lanes on the same
        // plane word or block, one atomic per run
        const bool tr = (za | oa_set | oa_clr | zb_clr | ob_set | ob_clr) != 0;
        if (__popcll(__ballot(tr)) <= ATOM_DIRECT) {
            if (za) atomicOr(&C.zb[a >> 5], za);
            if (oa_set) atomicOr(&C.ob[a >> 5], oa_set);
            if (oa_clr) atomicAnd(&C.ob[a >> 5], ~oa_clr);
            if (zb_clr) atomicAnd(&C.zb[b >> 5], ~zb_clr);
            if (ob_set) atomicOr(&C.ob[b >> 5], ob_set);
            if (ob_clr) atomicAnd(&C.ob[b >> 5], ~ob_clr);
            if (da) atomicAdd(&C.blk[a >> BLK_SHIFT], 1u);
            if (db) atomicSub(&C.blk[b >> BLK_SHIFT], 1u);
            continue;
        }
        // the planes: one atomic per plane word and run
        {
            const int64_t key = (za | oa_set | oa_clr) ? (a >> 5) : -1;
            bool head;
            const bool tail = run_tail(key, &head);
            uint32_t v[3] = {za, oa_set, oa_clr};
            run_or<3>(head, v);
            if (tail && key >= 0) {
                if (v[0]) atomicOr(&C.zb[key], v[0]);
                if (v[1]) atomicOr(&C.ob[key], v[1]);
                if (v[2]) atomicAnd(&C.ob[key], ~v[2]);
            }
        }
        {
            const int64_t key = (zb_clr | ob_set | ob_clr) ? (b >> 5) : -1;
            bool head;
            const bool tail = run_tail(key, &head);
            uint32_t v[3] = {zb_clr, ob_set, ob_clr};
            run_or<3>(head, v);
            if (tail && key >= 0) {
                if (v[0]) atomicAnd(&C.zb[key], ~v[0]);
                if (v[1]) atomicOr(&C.ob[key], v[1]);
                if (v[2]) atomicAnd(&C.ob[key], ~v[2]);
            }
        }
        // the uncovered-count blocks
        {
            const int64_t key = da ? (a >> BLK_SHIFT) : -1;
            bool head;
            const bool tail = run_tail(key, &head);
            const int v = run_add(head, da);
            if (tail && key >= 0 && v) atomicAdd(&C.blk[key], (uint32_t)v);
        }
        {
            const int64_t key = db ? (b >> BLK_SHIFT) : -1;
            bool head;
            const bool tail = run_tail(key, &head);
            const int v = run_add(head, -db);
            if (tail && key >= 0 && v) atomicSub(&C.blk[key], (uint32_t)v);
        }
    }
    return d;
}

// The selection's results, published by commit workgroup 0 into the p_
// fields (the other workgroups may not have read the control block yet, and
// read the cursors, generator state and `selected` to make the same
// selection).  Thread 0.
__device__ __forceinline__ void sel_publish(Ctl* g, const Ctl* l)
{
    for (int k = 0; k < 4; k++) g->p_rst[k] = l->rst[k];
    g->p_rgen = l->rgen;
    g->p_ruse = l->ruse;
    g->p_rlow = l->rlow;
    g->p_bi = l->bi; g->p_bp = l->bp; g->p_bv = l->bv; g->p_oldv = l->oldv;
    g->p_dsel = l->dsel;
    g->p_nties = l->nties;
    g->p_stamp = l->stamp;
    g->p_base = l->base;
    g->p_bpk = l->bpk;
    g->p_dv = l->dv;
    g->p_ncw = l->ncw;
    g->p_t_eve = l->t_eve;
    g->p_ok = 1;
}

// The post kernel's start: the published selection into the control block
// (its LDS copy).  Thread 0.
__device__ __forceinline__ void sel_apply(Ctl* l)
{
    for (int k = 0; k < 4; k++) l->rst[k] = l->p_rst[k];
    l->rgen = l->p_rgen;
    l->ruse = l->p_ruse;
    l->rlow = l->p_rlow;
    l->bi = l->p_bi; l->bp = l->p_bp; l->bv = l->p_bv; l->oldv = l->p_oldv;
    l->dsel = l->p_dsel;
    l->nties = l->p_nties;
    l->stamp = l->p_stamp;
    l->base = l->p_base;
    l->bpk = l->p_bpk;
    l->dv = l->p_dv;
    l->ncw = l->p_ncw;
    l->t_eve = l->p_t_eve;
    l->nstale = 0;          // the evaluator has recounted them
    l->selected = 1;
    l->p_ok = 0;
}

// The selection and commit of a search step, on nblk workgroups; this one is
// number bid.  Every workgroup makes the selection itself, from the same
// evaluator results and ring draws, so it needs no kernel of its own and no
// workgroup waits for another; workgroup 0 publishes it.  Then the pattern
// updates.  The step's bookkeeping and the next begin run in post_body,
// launched after it, so these workgroups end without a device-scope release
// of their own.
__device__ __forceinline__ void commit_body(const Cell& C, int bid, int nblk)
{
    // the control block read from an LDS copy (one round trip), written in
    // global memory (the uncovered delta, the list's inserts, the timers)
    Ctl* ctl = C.ctl;
    __shared__ Ctl lc;
    __shared__ SelSh sel;
    const int64_t tc0 = clk();
    desc_touch(C);
    // wave 0: the one-wave selection's loads, issued before (and arriving
    // with) the control block's
    SelPre pre;
    if (threadIdx.x < 64) sel_preload(C, threadIdx.x, pre);
    const int64_t gt = (int64_t)bid * WG + threadIdx.x;
    const uint32_t sub0 = sub_base(C, gt);
    __shared__ int64_t spw[MAXN];
    pw_to_lds(C, spw);
    ctl_to_lds(&lc, ctl);
    if (!lc.begun || lc.done || lc.selected) return;
    wg_select(C, &lc, sel, pre);
    __syncthreads();
    const int64_t te = clk();
    __shared__ int64_t ctab[MAXN * 16];
    commit_table(C, spw, lc.bpk, ctab);
    const int64_t d = commit_units(C, &lc, gt, (int64_t)nblk * WG, ctab, lc.th_on != 0, sub0);
    // the selection's publication (the post kernel reads it), here rather than
    // before the table's barrier, which would wait for its stores
    if (bid == 0 && threadIdx.x == 0) {
        if (lc.t_cme) tph_add(ctl, 8, tc0 - lc.t_cme);
        tph_add(ctl, 0, te - tc0);
        lc.t_eve = te;
        sel_publish(ctl, &lc);
    }
    // the uncovered delta, summed per wave with no workgroup barrier: most
    // waves have none
    if (__ballot(d != 0)) {
        const int64_t ws = wave_sum64(d);
        if ((threadIdx.x & 63) == 0 && ws)
            atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->dunc), (unsigned long long)ws);
    }
}

// After the commit, one workgroup: apply the move, stamp tabu, track the best
// code, kick, trace, and begin the next step.
__device__ __forceinline__ void post_body(const Cell& C)
{
    __shared__ NarrowSh sh;
    __shared__ Ctl lc;
    const int64_t tp0 = clk();
    desc_touch(C);
    __shared__ int64_t spw[MAXN];
    pw_to_lds(C, spw);
    ctl_to_lds(&lc, C.ctl);
    Ctl* L = &lc;
    if (!L->begun || L->done || !L->p_ok) return;
    // (no barrier between these: sel_apply writes only fields that
    // wg_sel_after does not read, and thread 0, which wrote them, is the one
    // that reads them first in wg_post)
    if (threadIdx.x == 0) {
        sel_apply(L);
        if (L->t_eve) L->tph[9] += tp0 - L->t_eve;
    }
    wg_sel_after(C, L, sh.red);
    wg_post(C, L, sh);
    wg_begin(C, L, sh, spw);
    if (threadIdx.x == 0) L->t_cme = clk();
    ctl_from_lds(C.ctl, &lc);
}

// A single move requested through the C ABI (k_set_move first): the pattern
// updates, then the last workgroup to arrive applies the move only.
__device__ __forceinline__ void commit_api_body(const Cell& C, int bid, int nblk)
{
    Ctl* ctl = C.ctl;
    __shared__ WgRed red;
    __shared__ int last;
    if (!ctl->begun || ctl->done || !ctl->selected) return;
    const int64_t gt = (int64_t)bid * WG + threadIdx.x;
    __shared__ int64_t ctab[MAXN * 16];
    __shared__ int64_t spw[MAXN];
    pw_to_lds(C, spw);
    __syncthreads();
    commit_table(C, spw, C.cpk[ctl->bi], ctab);
    const int64_t d = commit_units(C, ctl, gt, (int64_t)nblk * WG, ctab, false, sub_base(C, gt));
    const int64_t ds = wg_sum(d, red);
    if (threadIdx.x == 0 && ds) atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->dunc), (unsigned long long)ds);
    if (!wg_arrive_last(C.arrive, &ctl->cmdone, &last, bid, nblk)) return;
    __shared__ Ctl lc;
    ctl_to_lds(&lc, ctl);
    Ctl* L = &lc;
    if (threadIdx.x == 0) {
        L->cmdone = 0;
        const int64_t dunc = L->dunc;
        L->dunc = 0;
        L->unc += dunc;
        L->dsel = dunc;
        C.code[(int64_t)L->bi * C.n + L->bp] = (uint8_t)L->bv;
        C.cpk[L->bi] = nib_put(C.cpk[L->bi], L->bp, L->bv);
        C.cidx[L->bi] += L->dv;
        log_dirty(C, L, L->bi);
        if (L->ul_ovf) { L->ul_valid = 0; L->ul_ovf = 0; L->ul_n = 0; }
        L->begun = 0;
        L->selected = 0;
    }
    ctl_from_lds(C.ctl, &lc);
}

__global__ void __launch_bounds__(WG) k_commit(const Cell C)
{
    commit_body(C, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(WG) k_post(const Cell C)
{
    post_body(C);
}

__global__ void __launch_bounds__(WG) k_commit_api(const Cell C)
{
    commit_api_body(C, blockIdx.x, gridDim.x);
}

// set the run's stop and begin the first step (one workgroup)
// start a run of nsteps; notrace: the run records no trace (batched runs)
__device__ __forceinline__ void begin_run_body(const Cell& C, int64_t nsteps, bool notrace)
{
    __shared__ NarrowSh sh;
    __shared__ Ctl lc;
    __shared__ int64_t spw[MAXN];
    pw_to_lds(C, spw);
    ctl_to_lds(&lc, C.ctl);
    Ctl* ctl = &lc;
    if (threadIdx.x == 0) {
        ctl->it_stop = ctl->it + nsteps - (ctl->begun ? 1 : 0);
        ctl->tpos = 0;
        if (notrace) {
            ctl->trace = nullptr;
            ctl->tcap = 0;
        }
    }
    __syncthreads();
    wg_begin(C, ctl, sh, spw);
    if (threadIdx.x == 0) ctl->t_cme = clk();
    ctl_from_lds(C.ctl, &lc);
}

__global__ void __launch_bounds__(WG) k_begin_run(const Cell C, int64_t nsteps)
{
    begin_run_body(C, nsteps, false);
}

// --------------------------------------------------------- init kernels
// uniform random digits for codewords have..M-1, row-major (:1235-1238):
// digit k is rmod(q) of the run's draw k.  The control block is fresh (no
// draw generated or used), so the draws skip the ring: chunks of RNG_CHUNK
// draws, wave w making the chunk's part w (RNG_PART draws) from the state
// J_w s (as ring_fill_parts); the wave that makes the chunk's last draw keeps
// the state after it.  Afterwards rgen = ruse = rlow = the draws used.  One
// workgroup of RNG_SPLIT waves.
static_assert(RNG_PART == 64, "k_init_random: lane k keeps draw k of a part");
__global__ void __launch_bounds__(RNG_SPLIT * 64) k_init_random(const Cell C, int have)
{
    Ctl* ctl = C.ctl;
    __shared__ uint64_t ss[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t need = (int64_t)(C.M - have) * C.n;
    uint8_t* out = C.code + (int64_t)have * C.n;
    if (threadIdx.x < 4) ss[threadIdx.x] = ctl->rst[threadIdx.x];
    __syncthreads();
    for (int64_t done = 0; done < need; done += RNG_CHUNK) {
        const uint64_t s0[4] = {uniform64(ss[0]), uniform64(ss[1]), uniform64(ss[2]), uniform64(ss[3])};
        __syncthreads();
        const int64_t gb = done + (int64_t)RNG_PART * w;
        if (gb < need) {
            uint64_t st[4] = {s0[0], s0[1], s0[2], s0[3]};
            if (w > 0) {
                const uint64_t* Mj = C.jump + (size_t)(w - 1) * 256 * 4;
                uint64_t o[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint64_t* row = Mj + (size_t)(64 * k + lane) * 4;
                    const uint64_t x = (row[0] & s0[0]) ^ (row[1] & s0[1]) ^ (row[2] & s0[2]) ^ (row[3] & s0[3]);
                    o[k] = __ballot(__popcll(x) & 1);
                }
#pragma unroll
                for (int k = 0; k < 4; k++) st[k] = uniform64(o[k]);
            }
            const int nb = (need - gb < RNG_PART) ? (int)(need - gb) : RNG_PART;
            uint64_t mine = 0;
            for (int k = 0; k < nb; k++) {
                const uint64_t r = xo_next(st);
                if (lane == k) mine = r;
            }
            if (lane < nb) out[gb + lane] = (uint8_t)(mine % (uint64_t)C.q);
            if (lane == 0 && (gb + nb == need || w == RNG_SPLIT - 1)) {
                ss[0] = st[0]; ss[1] = st[1]; ss[2] = st[2]; ss[3] = st[3];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ctl->rst[0] = ss[0]; ctl->rst[1] = ss[1]; ctl->rst[2] = ss[2]; ctl->rst[3] = ss[3];
        ctl->rgen = need;
        ctl->ruse = need;
        ctl->rlow = need;
    }
}

// digits -> packed digits and word index for every codeword
__global__ void __launch_bounds__(WG) k_pack_code(const Cell C)
{
    for (int i = blockIdx.x * WG + threadIdx.x; i < C.M; i += gridDim.x * WG) {
        uint64_t pk = 0;
        int64_t idx = 0;
        for (int j = 0; j < C.n; j++) {
            const int v = C.code[(int64_t)i * C.n + j];
            pk |= (uint64_t)v << (4 * j);
            idx += (int64_t)v * C.pw[j];
        }
        C.cpk[i] = pk;
        C.cidx[i] = idx;
    }
}

// greedy placement of codeword i on the picked uncovered word (:1219-1233)
__global__ void __launch_bounds__(WG) k_greedy_pick(const Cell C, int i)
{
    __shared__ NarrowSh sh;
    __shared__ Ctl lc;
    ctl_to_lds(&lc, C.ctl);
    Ctl* ctl = &lc;
    int64_t w = wg_pick(C, ctl, sh);
    if (w < 0) wg_ensure(C, ctl, 1);
    if (threadIdx.x == 0) {
        if (w < 0) {
            w = (int64_t)(ring_at(C, ctl->ruse) % (uint64_t)C.N);
            ctl->ruse += 1;
        }
        const uint64_t pk = pack_index(C, w);
        for (int j = 0; j < C.n; j++) C.code[(int64_t)i * C.n + j] = (uint8_t)nib_get(pk, j);
        C.cpk[i] = pk;
        C.cidx[i] = w;
    }
    ctl_from_lds(C.ctl, &lc);
}

// +-1 ball of codeword i (incremental, whole grid)
__global__ void __launch_bounds__(WG) k_ball_cw(const Cell C, int i, int delta)
{
    Ctl* ctl = C.ctl;
    __shared__ WgRed red;
    const bool list_on = ctl->ul_valid != 0;
    const int64_t base = C.cidx[i];
    const uint64_t pk = C.cpk[i];
    int64_t d = 0;
    for (int64_t g = (int64_t)blockIdx.x * WG + threadIdx.x; g < C.V; g += (int64_t)gridDim.x * WG) {
        uint64_t wpk;
        const int64_t w = ball_word_at(C, base, pk, g, &wpk);
        d += ball_word(C, ctl, w, wpk, delta, list_on);
    }
    const int64_t s = wg_sum(d, red);
    if (threadIdx.x == 0 && s) atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->unc), (unsigned long long)s);
}

// words of each given centre's ball with cnt == 1 (remove-and-repair, :927-946);
// a negative centre has left the list and scores none
__global__ void __launch_bounds__(WG) k_unique_cover(const Cell C, const int64_t* centers,
                                                     int k, int64_t* out)
{
    __shared__ WgRed red;
    for (int c = blockIdx.x; c < k; c += gridDim.x) {
        const int64_t base = centers[c];
        if (base < 0) continue;
        const uint64_t pk = pack_index(C, base);
        int64_t u = 0;
        for (int64_t g = threadIdx.x; g < C.V; g += WG)
            u += bit_of(C.ob, ball_word_at(C, base, pk, g, nullptr));
        u = wg_sum(u, red);
        if (threadIdx.x == 0) out[c] = u;
    }
}

// the descent's choice (:1177-1189): the first listed centre with the fewest
// uniquely covered words leaves the list (its entry becomes -1) and its word
// goes to sel[0] for k_ball_drop; -1 there when none is listed.  One
// workgroup; the key (count << 32 | position) makes the first of equal
// counts the smallest.
__global__ void __launch_bounds__(WG) k_rr_pick(int64_t* __restrict__ centers,
                                                const int64_t* __restrict__ uc, int k,
                                                int64_t* __restrict__ sel)
{
    __shared__ WgRed red;
    int64_t key = INT64_MAX;
    for (int c = threadIdx.x; c < k; c += WG)
        if (centers[c] >= 0) {
            const int64_t kc = (uc[c] << 32) | c;
            key = kc < key ? kc : key;
        }
    key = wg_min(key, red);
    if (threadIdx.x == 0) {
        if (key == INT64_MAX) {
            sel[0] = -1;
        } else {
            const int c = (int)(key & 0xffffffff);
            sel[0] = centers[c];
            centers[c] = -1;
        }
    }
}

// whole-grid list rebuild from the zb plane
__global__ void __launch_bounds__(WG) k_list_rebuild(const Cell C)
{
    Ctl* ctl = C.ctl;
    for (int64_t x = (int64_t)blockIdx.x * WG + threadIdx.x; x < C.nzw; x += (int64_t)gridDim.x * WG) {
        const uint32_t z = C.zb[x];
        C.inl[x] = z;
        uint32_t m = z;
        while (m) {
            const int b = __ffs(m) - 1;
            m &= m - 1;
            const int64_t w = x * 32 + b;
            const unsigned long long pos = atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->ul_n), 1ULL);
            if ((int64_t)pos < C.ulcap) { C.ul_idx[pos] = w; C.ul_pk[pos] = pack_index(C, w); }
            else atomicOr(&ctl->ul_ovf, 1);
        }
    }
}

// first uncovered word at or after start (one workgroup)
__global__ void __launch_bounds__(WG) k_first_from(const Cell C, int64_t start, int64_t* out)
{
    __shared__ NarrowSh sh;
    const int64_t w = wg_first_from(C, start, sh);
    if (threadIdx.x == 0) *out = w;
}

// ----------------------------------------- coverer XOR and leaving counts
// xor32 (zeroed first): every codeword's index XORed into the words of its ball
__global__ void __launch_bounds__(WG) k_xor_build(const Cell C)
{
    const int64_t tot = (int64_t)C.M * C.V;
    for (int64_t x = (int64_t)blockIdx.x * WG + threadIdx.x; x < tot; x += (int64_t)gridDim.x * WG) {
        const int i = (int)(x / C.V);
        const int64_t g = x - (int64_t)i * C.V;
        const int64_t w = ball_word_at(C, C.cidx[i], C.cpk[i], g, nullptr);
        atomicXor(&C.xor32[w >> 1], (uint32_t)i << (16 * (uint32_t)(w & 1)));
    }
}

// th (zeroed first): T and H_j of every codeword, one (codeword, R-subset)
// sphere per wave
__global__ void __launch_bounds__(WG) k_th_all(const Cell C)
{
    __shared__ WalkTab wtab[WAVES];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t nu = (int64_t)C.M * C.F;
    for (int64_t u = (int64_t)blockIdx.x * WAVES + wv; u < nu; u += (int64_t)gridDim.x * WAVES) {
        const int i = (int)(u / C.F);
        const int f = (int)(u - (int64_t)i * C.F);
        const uint32_t sub = C.fsub[f];
        const int64_t c1 = sphere_count(C, C.ob, C.cidx[i], C.cpk[i], sub, wtab[wv]);
        if (lane == 0 && c1) {
            int32_t* t = C.th + (int64_t)i * THW;
            atomicAdd(&t[0], (int)c1);
            for (int k = 0; k < C.R; k++) atomicAdd(&t[1 + ((sub >> (4 * k)) & 15)], (int)c1);
        }
    }
}

// the search's use of th from here on: on, no codeword pending a recount
__global__ void k_th_setup(const Cell C, int on)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        C.ctl->th_on = on;
        C.ctl->nstale = 0;
    }
}

// --------------------------------------------- evaluator through the ABI
// codewords slot_cw[0..k) against word u (ctl->upk), all differing positions
__global__ void __launch_bounds__(WG) k_eval_probe(const Cell C, int k)
{
    __shared__ WalkTab wtab[WAVES];
    WalkTab& tb = wtab[threadIdx.x >> 6];
    const Ctl* ctl = C.ctl;
    const uint64_t upk = ctl->upk;
    const int64_t wid = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * WAVES;
    const int64_t nl = (int64_t)k * C.F, ne = (int64_t)k * C.n * C.E;
    for (int64_t g = wid; g < nl + ne; g += nw) {
        if (g < nl) {
            const int slot = (int)(g / C.F);
            unit_leave(C, slot, (int)(g - (int64_t)slot * C.F), tb);
        } else {
            const int64_t h = g - nl;
            const int slot = (int)(h / ((int64_t)C.n * C.E));
            const int64_t r = h - (int64_t)slot * C.n * C.E;
            const int p = (int)(r / C.E);
            const int e = (int)(r - (int64_t)p * C.E);
            const int up = nib_get(upk, p);
            if (nib_get(C.cpk[C.slot_cw[slot]], p) == up) continue;   // uniform per wave
            unit_walk(C, slot, p, up, e, tb);
        }
    }
}

// every move of the codewords slot_cw[0..k): the leaving units as above, and
// per (codeword, position p, value v != c_p, R-subset avoiding p) the words
// with cnt == 0 of that subset's sphere around the moved codeword, counted
// into negw[slot][p][v]
__global__ void __launch_bounds__(WG) k_eval_probe_all(const Cell C, int k, int32_t* negw)
{
    __shared__ WalkTab wtab[WAVES];
    WalkTab& tb = wtab[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int64_t wid = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * WAVES;
    const int64_t per = (int64_t)C.n * C.q1 * C.E;
    const int64_t nl = (int64_t)k * C.F, ne = (int64_t)k * per;
    for (int64_t g = wid; g < nl + ne; g += nw) {
        if (g < nl) {
            const int slot = (int)(g / C.F);
            unit_leave(C, slot, (int)(g - (int64_t)slot * C.F), tb);
            continue;
        }
        const int64_t h = g - nl;
        const int slot = (int)(h / per);
        const int64_t r = h - (int64_t)slot * per;
        const int pv = (int)(r / C.E);
        const int e = (int)(r - (int64_t)pv * C.E);
        const int p = pv / C.q1, t = pv - p * C.q1;
        const int i = C.slot_cw[slot];
        const uint64_t pk0 = C.cpk[i];
        const int cp = nib_get(pk0, p);
        const int v = t + (t >= cp);
        const int64_t base = C.cidx[i] + (int64_t)(v - cp) * C.pw[p];
        const int64_t c0 = sphere_count(C, C.zb, base, nib_put(pk0, p, v), C.psub[(int64_t)p * C.E + e], tb);
        if (lane == 0 && c0) atomicAdd(&negw[(int64_t)slot * WIDE_W + p * MAXQ + v], (int)c0);
    }
}

__global__ void __launch_bounds__(WG) k_eval_finish_all(const Cell C, int k, const int32_t* negw, int64_t* dvec)
{
    const int nq = C.n * C.q;
    for (int x = blockIdx.x * WG + threadIdx.x; x < k * nq; x += gridDim.x * WG) {
        const int slot = x / nq, pv = x - slot * nq, p = pv / C.q, v = pv - p * C.q;
        const int32_t* a = C.acc + (int64_t)slot * ACCW;
        if (nib_get(C.cpk[C.slot_cw[slot]], p) == v) dvec[x] = DSEL_RANDOM;
        else dvec[x] = (int64_t)(a[0] - a[1 + p]) - (int64_t)negw[(int64_t)slot * WIDE_W + p * MAXQ + v];
    }
}

__global__ void __launch_bounds__(WG) k_eval_finish(const Cell C, int k, int64_t* dvec)
{
    const uint64_t upk = C.ctl->upk;
    for (int x = blockIdx.x * WG + threadIdx.x; x < k * C.n; x += gridDim.x * WG) {
        const int slot = x / C.n, p = x - slot * C.n;
        const int32_t* a = C.acc + (int64_t)slot * ACCW;
        if (nib_get(C.cpk[C.slot_cw[slot]], p) == nib_get(upk, p)) dvec[x] = DSEL_RANDOM;
        else dvec[x] = (int64_t)(a[0] - a[1 + p]) - (int64_t)a[1 + MAXN + p];
    }
}

// set up a single move for k_commit_api
__global__ void k_set_move(const Cell C, int i, int p, int v)
{
    Ctl* ctl = C.ctl;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int oldv = C.code[(int64_t)i * C.n + p];
    ctl->bi = i; ctl->bp = p; ctl->bv = v; ctl->oldv = oldv;
    ctl->base = C.cidx[i];
    ctl->bpk = C.cpk[i];
    ctl->dv = (int64_t)(v - oldv) * C.pw[p];
    ctl->begun = 1;
    ctl->selected = 1;
    ctl->done = 0;
}

// Stop conditions of the search loop (one thread).  The deadline is `ticks`
// of the device clock from now (ticks <= 0: none).  A new target restarts the
// time-to-target record; a target the best code already meets is met now
// (the reference's check before its first step, covsearch2e.c:1292-1293), and
// a search stopped at a target it no longer meets may go on.
__global__ void k_set_limits(const Cell C, int64_t target, int64_t ticks)
{
    Ctl* ctl = C.ctl;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t now = clk();
    if (target != ctl->target) {
        ctl->ttt_it = -1;
        ctl->t_ttt = 0;
    }
    ctl->target = target;
    ctl->deadline = ticks > 0 ? now + ticks : 0;
    if (ctl->done == 2 && (target < 0 || ctl->best_unc > target)) ctl->done = 0;
    if (target >= 0 && ctl->ttt_it < 0 && ctl->best_unc <= target) {
        ctl->ttt_it = ctl->it;
        ctl->t_ttt = now;
    }
}

// the device clock, for the host's reading of the time-to-target record
__global__ void k_clock(int64_t* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = clk();
}

}  // namespace cov
