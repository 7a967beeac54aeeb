// cov_common.h -- shared definitions of the MI355X covering-code engine.
//
// Data layout in HBM, per cell (one (q,n,R,M) search state):
//   cnt   uint16[N]   coverage multiplicity, MSB-first base-q word index;
//                     allocated as uint32 pairs so ball updates can use 32-bit
//                     atomics on the containing word (cnt <= M <= 65535 never
//                     carries or borrows across the 16-bit lanes)
//   zb,ob uint32[N/32] bit planes [cnt==0] and [cnt==1]: the evaluator reads
//                     these (1 bit per word) instead of cnt, so a K6(9,4) plane
//                     is 1.26 MB and stays resident in one XCD's 4 MB L2
//   blk   uint32[N/4096] uncovered words per 4096-word block: turns the cyclic
//                     "first uncovered at or after start" scan into a short
//                     walk over block counts
//   code  uint8[M*n], cpk uint64[M] (digits packed 4 bits per position, for
//                     one-xor Hamming distances), cidx int64[M], best uint8[M*n]
//   tabu  int32[M*n*q]
//   ul_*  uncovered-word list (index + packed digits) with a membership plane,
//         used for the entering side of the evaluator once it is short
//   rbuf  uint64[RB] ring of xoshiro256++ draws; decisions consume draws by a
//         cursor, generation runs ahead off the critical path
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cov {

constexpr int MAXN = 16;
constexpr int MAXQ = 16;
constexpr int MAXR = 8;
constexpr int WG = 256;              // threads per workgroup in every kernel
constexpr int WAVES = WG / 64;
constexpr int BLK_SHIFT = 12;        // words per uncovered-count block = 4096
constexpr int BLK_WORDS = 1 << BLK_SHIFT;
constexpr int RB = 1 << 14;          // rng ring size (draws)
constexpr int ACCW = 1 + 2 * MAXN;   // per-slot accumulators: t1, hit[16], neg[16]
constexpr int LIST_CHUNK = 64;       // uncovered-list entries per evaluator unit: one round of a wave, two
                                     // dependent loads (entry, then its plane bit), so no unit is a long chain
constexpr int64_t DSEL_RANDOM = 1LL << 60;
constexpr int TRW = 16;
constexpr int DIRTY_CAP = 4096;      // logged codeword changes before a full best copy
// The leaving side's counts per codeword c, kept up to date instead of
// recomputed each step: th[c*THW] = T(c) = #{w : d(w,c) = R, cnt(w) = 1} and
// th[c*THW + 1 + j] = H_j(c), the same words with w_j != c_j.  A word whose
// cnt crosses 1 changes them for its unique coverer only, which the XOR of
// its coverers' indices names; codewords that move are recounted whole.
constexpr int THW = 1 + MAXN;
constexpr int STALE_MAX = 8;         // codewords recounted per step at most (the move and the kicks)
// Workgroup arrival: up to ARR_G sub-counters per kernel, each on its own
// 256-byte line (atomics to one line are serialised where they are performed)
constexpr int ARR_G = 32;
constexpr int ARR_STRIDE = 64;       // int32 slots between sub-counters
constexpr uint64_t NIB1 = 0x1111111111111111ULL;
constexpr int WIDE_W = MAXN * MAXQ;  // --wide entering counts per slot: [position][value]
constexpr int64_t UL_REF_HALF = (int64_t)1 << 19;   // the reference's list rebuild threshold (ULCAP / 2)

// Device-resident search state.  Scalars written by the narrow phases (one
// workgroup) and read by the wide kernels of the same step.
struct alignas(16) Ctl {
    uint64_t rst[4];        // generator state after rgen draws
    int64_t rgen;           // draws generated into the ring
    int64_t ruse;           // draws consumed
    int64_t rlow;           // oldest draw still to be read this step
    int64_t it, it_stop, since, nkicks, unc, best_unc;
    int64_t stall;
    int64_t dunc;           // commit accumulator (device atomics)
    int64_t ul_n;
    int64_t u;
    uint64_t upk;
    int64_t sh_pos;         // ring position of the first shuffle draw
    int64_t n_leave, n_enter, nchunk;
    int64_t dsel, base, dv;
    int64_t tpos;           // trace records written this run
    int64_t eval_bytes, commit_bytes;
    int32_t done, begun, selected, fault;
    int32_t tenure, kick, cand, budget;
    int32_t ul_valid, ul_ovf;
    int32_t dtar, ncand, take, nq, use_list;
    int32_t bi, bp, bv, oldv, nties, stamp, ncw;
    int32_t cmdone;
    uint64_t bpk;           // packed digits of codeword bi before the move (set with bi)
    int32_t th_on;          // the leaving side comes from th (else the per-step walks)
    int32_t nstale;         // codewords whose th rows the next evaluator recounts
    int32_t stale[STALE_MAX];
    int32_t ndirty;         // codewords logged in dirty since best was last made equal to code
    int64_t stale_idx[STALE_MAX];   // their word indices and packed digits after the step,
    uint64_t stale_pk[STALE_MAX];   // so the evaluator's recount reads them with the block
    // wall_clock64 ticks spent per phase (diagnostics): 0 select, 1 post,
    // 2 begin, 3 pick, 4 list upkeep, 5 shuffle helper, 6 rng helper, 7 kick,
    // 8 previous begin's end -> selection start (the evaluator's span),
    // 9 selection end -> post start (the commit's span), 10 gap previous
    // begin's end -> evaluator start, 11 gap selection end -> commit start
    int64_t tph[12];
    int64_t t_eve, t_cme;
    // stop conditions checked before every step, as the reference checks them
    // at the top of its loop: stop once best_unc <= target (target < 0: none),
    // and do not begin a step once the device clock has passed deadline (0:
    // none).  ttt_it is the iteration whose move first brought best_unc to
    // the target (-1: not yet), t_ttt the device clock then.
    int64_t target, deadline, ttt_it, t_ttt;
    // algorithmic bytes of the reference's move per begun step: the scan of
    // the code (M n), both sides of every candidate codeword's evaluation
    // walked in full (2 B per counter read) and the commit's read-modify-
    // write of both spheres (8 B per pattern); and the steps counted
    int64_t move_bytes, move_steps;
    // --wide: every move of a candidate codeword is a candidate while the
    // uncovered list is valid and short (wnow, decided per step); per =
    // candidates per candidate codeword (n (q-1) when wnow, else dtar)
    int32_t wide, wnow, per, pad_w;
    // the evaluator's ring refill, decided by begin (the last writer before
    // the evaluator) so that every filling workgroup sees one decision:
    // fill_on, the first position and the generator state there
    int64_t fill_g0;
    uint64_t fill_s[4];
    int32_t fill_on, pad_f;
    int64_t* trace;         // per-run trace records (tcap of them), or null
    int64_t tcap;
    // the step's selection as commit workgroup 0 publishes it, applied by the
    // post kernel: kept apart from the fields the selection reads (cursors,
    // generator state, `selected`), which commit workgroups that start later
    // still read to make the same selection
    uint64_t p_rst[4];
    int64_t p_rgen, p_ruse, p_rlow, p_dsel, p_base, p_dv, p_t_eve;
    uint64_t p_bpk;
    int32_t p_bi, p_bp, p_bv, p_oldv, p_nties, p_stamp, p_ncw, p_ok;
};
static_assert(sizeof(Ctl) % 16 == 0, "Ctl is copied in 8-byte words");

// The single-workgroup phases work on a copy of the control block in LDS:
// one parallel load at entry and one store at exit instead of a dependent
// global round trip per field.  All threads call these.
__device__ inline void ctl_to_lds(Ctl* lds, const Ctl* g)
{
    const int64_t* s = reinterpret_cast<const int64_t*>(g);
    int64_t* d = reinterpret_cast<int64_t*>(lds);
    for (int x = threadIdx.x; x < (int)(sizeof(Ctl) / 8); x += WG) d[x] = s[x];
    __syncthreads();
}
__device__ inline void ctl_from_lds(Ctl* g, const Ctl* lds)
{
    __syncthreads();
    const int64_t* s = reinterpret_cast<const int64_t*>(lds);
    int64_t* d = reinterpret_cast<int64_t*>(g);
    for (int x = threadIdx.x; x < (int)(sizeof(Ctl) / 8); x += WG) d[x] = s[x];
    __syncthreads();
}

// One load from every 64-byte line of the cell descriptor, issued together at
// a kernel's start: the fields read later (in branches, after the first
// memory round trips) then hit the scalar cache instead of each missing in
// turn.
template <typename T>
__device__ __forceinline__ void desc_touch(const T& d)
{
    constexpr int nw = (int)(sizeof(T) / 4);
    const uint32_t* w = reinterpret_cast<const uint32_t*>(&d);
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < (nw + 15) / 16; k++) acc ^= w[16 * k < nw ? 16 * k : nw - 1];
    asm volatile("" ::"s"(acc));
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Device pointers of a cell.  In device code they carry the global address
// space, so every access through them is a global_* instruction whether the
// descriptor comes in a kernel argument or from a table in device memory.
#ifdef __HIP_DEVICE_COMPILE__
#define GPTR(T) __attribute__((address_space(1))) T*
#else
#define GPTR(T) T*
#endif

struct Cell {
    int q, n, R, M, q1, F, E;
    int64_t P;              // (q-1)^R patterns per position subset
    int64_t N, nzw, nblk, V, ncap, ulcap;
    int64_t ul_use;         // list length below which the evaluator reads the list, (R+1) E P / n
    int64_t pw[MAXN];
    uint32_t magic;         // ceil(2^32 / (q-1)): t / (q-1) = umulhi(t, magic)
    uint64_t nmag, mmag;    // floor(2^64 / N), floor(2^64 / M) (0 for M = 1): see mod_magic
    // sphere walks by tables: pattern t of an R-subset = hi[t / wnl] + lo[t % wnl],
    // lo over the wa innermost positions of the subset (wnl = (q-1)^wa words),
    // hi over the other R - wa (wnh words); wok = 0 when they do not fit
    int32_t wa, wnl, wnh, wok;
    uint32_t wmag;          // t / wnl = umulhi(t, wmag) (wnl > 1)
    GPTR(uint32_t) cnt32;        // cnt as uint32 pairs
    GPTR(uint32_t) zb;
    GPTR(uint32_t) ob;
    GPTR(uint32_t) blk;
    GPTR(uint8_t) code;
    GPTR(uint64_t) cpk;
    GPTR(int64_t) cidx;
    GPTR(uint8_t) best;
    GPTR(int32_t) tabu;
    GPTR(uint32_t) fsub;         // C(n,R) R-subsets of all positions, 4 bits each
    GPTR(uint32_t) psub;         // n x C(n-1,R) R-subsets avoiding p
    GPTR(uint32_t) bsub;         // subsets of size 0..R, grouped by size
    int64_t bsub_off[MAXR + 2];   // first subset of each size in bsub
    int64_t bpat_off[MAXR + 2];   // first ball word of each size
    GPTR(int64_t) ul_idx;
    GPTR(uint64_t) ul_pk;
    GPTR(uint32_t) inl;          // list membership plane
    GPTR(uint32_t) shuf;         // shuffled candidate order (original indices)
    GPTR(int32_t) acc;           // slot accumulators, ACCW per slot
    GPTR(int32_t) slot_cw;       // slot -> codeword
    GPTR(uint64_t) slot_cpk;     // slot -> its codeword's packed digits (as cpk)
    GPTR(uint64_t) rbuf;
    GPTR(int64_t) scratch64;     // small reduction outputs
    GPTR(int32_t) arrive;        // workgroup arrival sub-counters of the single-move commit
    GPTR(int32_t) dirty;         // codewords changed since best == code (DIRTY_CAP of them)
    GPTR(uint64_t) jump;         // the rng jump matrices J_1 .. J_{RNG_SPLIT-1}
    GPTR(uint32_t) xor32;        // per word, the XOR of the indices of its covering codewords (uint16 pairs)
    GPTR(int32_t) th;            // T and H_j per codeword, THW each
    GPTR(int32_t) cst;           // staged selection reads: codeword, slot, digits, tabu stamp (4 each)
    GPTR(int64_t) cbs;           // staged selection reads: word index of the candidate's codeword
    GPTR(uint64_t) cps;          // staged selection reads: packed digits of the candidate's codeword
    GPTR(uint64_t) csr;          // staged selection reads: the ring draws from the selection's cursor (SEL_WAVE + 1)
    GPTR(int32_t) dth;           // th changes made since the last evaluator (THW per codeword), which folds them in
    GPTR(uint8_t) dist;          // distance transform output (N bytes), lazily
    GPTR(int32_t) accw;          // --wide: entering counts per slot, position and value (WIDE_W each), lazily
    GPTR(Ctl) ctl;
};

// The uncovered list only serves the evaluator, which reads it while it is
// shorter than ul_use: keep it while at most this many words are uncovered
// (and it fits), rebuild it below half of that.  Speed only: the list never
// changes a value.
__host__ __device__ inline int64_t list_keep(const Cell& C)
{
    const int64_t cap = C.ulcap / 2;
    return C.ul_use < cap ? C.ul_use : cap;
}

// ------------------------------------------------------------- helpers
__device__ __forceinline__ uint16_t cnt_get(const Cell& C, int64_t w)
{
    return reinterpret_cast<const uint16_t*>(C.cnt32)[w];
}
__device__ __forceinline__ void cnt_put(const Cell& C, int64_t w, uint16_t v)
{
    reinterpret_cast<uint16_t*>(C.cnt32)[w] = v;
}
__device__ __forceinline__ uint32_t xor_get(const Cell& C, int64_t w)
{
    return reinterpret_cast<const uint16_t*>(C.xor32)[w];
}
__device__ __forceinline__ void xor_put(const Cell& C, int64_t w, uint32_t v)
{
    reinterpret_cast<uint16_t*>(C.xor32)[w] = (uint16_t)v;
}
__device__ __forceinline__ uint32_t bit_of(const uint32_t* plane, int64_t w)
{
    return (plane[w >> 5] >> (w & 31)) & 1u;
}
__device__ __forceinline__ void bit_set(uint32_t* plane, int64_t w)
{
    atomicOr(&plane[w >> 5], 1u << (w & 31));
}
__device__ __forceinline__ void bit_clr(uint32_t* plane, int64_t w)
{
    atomicAnd(&plane[w >> 5], ~(1u << (w & 31)));
}

// The word-index arithmetic below is also compiled for the host, where a
// bounds-checked emulation runs it over the same plan (cov_plan.h).

// high half of a 32 x 32-bit product (v_mul_hi_u32 on the device)
__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

// high half of a 64 x 64-bit product
__host__ __device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// x mod d for a fixed d >= 2 with mag = floor(2^64 / d): the estimate
// mulhi(x, mag) is floor(x / d) or one less (mag > 2^64/d - 1, so
// x mag / 2^64 > x/d - 1), hence one conditional subtraction.  Exact for
// every 64-bit x; a few multiplies instead of the 64-bit division routine.
__host__ __device__ __forceinline__ uint64_t mod_magic(uint64_t x, uint64_t d, uint64_t mag)
{
    if (d == 1) return 0;
    uint64_t r = x - mulhi64(x, mag) * d;
    return r >= d ? r - d : r;
}
// x mod d for 1 <= d <= 65536 in 32-bit steps: with x = hi 2^32 + lo,
// x mod d = ((hi mod d)(2^32 mod d) + lo mod d) mod d, every term below
// d (d - 1) < 2^32, and the first three remainders independent of each other
__host__ __device__ __forceinline__ uint32_t mod_u16(uint64_t x, uint32_t d)
{
    const uint32_t hi = (uint32_t)(x >> 32), lo = (uint32_t)x;
    const uint32_t p32 = (0xFFFFFFFFu % d + 1u) % d;
    return ((hi % d) * p32 + lo % d) % d;
}
// x mod d for any d >= 1 (the short form when d allows)
__host__ __device__ __forceinline__ uint64_t mod_any(uint64_t x, uint64_t d)
{
    return d <= 65536 ? (uint64_t)mod_u16(x, (uint32_t)d) : x % d;
}

// floor(2^64 / d) for d >= 2, on the host (0 for d = 1)
inline uint64_t magic64(uint64_t d)
{
    return d <= 1 ? 0 : (uint64_t)(((unsigned __int128)1 << 64) / d);
}

// nibble-wise "differs" mask: bit 4j set iff digit j of a and b differ
__host__ __device__ __forceinline__ uint64_t nib_ne(uint64_t a, uint64_t b)
{
    uint64_t x = a ^ b;
    return (x | (x >> 1) | (x >> 2) | (x >> 3)) & NIB1;
}
__host__ __device__ __forceinline__ int nib_dist(uint64_t a, uint64_t b)
{
    return __builtin_popcountll(nib_ne(a, b));
}
// the t-th (from 0, increasing) position where digits a and b differ
__host__ __device__ __forceinline__ int slot_diff(uint64_t a, uint64_t b, int t)
{
    uint64_t ne = nib_ne(a, b);
    for (int k = 0; k < t; k++) ne &= ne - 1;
    return (__builtin_ctzll(ne)) >> 2;
}

// word w (digits wpk) became, or stopped being, covered by codeword o alone
// (sign +1 / -1): o's leaving counts change when d(w, o) = R.  The change goes
// to dth, not th: th stays as the evaluator left it while the commit's
// workgroups select from it; the next evaluator folds dth into th.
__device__ __forceinline__ void th_event(const Cell& C, int o, uint64_t wpk, int sign)
{
    if ((unsigned)o >= (unsigned)C.M) return;
    const uint64_t ne = nib_ne(wpk, C.cpk[o]);
    if (__popcll(ne) != C.R) return;
    int32_t* t = C.dth + (int64_t)o * THW;
    atomicAdd(&t[0], sign);
    uint64_t m = ne;
    while (m) {
        const int b = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        atomicAdd(&t[1 + (b >> 2)], sign);
    }
}
__host__ __device__ __forceinline__ int nib_get(uint64_t pk, int j)
{
    return (int)((pk >> (4 * j)) & 15u);
}
__host__ __device__ __forceinline__ uint64_t nib_put(uint64_t pk, int j, int v)
{
    return (pk & ~(15ULL << (4 * j))) | ((uint64_t)v << (4 * j));
}

// digits of a word index, packed (32-bit arithmetic when q^n fits)
__host__ __device__ inline uint64_t pack_index(const Cell& C, int64_t w)
{
    uint64_t pk = 0;
    if (C.N <= 0xFFFFFFFFLL) {
        uint32_t x = (uint32_t)w;
        const uint32_t q = (uint32_t)C.q;
        for (int j = C.n - 1; j >= 0; j--) {
            const uint32_t y = x / q;
            pk |= (uint64_t)(x - y * q) << (4 * j);
            x = y;
        }
        return pk;
    }
    for (int j = C.n - 1; j >= 0; j--) {
        pk |= (uint64_t)(w % C.q) << (4 * j);
        w /= C.q;
    }
    return pk;
}

// relaxed agent-scope load: sees device atomics of this kernel without
// relying on the vector L1 (used by the single-workgroup phases)
__device__ __forceinline__ uint32_t ld_acq32(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__host__ __device__ __forceinline__ uint32_t div_q1(const Cell& C, uint32_t t)
{
    return C.q1 == 1 ? t : mulhi32(t, C.magic);
}

// Word of pattern t (0 <= t < (q-1)^r) on the position subset `sub` (r
// positions, 4 bits each, increasing) around the word with digits pk at base:
// the last position varies fastest, and the k-th value of a position skips
// the centre digit (v = k + (k >= c)).  Also returns the packed digits.
__host__ __device__ __forceinline__ int64_t pattern_word(const Cell& C, int64_t base, uint64_t pk,
                                                uint32_t sub, int r, uint32_t t,
                                                uint64_t* wpk)
{
    int64_t a = base;
    uint64_t p2 = pk;
    for (int i = r - 1; i >= 0; i--) {
        const uint32_t tq = div_q1(C, t);
        const int k = (int)(t - tq * (uint32_t)C.q1);
        t = tq;
        const int j = (sub >> (4 * i)) & 15;
        const int c = nib_get(pk, j);
        const int v = k + (k >= c);
        a += (int64_t)(v - c) * C.pw[j];
        p2 = nib_put(p2, j, v);
    }
    if (wpk) *wpk = p2;
    return a;
}

// the packed digits alone of pattern t (pattern_word's second result, with
// no access to the powers)
__host__ __device__ __forceinline__ uint64_t pattern_digits(const Cell& C, uint64_t pk, uint32_t sub, int r, uint32_t t)
{
    uint64_t p2 = pk;
    for (int i = r - 1; i >= 0; i--) {
        const uint32_t tq = div_q1(C, t);
        const int k = (int)(t - tq * (uint32_t)C.q1);
        t = tq;
        const int j = (sub >> (4 * i)) & 15;
        const int c = nib_get(pk, j);
        p2 = nib_put(p2, j, k + (k >= c));
    }
    return p2;
}

// Offsets of the patterns of the R-subset `sub` around the word with digits
// pk, as two tables of one wave: lo over the wa innermost positions of the
// subset, hi over the rest, so that pattern t is the word
// base + hi[t / wnl] + lo[t % wnl] (the same word as pattern_word).  The wave
// builds them in its LDS slice; all 64 lanes.
constexpr int WT_MAX = 128;
struct WalkTab {
    int64_t lo[WT_MAX];
    int64_t hi[WT_MAX];
};
__host__ __device__ __forceinline__ int64_t walk_part(const Cell& C, uint64_t pk, uint32_t sub, int i1, int i0,
                                                     uint32_t x)
{
    int64_t s = 0;
    for (int i = i1; i >= i0; i--) {
        const uint32_t xq = div_q1(C, x);
        const int k = (int)(x - xq * (uint32_t)C.q1);
        x = xq;
        const int j = (sub >> (4 * i)) & 15;
        const int c = nib_get(pk, j);
        s += (int64_t)(k + (k >= c) - c) * C.pw[j];
    }
    return s;
}
// entries first, first + stride, ... of both tables
__host__ __device__ __forceinline__ void walk_fill(const Cell& C, uint64_t pk, uint32_t sub, int64_t* lo, int64_t* hi,
                                                   int first, int stride)
{
    const int r = C.R, a = C.wa;
    for (int l = first; l < C.wnl; l += stride) lo[l] = walk_part(C, pk, sub, r - 1, r - a, (uint32_t)l);
    for (int h = first; h < C.wnh; h += stride) hi[h] = walk_part(C, pk, sub, r - a - 1, 0, (uint32_t)h);
}
__device__ __forceinline__ void walk_tables(const Cell& C, uint64_t pk, uint32_t sub, WalkTab& tb)
{
    walk_fill(C, pk, sub, tb.lo, tb.hi, threadIdx.x & 63, 64);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__host__ __device__ __forceinline__ int64_t walk_word(const Cell& C, const int64_t* lo, const int64_t* hi,
                                                     int64_t base, uint32_t t)
{
    const uint32_t h = C.wnl == 1 ? t : mulhi32(t, C.wmag);
    return base + hi[h] + lo[t - h * (uint32_t)C.wnl];
}
__device__ __forceinline__ int64_t walk_word(const Cell& C, const WalkTab& tb, int64_t base, uint32_t t)
{
    return walk_word(C, tb.lo, tb.hi, base, t);
}
// the wave is done with its tables (before the next unit rebuilds them)
__device__ __forceinline__ void walk_release()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ----------------------------------------------------------------- rng
__host__ __device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// xoshiro256++ step (Blackman & Vigna)
__host__ __device__ __forceinline__ uint64_t xo_next(uint64_t s[4])
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

// Generate draws until rgen >= upto, at most `budget`, never overwriting a draw
// still to be read (positions below rlow + RB).  Called by all 64 lanes of one
// wave: the xoshiro state is wave-uniform (scalar unit), lane k keeps the k-th
// draw of each group of 64 and the group is stored at once.
__device__ inline int64_t wave_ring_fill(const Cell& C, Ctl* ctl, int64_t upto, int64_t budget)
{
    const int lane = threadIdx.x & 63;
    // wave-uniform state in scalar registers
    uint64_t s[4] = {uniform64(ctl->rst[0]), uniform64(ctl->rst[1]), uniform64(ctl->rst[2]),
                     uniform64(ctl->rst[3])};
    const int64_t g0 = (int64_t)uniform64((uint64_t)ctl->rgen);
    int64_t lim = (int64_t)uniform64((uint64_t)(ctl->rlow + RB));
    upto = (int64_t)uniform64((uint64_t)upto);
    budget = (int64_t)uniform64((uint64_t)budget);
    if (upto < lim) lim = upto;
    if (budget < lim - g0) lim = g0 + budget;     // (budget may be INT64_MAX: no g0 + budget)
    int64_t g = g0;
    while (g < lim) {
        const int nb = (lim - g < 64) ? (int)(lim - g) : 64;
        uint64_t mine = 0;
        for (int k = 0; k < nb; k++) {
            const uint64_t r = xo_next(s);
            if (lane == k) mine = r;
        }
        if (lane < nb) C.rbuf[(g + lane) & (RB - 1)] = mine;
        g += nb;
    }
    if (lane == 0 && g != g0) {
        ctl->rst[0] = s[0]; ctl->rst[1] = s[1]; ctl->rst[2] = s[2]; ctl->rst[3] = s[3];
        ctl->rgen = g;
    }
    return g - g0;
}

// Many waves at once.  The generator's state transition is linear over
// GF(2), so the state RNG_PART * w draws ahead is J_w s for a fixed 256 x 256
// bit matrix J_w = T^(RNG_PART w).  A matrix is 256 rows of 4 words: bit j of
// row b is set when bit b of the result depends on bit j of the state, bit b
// of a state being bit b % 64 of word b / 64.  A part's draws are a serial
// chain (about 110 cycles each on the scalar unit), so the chunk is cut in
// parts of one group of 64.
constexpr int RNG_SPLIT = 16;                   // parts of a fill
constexpr int RNG_PART = 64;                    // draws per part
constexpr int RNG_CHUNK = RNG_SPLIT * RNG_PART; // draws per fill
constexpr int RNG_WGS = RNG_SPLIT / 4;          // filling workgroups (of 4 waves) when the grid allows
constexpr int RNG_JUMP_WORDS = (RNG_SPLIT - 1) * 256 * 4;

// J_1 .. J_{RNG_SPLIT-1}, on the host (out: RNG_JUMP_WORDS words)
inline void rng_jump_tables(uint64_t* out)
{
    typedef uint64_t Mat[256][4];
    static Mat T, P, Q;
    // T: one step of the state transition (the output function is not needed)
    for (int j = 0; j < 256; j++) {
        uint64_t st[4] = {0, 0, 0, 0};
        st[j >> 6] = (uint64_t)1 << (j & 63);
        (void)xo_next(st);
        for (int b = 0; b < 256; b++) {
            if (j == 0) T[b][0] = T[b][1] = T[b][2] = T[b][3] = 0;
            if ((st[b >> 6] >> (b & 63)) & 1) T[b][j >> 6] |= (uint64_t)1 << (j & 63);
        }
    }
    // (A B) row b = xor of the rows j of B for the bits j of A's row b
    auto mul = [](const Mat& A, const Mat& B, Mat& C) {
        for (int b = 0; b < 256; b++) {
            uint64_t r[4] = {0, 0, 0, 0};
            for (int j = 0; j < 256; j++)
                if ((A[b][j >> 6] >> (j & 63)) & 1)
                    for (int k = 0; k < 4; k++) r[k] ^= B[j][k];
            for (int k = 0; k < 4; k++) C[b][k] = r[k];
        }
    };
    // P = T^RNG_PART by squaring (RNG_PART is a power of two)
    for (int b = 0; b < 256; b++)
        for (int k = 0; k < 4; k++) P[b][k] = T[b][k];
    for (int e = 1; e < RNG_PART; e <<= 1) {
        mul(P, P, Q);
        for (int b = 0; b < 256; b++)
            for (int k = 0; k < 4; k++) P[b][k] = Q[b][k];
    }
    // J_1 = P, J_w = J_{w-1} P
    for (int b = 0; b < 256; b++)
        for (int k = 0; k < 4; k++) Q[b][k] = P[b][k];
    for (int w = 1; w < RNG_SPLIT; w++) {
        for (int b = 0; b < 256; b++)
            for (int k = 0; k < 4; k++) out[((size_t)(w - 1) * 256 + b) * 4 + k] = Q[b][k];
        if (w + 1 < RNG_SPLIT) {
            static Mat R;
            mul(Q, P, R);
            for (int b = 0; b < 256; b++)
                for (int k = 0; k < 4; k++) Q[b][k] = R[b][k];
        }
    }
}

// The fill decided by begin (fill_on): RNG_CHUNK draws from fill_g0
// (state fill_s), part w the RNG_PART draws from fill_g0 + RNG_PART w, its
// state J_w fill_s (lane l computes bits l, l+64, l+128, l+192 of it, one
// ballot per word).  Waves wi = 0 .. nwaves-1 of the filling workgroups take
// parts wi, wi + nwaves, ...; the wave with the last part stores the state
// after the chunk.  Nothing else reads or writes rgen or rst while the
// evaluator runs.  The decision is read from rd, the state stored to ctl.
// All threads of a filling workgroup.
__device__ inline void ring_fill_parts(const Cell& C, const Ctl* rd, Ctl* ctl, int wi, int nwaves)
{
    const int lane = threadIdx.x & 63;
    const int64_t g0 = (int64_t)uniform64((uint64_t)rd->fill_g0);
    const uint64_t s0[4] = {uniform64(rd->fill_s[0]), uniform64(rd->fill_s[1]), uniform64(rd->fill_s[2]),
                            uniform64(rd->fill_s[3])};
    for (int w = wi; w < RNG_SPLIT; w += nwaves) {
        uint64_t s[4] = {s0[0], s0[1], s0[2], s0[3]};
        if (w > 0) {
            const uint64_t* M = C.jump + (size_t)(w - 1) * 256 * 4;
            uint64_t o[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t* row = M + (size_t)(64 * k + lane) * 4;
                const uint64_t x = (row[0] & s0[0]) ^ (row[1] & s0[1]) ^ (row[2] & s0[2]) ^ (row[3] & s0[3]);
                o[k] = __ballot(__popcll(x) & 1);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) s[k] = uniform64(o[k]);
        }
        const int64_t gb = g0 + (int64_t)RNG_PART * w;
        for (int grp = 0; grp < RNG_PART / 64; grp++) {
            uint64_t mine = 0;
            for (int k = 0; k < 64; k++) {
                const uint64_t r = xo_next(s);
                if (lane == k) mine = r;
            }
            C.rbuf[(gb + 64 * grp + lane) & (RB - 1)] = mine;
        }
        if (w == RNG_SPLIT - 1 && lane == 0) {
            ctl->rst[0] = s[0]; ctl->rst[1] = s[1]; ctl->rst[2] = s[2]; ctl->rst[3] = s[3];
            ctl->rgen = g0 + RNG_CHUNK;
        }
    }
}

__device__ __forceinline__ int64_t clk() { return (int64_t)wall_clock64(); }

// a phase timer of the control block in global memory, added without a read
// (the wide kernels; the one-workgroup phases add to their LDS copy)
__device__ __forceinline__ void tph_add(Ctl* ctl, int k, int64_t ticks)
{
    atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->tph[k]), (unsigned long long)ticks);
}

__device__ __forceinline__ uint64_t ring_at(const Cell& C, int64_t pos)
{
    return C.rbuf[pos & (RB - 1)];
}

// ------------------------------------------------- workgroup reductions
struct WgRed {
    int64_t v[WG / 64];
    int64_t out;
};

__device__ __forceinline__ int64_t wave_sum64(int64_t x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}
__device__ __forceinline__ int64_t wave_min64(int64_t x)
{
    for (int o = 32; o > 0; o >>= 1) {
        int64_t y = __shfl_xor(x, o, 64);
        x = y < x ? y : x;
    }
    return x;
}
__device__ __forceinline__ int64_t wave_max64(int64_t x)
{
    for (int o = 32; o > 0; o >>= 1) {
        int64_t y = __shfl_xor(x, o, 64);
        x = y > x ? y : x;
    }
    return x;
}

// the value of the wave's first lane (lowest index) with `has` set, else
// `none`: a ballot and two lane reads (all 64 lanes)
__device__ __forceinline__ int64_t wave_first64(bool has, int64_t v, int64_t none)
{
    const uint64_t m = __ballot(has);
    if (!m) return none;
    const int l = __ffsll((unsigned long long)m) - 1;
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((uint64_t)v >> 32), l);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// all threads of the workgroup must call these
__device__ inline int64_t wg_sum(int64_t x, WgRed& r)
{
    x = wave_sum64(x);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) r.v[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t s = 0;
        for (int i = 0; i < WG / 64; i++) s += r.v[i];
        r.out = s;
    }
    __syncthreads();
    const int64_t res = r.out;
    __syncthreads();
    return res;
}
__device__ inline int64_t wg_min(int64_t x, WgRed& r)
{
    x = wave_min64(x);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) r.v[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t s = r.v[0];
        for (int i = 1; i < WG / 64; i++) s = r.v[i] < s ? r.v[i] : s;
        r.out = s;
    }
    __syncthreads();
    const int64_t res = r.out;
    __syncthreads();
    return res;
}
__device__ inline int64_t wg_max(int64_t x, WgRed& r)
{
    x = wave_max64(x);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) r.v[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t s = r.v[0];
        for (int i = 1; i < WG / 64; i++) s = r.v[i] > s ? r.v[i] : s;
        r.out = s;
    }
    __syncthreads();
    const int64_t res = r.out;
    __syncthreads();
    return res;
}

// inclusive prefix sum over the workgroup (thread order); total in *tot
struct WgScan {
    int64_t w[WG / 64];
};
__device__ inline int64_t wg_incl_sum(int64_t x, WgScan& s, int64_t* tot)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s.w[wv] = x;
    __syncthreads();
    int64_t pre = 0, all = 0;
    for (int i = 0; i < WG / 64; i++) {
        if (i < wv) pre += s.w[i];
        all += s.w[i];
    }
    __syncthreads();
    *tot = all;
    return x + pre;
}
__device__ inline int64_t wg_incl_min(int64_t x, WgScan& s)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(x, o, 64);
        if (lane >= o && y < x) x = y;
    }
    if (lane == 63) s.w[wv] = x;
    __syncthreads();
    for (int i = 0; i < wv; i++)
        if (s.w[i] < x) x = s.w[i];
    __syncthreads();
    return x;
}
__device__ inline int64_t wg_incl_max(int64_t x, WgScan& s)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(x, o, 64);
        if (lane >= o && y > x) x = y;
    }
    if (lane == 63) s.w[wv] = x;
    __syncthreads();
    for (int i = 0; i < wv; i++)
        if (s.w[i] > x) x = s.w[i];
    __syncthreads();
    return x;
}

}  // namespace cov
