// cov_cover.h -- coverage counts (kernel b) and uncovered-set reductions (kernel d).
//
// Recount.  cnt[w] = #{codewords within distance R of w}.  Instead of M*V
// global atomics, the word space is cut into tiles of T = q^(n-t) consecutive
// words that share their first t digits (MSB-first, so a tile is contiguous).
// A workgroup owns one tile in LDS: a codeword whose prefix is at distance
// d <= R from the tile's prefix contributes its suffix ball of radius R - d,
// added with LDS atomics; the finished tile is written to HBM once, coalesced.
// HBM traffic is 2N bytes of counters plus M*8 bytes of codeword digits per
// tile, against 4*M*V bytes for global read-modify-writes.
//
// Ball update.  +-1 over the radius-R ball of given centres with 32-bit
// atomics on the word holding the 16-bit counter (cnt <= M <= 65535: a lane
// never carries into or borrows from its neighbour).  With one centre every
// word is distinct, so the [cnt==0]/[cnt==1] planes, block counts and the
// uncovered list are updated from each atomic's old value; with several
// centres the planes are re-derived from cnt afterwards.
#pragma once
#include "cov_common.h"

namespace cov {

constexpr int TILE_MAX = 16384;   // words per recount tile (64 KB of LDS counters)

// tables of the suffix ball used by the recount (positions t..n-1)
struct RecountPlan {
    int t;                 // prefix digits fixed per tile
    int64_t T;             // words per tile
    int64_t ntile;
    uint32_t* ssub;        // suffix subsets, grouped by size
    int64_t ssub_off[MAXR + 2];
    int64_t spat[MAXR + 2];     // cumulative suffix-ball sizes by radius
    int64_t spr[MAXR + 2];      // (q-1)^r
    float sinv[MAXR + 2];       // 1/(q-1)^r
};

// x / d for x < 2^24 via a float reciprocal estimate and one correction
__device__ __forceinline__ uint32_t udiv_small(uint32_t x, uint32_t d, float inv, uint32_t* rem)
{
    int32_t qf = (int32_t)((float)x * inv);
    int32_t r = (int32_t)x - qf * (int32_t)d;
    if (r < 0) { qf--; r += (int32_t)d; }
    if (r >= (int32_t)d) { qf++; r -= (int32_t)d; }
    *rem = (uint32_t)r;
    return (uint32_t)qf;
}

constexpr int RC_SSUB_LDS = 1024;   // suffix subsets the recount keeps in LDS (else read from memory)

__global__ void __launch_bounds__(WG) k_recount(const Cell C, RecountPlan plan)
{
    extern __shared__ uint32_t tile[];
    __shared__ int64_t rc_off[WAVES][64];
    __shared__ uint64_t rc_pk[WAVES][64];
    __shared__ int64_t rc_sb[WAVES][64];
    // the per-item tables, indexed per lane: in LDS rather than read from the
    // kernel arguments with a divergent index
    __shared__ int64_t rc_spat[MAXR + 2], rc_spr[MAXR + 2], rc_soff[MAXR + 2], rc_pw[MAXN];
    __shared__ float rc_sinv[MAXR + 2];
    __shared__ uint32_t rc_ssub[RC_SSUB_LDS];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int t = plan.t;
    const int64_t T = plan.T;
    const int64_t nss = plan.ssub_off[MAXR + 1];
    const bool ss_lds = nss <= RC_SSUB_LDS;
    if (tid < MAXR + 2) {
        rc_spat[tid] = plan.spat[tid];
        rc_spr[tid] = plan.spr[tid];
        rc_soff[tid] = plan.ssub_off[tid];
        rc_sinv[tid] = plan.sinv[tid];
    }
    if (tid < MAXN) rc_pw[tid] = tid < C.n ? C.pw[tid] : 0;
    if (ss_lds)
        for (int64_t x = tid; x < nss; x += WG) rc_ssub[x] = plan.ssub[x];
    for (int64_t tau = blockIdx.x; tau < plan.ntile; tau += gridDim.x) {
        for (int64_t x = tid; x < T; x += WG) tile[x] = 0;
        // packed prefix digits of the tile (positions 0..t-1)
        uint64_t ppk = 0;
        {
            int64_t r = tau;
            for (int j = t - 1; j >= 0; j--) {
                ppk |= (uint64_t)(r % C.q) << (4 * j);
                r /= C.q;
            }
        }
        const uint64_t pmask = t ? ((t >= 16) ? ~0ULL : ((1ULL << (4 * t)) - 1)) : 0ULL;
        __syncthreads();
        // Each wave takes 64 codewords at a time, one per lane; those whose
        // prefix is within R of the tile's contribute their suffix ball of
        // radius R - d.  The wave's balls are laid end to end (a wave prefix
        // sum of their sizes) and the lanes take consecutive items of the
        // whole, each finding its codeword by a binary search of the starts.
        for (int i0 = wv * 64; i0 < C.M; i0 += WAVES * 64) {
            const int i = i0 + lane;
            uint64_t cpk = 0;
            int64_t vs = 0, sbase = 0;
            if (i < C.M) {
                cpk = C.cpk[i];
                const int64_t ci = C.cidx[i];
                const int dpre = __popcll(nib_ne(cpk & pmask, ppk));
                if (dpre <= C.R) {
                    vs = plan.spat[C.R - dpre + 1];
                    sbase = ci % T;
                }
            }
            int64_t incl = vs;
            for (int o = 1; o < 64; o <<= 1) {
                const int64_t y = __shfl_up(incl, o, 64);
                if (lane >= o) incl += y;
            }
            const int64_t total = __shfl(incl, 63, 64);
            rc_off[wv][lane] = incl - vs;
            rc_pk[wv][lane] = cpk;
            rc_sb[wv][lane] = sbase;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int64_t f = lane; f < total; f += 64) {
                // the last codeword whose ball starts at or before f (its ball is not empty)
                int j = 0;
#pragma unroll
                for (int step = 32; step; step >>= 1)
                    if (rc_off[wv][j + step] <= f) j += step;
                const int64_t loc0 = f - rc_off[wv][j];
                int r = 0;
                while (loc0 >= rc_spat[r + 1]) r++;
                const uint32_t loc = (uint32_t)(loc0 - rc_spat[r]);
                uint32_t pat;
                const uint32_t sidx = r ? udiv_small(loc, (uint32_t)rc_spr[r], rc_sinv[r], &pat) : 0;
                if (!r) pat = 0;
                const int64_t si = rc_soff[r] + sidx;
                const uint32_t sub = ss_lds ? rc_ssub[si] : plan.ssub[si];
                // the word of pattern pat on subset sub (pattern_word's, with the
                // place values from LDS)
                const uint64_t cpk = rc_pk[wv][j];
                int64_t w = rc_sb[wv][j];
                uint32_t tt = pat;
                for (int e = r - 1; e >= 0; e--) {
                    const uint32_t tq = div_q1(C, tt);
                    const int k = (int)(tt - tq * (uint32_t)C.q1);
                    tt = tq;
                    const int jj = (sub >> (4 * e)) & 15;
                    const int c = nib_get(cpk, jj);
                    w += (int64_t)(k + (k >= c) - c) * rc_pw[jj];
                }
                atomicAdd(&tile[w], 1u);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        __syncthreads();
        uint16_t* cnt = reinterpret_cast<uint16_t*>(C.cnt32) + tau * T;
        for (int64_t x = tid; x < T; x += WG) cnt[x] = (uint16_t)tile[x];
        __syncthreads();
    }
}

// cnt -> planes, block counts and the uncovered count; invalidates the list
__global__ void __launch_bounds__(WG) k_derive(const Cell C)
{
    __shared__ WgRed red;
    __shared__ uint32_t bc[2];
    const int64_t groups_per_wg = WG;              // 256 groups of 32 words = 2 blocks
    int64_t my_unc = 0;
    for (int64_t g0 = (int64_t)blockIdx.x * groups_per_wg; g0 < C.nzw; g0 += (int64_t)gridDim.x * groups_per_wg) {
        if (threadIdx.x < 2) bc[threadIdx.x] = 0;
        __syncthreads();
        const int64_t g = g0 + threadIdx.x;
        uint32_t z = 0, o = 0;
        if (g < C.nzw) {
            const int64_t w0 = g * 32;
            const uint16_t* cnt = reinterpret_cast<const uint16_t*>(C.cnt32);
            if (w0 + 32 <= C.N) {
                const uint4* p4 = reinterpret_cast<const uint4*>(cnt + w0);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    uint4 v = p4[k];
                    uint32_t e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int h = 0; h < 4; h++) {
                        const uint32_t lo = e[h] & 0xFFFFu, hi = e[h] >> 16;
                        const int b = k * 8 + h * 2;
                        z |= (uint32_t)(lo == 0) << b;
                        o |= (uint32_t)(lo == 1) << b;
                        z |= (uint32_t)(hi == 0) << (b + 1);
                        o |= (uint32_t)(hi == 1) << (b + 1);
                    }
                }
            } else {
                for (int b = 0; b < 32 && w0 + b < C.N; b++) {
                    const uint16_t c = cnt[w0 + b];
                    z |= (uint32_t)(c == 0) << b;
                    o |= (uint32_t)(c == 1) << b;
                }
            }
            C.zb[g] = z;
            C.ob[g] = o;
            const int pc_ = __popc(z);
            my_unc += pc_;
            if (pc_) atomicAdd(&bc[threadIdx.x >> 7], (uint32_t)pc_);
        }
        __syncthreads();
        if (threadIdx.x < 2) {
            const int64_t b = (g0 >> 7) + threadIdx.x;
            if (b < C.nblk) C.blk[b] = bc[threadIdx.x];
        }
        __syncthreads();
    }
    const int64_t s = wg_sum(my_unc, red);
    if (threadIdx.x == 0 && s) atomicAdd(reinterpret_cast<unsigned long long*>(&C.ctl->unc),
                                         (unsigned long long)s);
}

// list insertion of a word that just became uncovered (call with valid list)
__device__ __forceinline__ void list_insert(const Cell& C, Ctl* ctl, int64_t w, uint64_t wpk)
{
    const uint32_t bit = 1u << (w & 31);
    const uint32_t old = atomicOr(&C.inl[w >> 5], bit);
    if (old & bit) return;
    const unsigned long long pos = atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->ul_n), 1ULL);
    if ((int64_t)pos < C.ulcap) {
        C.ul_idx[pos] = w;
        C.ul_pk[pos] = wpk;
    } else {
        atomicOr(&ctl->ul_ovf, 1);
    }
}

// one ball word, incremental (distinct words within the call); returns the
// change in the uncovered count
// With tr >= 0 the word's coverer XOR takes codeword tr in or out, and the
// codeword left as the word's only coverer, or no longer its only one, has its
// leaving counts updated.
__device__ __forceinline__ int ball_word(const Cell& C, Ctl* ctl, int64_t w, uint64_t wpk,
                                         int delta, bool list_on, int tr = -1)
{
    const uint32_t sh = (uint32_t)(w & 1) * 16u;
    uint32_t* p = &C.cnt32[w >> 1];
    int d = 0;
    uint32_t ox = 0;
    if (tr >= 0) ox = (atomicXor(&C.xor32[w >> 1], (uint32_t)tr << sh) >> sh) & 0xFFFFu;
    if (delta > 0) {
        const uint32_t old = atomicAdd(p, 1u << sh);
        const uint32_t oc = (old >> sh) & 0xFFFFu;
        if (tr >= 0 && oc == 1) th_event(C, (int)ox, wpk, -1);
        if (oc == 0) {
            bit_clr(C.zb, w); bit_set(C.ob, w);
            atomicSub(&C.blk[w >> BLK_SHIFT], 1u);
            d = -1;
        } else if (oc == 1) {
            bit_clr(C.ob, w);
        }
    } else {
        const uint32_t old = atomicSub(p, 1u << sh);
        const uint32_t oc = (old >> sh) & 0xFFFFu;
        if (tr >= 0 && oc == 2) th_event(C, (int)(ox ^ (uint32_t)tr), wpk, +1);
        if (oc == 1) {
            bit_set(C.zb, w); bit_clr(C.ob, w);
            atomicAdd(&C.blk[w >> BLK_SHIFT], 1u);
            if (list_on) list_insert(C, ctl, w, wpk);
            d = 1;
        } else if (oc == 2) {
            bit_set(C.ob, w);
        }
    }
    return d;
}

// decode ball word g (0 <= g < V) around (base, pk)
__host__ __device__ __forceinline__ int64_t ball_word_at(const Cell& C, int64_t base, uint64_t pk,
                                                int64_t g, uint64_t* wpk)
{
    int r = 0;
    while (g >= C.bpat_off[r + 1]) r++;
    const int64_t loc = g - C.bpat_off[r];
    int64_t pr = 1;
    for (int i = 0; i < r; i++) pr *= C.q1;
    const int64_t sidx = loc / pr;
    const uint32_t pat = (uint32_t)(loc - sidx * pr);
    const uint32_t sub = C.bsub[C.bsub_off[r] + sidx];
    return pattern_word(C, base, pk, sub, r, pat, wpk);
}

// incremental single-ball update over the whole grid (k == 1 centre)
__device__ inline void ball_one(const Cell& C, int64_t base, uint64_t pk, int delta, WgRed& red)
{
    Ctl* ctl = C.ctl;
    const bool list_on = ctl->ul_valid != 0;
    int64_t d = 0;
    for (int64_t g = (int64_t)blockIdx.x * WG + threadIdx.x; g < C.V; g += (int64_t)gridDim.x * WG) {
        uint64_t wpk;
        const int64_t w = ball_word_at(C, base, pk, g, &wpk);
        d += ball_word(C, ctl, w, wpk, delta, list_on);
    }
    const int64_t s = wg_sum(d, red);
    if (threadIdx.x == 0 && s) atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->unc),
                                         (unsigned long long)s);
}

__global__ void __launch_bounds__(WG) k_ball_one(const Cell C, int64_t base,
                                                 uint64_t pk, int delta)
{
    __shared__ WgRed red;
    ball_one(C, base, pk, delta, red);
}

// the remove-and-repair descent's removal: the ball of the word that
// k_rr_pick left in *sel (none when it is negative)
__global__ void __launch_bounds__(WG) k_ball_drop(const Cell C, const int64_t* __restrict__ sel)
{
    __shared__ WgRed red;
    const int64_t base = *sel;
    if (base < 0) return;
    ball_one(C, base, pack_index(C, base), -1, red);
}

// counts only, several centres (planes re-derived afterwards)
__global__ void __launch_bounds__(WG) k_ball_multi(const Cell C,
                                                   const int64_t* __restrict__ centers, int k,
                                                   int delta)
{
    const int64_t total = (int64_t)k * C.V;
    for (int64_t g = (int64_t)blockIdx.x * WG + threadIdx.x; g < total; g += (int64_t)gridDim.x * WG) {
        const int64_t ci = g / C.V;
        const int64_t base = centers[ci];
        const int64_t w = ball_word_at(C, base, pack_index(C, base), g - ci * C.V, nullptr);
        const uint32_t sh = (uint32_t)(w & 1) * 16u;
        if (delta > 0) atomicAdd(&C.cnt32[w >> 1], 1u << sh);
        else           atomicSub(&C.cnt32[w >> 1], 1u << sh);
    }
}

// (d) #{w : cnt[w] == 0}, streaming 16 counters per thread per step
__global__ void __launch_bounds__(WG) k_count_zero(const Cell C, int64_t* out)
{
    __shared__ WgRed red;
    const uint16_t* cnt = reinterpret_cast<const uint16_t*>(C.cnt32);
    const int64_t nvec = C.N / 8;
    int64_t z = 0;
    const uint4* v4 = reinterpret_cast<const uint4*>(cnt);
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < nvec; i += (int64_t)gridDim.x * WG) {
        const uint4 v = v4[i];
        const uint32_t e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int h = 0; h < 4; h++) z += ((e[h] & 0xFFFFu) == 0) + ((e[h] >> 16) == 0);
    }
    for (int64_t w = nvec * 8 + (int64_t)blockIdx.x * WG + threadIdx.x; w < C.N; w += (int64_t)gridDim.x * WG)
        z += (cnt[w] == 0);
    const int64_t s = wg_sum(z, red);
    if (threadIdx.x == 0 && s) atomicAdd(reinterpret_cast<unsigned long long*>(out), (unsigned long long)s);
}

// (d) ordered gather of the words a predicate selects: uncovered words
// (cnt == 0) or words farther than r from the code (dist > r)
struct SelZero {
    const uint16_t* cnt;
    __device__ __forceinline__ bool operator()(int64_t w) const { return cnt[w] == 0; }
};
struct SelFar {
    const uint8_t* dist;
    int r;
    __device__ __forceinline__ bool operator()(int64_t w) const { return dist[w] > r; }
};

// pass 1: selected words per 4096-word block
template <class Sel>
__global__ void __launch_bounds__(WG) k_block_count(const Cell C, Sel sel, int64_t* bz)
{
    __shared__ WgRed red;
    for (int64_t b = blockIdx.x; b < C.nblk; b += gridDim.x) {
        int64_t z = 0;
        const int64_t w0 = b * BLK_WORDS;
        for (int k = threadIdx.x; k < BLK_WORDS; k += WG) {
            const int64_t w = w0 + k;
            if (w < C.N) z += sel(w);
        }
        const int64_t s = wg_sum(z, red);
        if (threadIdx.x == 0) bz[b] = s;
    }
}

// pass 2: block b writes its selected words at out[off[b]...], ascending
template <class Sel>
__global__ void __launch_bounds__(WG) k_block_emit(const Cell C, Sel sel, const int64_t* off,
                                                   int64_t* out, int64_t cap)
{
    __shared__ WgScan sc;
    constexpr int PER = BLK_WORDS / WG;   // 16 consecutive words per thread
    for (int64_t b = blockIdx.x; b < C.nblk; b += gridDim.x) {
        const int64_t o = off[b];
        if (o >= cap) continue;            // uniform per workgroup
        const int64_t w0 = b * BLK_WORDS + (int64_t)threadIdx.x * PER;
        int64_t mine = 0;
        for (int k = 0; k < PER; k++) mine += (w0 + k < C.N && sel(w0 + k));
        int64_t tot;
        int64_t pos = o + wg_incl_sum(mine, sc, &tot) - mine;
        for (int k = 0; k < PER; k++) {
            const int64_t w = w0 + k;
            if (w < C.N && sel(w)) {
                if (pos < cap) out[pos] = w;
                pos++;
            }
        }
        __syncthreads();
    }
}

}  // namespace cov
