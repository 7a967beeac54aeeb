// cov_dist.h -- the uint8 axis-DP Hamming distance transform (kernel c).
//
// dist[w] = min over codewords c of d_H(w, c).  Start from D = 0 on codewords
// and 255 elsewhere; for each axis j, every fiber (the q words that differ only
// in digit j) becomes D[x] <- min(D[x], 1 + min over the fiber).  After all n
// axes, in any order, D is the exact minimum distance (the R-fold dilation of
// cov/verify_independent.py:128-141 reads the same information as #{D <= r}).
//
// The n axes are grouped into a few passes over HBM instead of n:
//   inner pass  the last `a` axes: contiguous tiles of q^a words in LDS; the
//               tile is built from the codeword bitmap (one bit per word), so
//               no pass initialises D
//   outer pass  a group of b consecutive axes [j0, j0+b): a tile is q^b rows of
//               Lc contiguous words, the rows q^(n-j0-b) words apart; every
//               axis of the group is applied in LDS
// and the last pass also histograms the distances it stores.  Global traffic
// is 16-byte vectors; each LDS tile row keeps its global alignment mod 16 so
// vectors move without shifts, and only the partial vectors at a row's ends are
// stored byte by byte.
//
// The per-thread phases are __host__ __device__ functions so that a host build
// can run them thread by thread over plain arrays.
#pragma once
#include "cov_common.h"

namespace cov {

constexpr int DT_HB = MAXN + 2;   // histogram bins: 0..n, and 255 in bin n+1

// one pass of the transform, planned on the host
struct DtPass {
    int64_t N;
    int64_t rows;        // outer: rows per tile (q^b); inner: 1
    int64_t S_in;        // outer: row stride in words; inner: q^a (one block)
    int64_t Lc;          // outer: words per row chunk; inner: words per tile
    int64_t nch;         // outer: chunks per row segment
    int64_t ntiles;
    int32_t P;           // outer: LDS bytes per row (power of two); inner: LDS bytes
    int32_t pshift;      // outer: log2(P / 16)
    int32_t xs_shift;    // outer: log2 of the power of two >= Lc
    int32_t q, nax, first, last, hn;
    uint32_t u[MAXN];    // per axis: LDS stride, in rows (outer) or bytes (inner)
    uint32_t mg[MAXN];   // floor(2^32 / u) + 1, for u > 1
};

__host__ __device__ inline uint32_t dt_div(uint32_t f, uint32_t u, uint32_t mg)
{
    // exact floor(f / u) for f, u < 2^16
#ifdef __HIP_DEVICE_COMPILE__
    return u == 1 ? f : __umulhi(f, mg);
#else
    return u == 1 ? f : (uint32_t)(((uint64_t)f * mg) >> 32);
#endif
}

__host__ __device__ inline uint4 dt_bits_to_bytes(uint32_t bits16)
{
    // bit k set (a codeword) -> byte k = 0, else 255
    uint32_t w[4];
    for (int k = 0; k < 4; k++) {
        uint32_t x = 0;
        for (int b = 0; b < 4; b++)
            x |= (((bits16 >> (4 * k + b)) & 1u) ? 0u : 0xFFu) << (8 * b);
        w[k] = x;
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// one fiber: q bytes at p[0], p[u], ..., p[(q-1)u]
__host__ __device__ inline void dt_fiber(uint8_t* p, int64_t u, int q)
{
    int m = 255;
    for (int v = 0; v < q; v++) {
        const int x = p[v * u];
        m = x < m ? x : m;
    }
    const int m1 = m < 255 ? m + 1 : 255;
    for (int v = 0; v < q; v++) {
        const int x = p[v * u];
        p[v * u] = (uint8_t)(x < m1 ? x : m1);
    }
}

// packed 16-bit histogram counters: bin x in word x >> 2, field x & 3
struct DtHist {
    uint64_t w[(DT_HB + 3) / 4];
};
__host__ __device__ inline void dt_hist_add(DtHist& h, int x, int n)
{
    const int b = x <= n ? x : n + 1;
    const uint64_t one = (uint64_t)1 << (16 * (b & 3));
    for (int k = 0; k < (DT_HB + 3) / 4; k++) h.w[k] += (k == (b >> 2)) ? one : 0;
}

// ------------------------------------------------------------ inner pass
// tile t covers words [b0, b0 + len); LDS byte s + (w - b0) holds word w,
// s = b0 & 15
__host__ __device__ inline void dt_inner_load(const DtPass& p, int64_t t, int tid, int nth, uint8_t* sm,
                                              const uint8_t* D, const uint32_t* cwb)
{
    const int64_t b0 = t * p.Lc;
    const int64_t len = (p.N - b0 < p.Lc) ? p.N - b0 : p.Lc;
    const int64_t s = b0 & 15, g0 = b0 - s;
    const int64_t nv = (s + len + 15) >> 4;
    for (int64_t v = tid; v < nv; v += nth) {
        const int64_t g = g0 + 16 * v;
        uint4 x;
        if (p.first) x = dt_bits_to_bytes((cwb[g >> 5] >> (g & 31)) & 0xFFFFu);
        else x = *reinterpret_cast<const uint4*>(D + g);
        *reinterpret_cast<uint4*>(sm + 16 * v) = x;
    }
}

__host__ __device__ inline void dt_inner_axis(const DtPass& p, int64_t t, int ax, int tid, int nth, uint8_t* sm)
{
    const int64_t b0 = t * p.Lc;
    const int64_t len = (p.N - b0 < p.Lc) ? p.N - b0 : p.Lc;
    const int64_t s = b0 & 15;
    const uint32_t u = p.u[ax], mg = p.mg[ax];
    const uint32_t nf = (uint32_t)(len / p.q);
    for (uint32_t f = tid; f < nf; f += nth) {
        const uint32_t hi = dt_div(f, u, mg), lo = f - hi * u;
        dt_fiber(sm + s + (int64_t)hi * u * p.q + lo, u, p.q);
    }
}

__host__ __device__ inline void dt_inner_store(const DtPass& p, int64_t t, int tid, int nth, const uint8_t* sm,
                                               uint8_t* D, DtHist& h)
{
    const int64_t b0 = t * p.Lc;
    const int64_t len = (p.N - b0 < p.Lc) ? p.N - b0 : p.Lc;
    const int64_t s = b0 & 15, g0 = b0 - s, e = b0 + len;
    const int64_t nv = (s + len + 15) >> 4;
    for (int64_t v = tid; v < nv; v += nth) {
        const int64_t g = g0 + 16 * v;
        const uint8_t* src = sm + 16 * v;
        if (g >= b0 && g + 16 <= e) {
            *reinterpret_cast<uint4*>(D + g) = *reinterpret_cast<const uint4*>(src);
            if (p.last)
                for (int k = 0; k < 16; k++) dt_hist_add(h, src[k], p.hn);
        } else {
            for (int k = 0; k < 16; k++)
                if (g + k >= b0 && g + k < e) {
                    D[g + k] = src[k];
                    if (p.last) dt_hist_add(h, src[k], p.hn);
                }
        }
    }
}

// ------------------------------------------------------------ outer pass
// tile t = (o, c): rows r < p.rows start at rs_r = o*rows*S_in + r*S_in + c*Lc
// and hold len words; LDS row r is at r*P, word x of it at r*P + (rs_r & 15) + x
struct DtTile {
    int64_t rs0, len;
};
__host__ __device__ inline DtTile dt_outer_tile(const DtPass& p, int64_t t)
{
    const int64_t o = t / p.nch, c = t - o * p.nch;
    const int64_t x0 = c * p.Lc;
    DtTile T;
    T.rs0 = o * p.rows * p.S_in + x0;
    T.len = (p.S_in - x0 < p.Lc) ? p.S_in - x0 : p.Lc;
    return T;
}

__host__ __device__ inline void dt_outer_load(const DtPass& p, const DtTile& T, int tid, int nth, uint8_t* sm,
                                              uint8_t* shr, const uint8_t* D)
{
    const int64_t nslot = p.rows << p.pshift;
    const int64_t vmask = ((int64_t)1 << p.pshift) - 1;
    for (int64_t e = tid; e < nslot; e += nth) {
        const int64_t r = e >> p.pshift, v = e & vmask;
        const int64_t rs = T.rs0 + r * p.S_in;
        const int64_t s = rs & 15;
        if (v == 0) shr[r] = (uint8_t)s;
        if (v < ((s + T.len + 15) >> 4))
            *reinterpret_cast<uint4*>(sm + r * p.P + 16 * v) = *reinterpret_cast<const uint4*>(D + (rs - s) + 16 * v);
    }
}

__host__ __device__ inline void dt_outer_axis(const DtPass& p, const DtTile& T, int ax, int tid, int nth,
                                              uint8_t* sm, const uint8_t* shr)
{
    const uint32_t u = p.u[ax], mg = p.mg[ax];
    const int64_t nslot = (p.rows / p.q) << p.xs_shift;
    const int64_t xmask = ((int64_t)1 << p.xs_shift) - 1;
    for (int64_t e = tid; e < nslot; e += nth) {
        const int64_t x = e & xmask;
        if (x >= T.len) continue;
        const uint32_t fib = (uint32_t)(e >> p.xs_shift);
        const uint32_t hi = dt_div(fib, u, mg), lo = fib - hi * u;
        const int64_t r0 = (int64_t)hi * u * p.q + lo;
        int m = 255;
        for (int v = 0; v < p.q; v++) {
            const int64_t r = r0 + (int64_t)v * u;
            const int y = sm[r * p.P + shr[r] + x];
            m = y < m ? y : m;
        }
        const int m1 = m < 255 ? m + 1 : 255;
        for (int v = 0; v < p.q; v++) {
            const int64_t r = r0 + (int64_t)v * u;
            uint8_t* q8 = sm + r * p.P + shr[r] + x;
            const int y = *q8;
            *q8 = (uint8_t)(y < m1 ? y : m1);
        }
    }
}

__host__ __device__ inline void dt_outer_store(const DtPass& p, const DtTile& T, int tid, int nth,
                                               const uint8_t* sm, uint8_t* D, DtHist& h)
{
    const int64_t nslot = p.rows << p.pshift;
    const int64_t vmask = ((int64_t)1 << p.pshift) - 1;
    for (int64_t e = tid; e < nslot; e += nth) {
        const int64_t r = e >> p.pshift, v = e & vmask;
        const int64_t rs = T.rs0 + r * p.S_in;
        const int64_t s = rs & 15, re = rs + T.len;
        if (v >= ((s + T.len + 15) >> 4)) continue;
        const int64_t g = rs - s + 16 * v;
        const uint8_t* src = sm + r * p.P + 16 * v;
        if (g >= rs && g + 16 <= re) {
            *reinterpret_cast<uint4*>(D + g) = *reinterpret_cast<const uint4*>(src);
            if (p.last)
                for (int k = 0; k < 16; k++) dt_hist_add(h, src[k], p.hn);
        } else {
            for (int k = 0; k < 16; k++)
                if (g + k >= rs && g + k < re) {
                    D[g + k] = src[k];
                    if (p.last) dt_hist_add(h, src[k], p.hn);
                }
        }
    }
}

// ---------------------------------------------------------------- kernels
__device__ inline void dt_hist_flush(const DtHist& h, int hn, unsigned int* lh)
{
    for (int b = 0; b <= hn + 1; b++) {
        const uint32_t c = (uint32_t)((h.w[b >> 2] >> (16 * (b & 3))) & 0xFFFFu);
        if (c) atomicAdd(&lh[b], c);
    }
}

__global__ void __launch_bounds__(WG) k_dt_inner(const DtPass p, uint8_t* D, const uint32_t* cwb,
                                                 unsigned long long* hist)
{
    extern __shared__ __align__(16) uint8_t sm[];
    __shared__ unsigned int lh[DT_HB];
    if (threadIdx.x < DT_HB) lh[threadIdx.x] = 0;
    for (int64_t t = blockIdx.x; t < p.ntiles; t += gridDim.x) {
        DtHist h{};
        __syncthreads();
        dt_inner_load(p, t, threadIdx.x, WG, sm, D, cwb);
        __syncthreads();
        for (int ax = 0; ax < p.nax; ax++) {
            dt_inner_axis(p, t, ax, threadIdx.x, WG, sm);
            __syncthreads();
        }
        dt_inner_store(p, t, threadIdx.x, WG, sm, D, h);
        if (p.last) dt_hist_flush(h, p.hn, lh);
    }
    if (p.last) {
        __syncthreads();
        if (threadIdx.x < p.hn + 2 && lh[threadIdx.x])
            atomicAdd(&hist[threadIdx.x], (unsigned long long)lh[threadIdx.x]);
    }
}

__global__ void __launch_bounds__(WG) k_dt_outer(const DtPass p, uint8_t* D, unsigned long long* hist)
{
    extern __shared__ __align__(16) uint8_t sm[];
    __shared__ unsigned int lh[DT_HB];
    __shared__ uint8_t shr[1024];
    if (threadIdx.x < DT_HB) lh[threadIdx.x] = 0;
    for (int64_t t = blockIdx.x; t < p.ntiles; t += gridDim.x) {
        const DtTile T = dt_outer_tile(p, t);
        DtHist h{};
        __syncthreads();
        dt_outer_load(p, T, threadIdx.x, WG, sm, shr, D);
        __syncthreads();
        for (int ax = 0; ax < p.nax; ax++) {
            dt_outer_axis(p, T, ax, threadIdx.x, WG, sm, shr);
            __syncthreads();
        }
        dt_outer_store(p, T, threadIdx.x, WG, sm, D, h);
        if (p.last) dt_hist_flush(h, p.hn, lh);
    }
    if (p.last) {
        __syncthreads();
        if (threadIdx.x < p.hn + 2 && lh[threadIdx.x])
            atomicAdd(&hist[threadIdx.x], (unsigned long long)lh[threadIdx.x]);
    }
}

// codeword bitmap: set (or clear) the bit of every codeword
__global__ void __launch_bounds__(WG) k_cwb_mark(const Cell C, uint32_t* cwb, int set)
{
    for (int i = blockIdx.x * WG + threadIdx.x; i < C.M; i += gridDim.x * WG) {
        const int64_t w = C.cidx[i];
        const uint32_t b = 1u << (w & 31);
        if (set) atomicOr(&cwb[w >> 5], b);
        else atomicAnd(&cwb[w >> 5], ~b);
    }
}

// ------------------------------------------------------------- planning
// LDS bytes of a tile (the autotuner may change it)
constexpr int DT_LDS_DEFAULT = 32768;
constexpr int DT_LDS_MAX = 65536 - 2048;

inline void dt_fill_axes(DtPass& p, int q, int b, uint32_t unit)
{
    // axis k of the group (k = 0 outermost) has LDS stride q^(b-1-k) * unit
    p.nax = b;
    for (int k = 0; k < b; k++) {
        uint64_t u = unit;
        for (int i = 0; i < b - 1 - k; i++) u *= (uint64_t)q;
        p.u[k] = (uint32_t)u;
        p.mg[k] = u > 1 ? (uint32_t)((((uint64_t)1 << 32) / u) + 1) : 0u;
    }
}

// the passes for a cell; returns their number (0 on a bad budget)
inline int dt_plan(int q, int n, int64_t N, int budget, DtPass* out)
{
    if (budget < 4096 || budget > DT_LDS_MAX) return 0;
    // inner group: the largest q^a that fits, keeping >= 512 tiles on big spaces
    int a = 0;
    int64_t T = 1;
    while (a < n && T * q + 32 <= budget && (N / (T * q) >= 512 || N <= 64 * (int64_t)budget)) {
        T *= q;
        a++;
    }
    if (a == 0) { T = q; a = 1; }   // q <= 16 always fits
    int np = 0;
    const int m = n - a;
    // outer groups: at most bmax axes each, rows of >= 96 payload bytes
    int bmax = 0;
    for (int64_t B = q; B * 128 <= budget && bmax < m; B *= q) bmax++;
    if (m > 0 && bmax == 0) bmax = 1;
    const int ngroups = m > 0 ? (m + bmax - 1) / bmax : 0;
    {
        DtPass& p = out[np++];
        p = DtPass{};
        p.N = N; p.rows = 1; p.S_in = T; p.q = q; p.hn = n;
        // whole blocks of q^a words per tile, keeping >= 512 tiles where the space allows
        int64_t per = (budget - 32) / T;
        const int64_t per_par = N / (T * 512);
        if (per > per_par) per = per_par;
        if (per < 1) per = 1;
        p.Lc = per * T;
        p.ntiles = (N + p.Lc - 1) / p.Lc;
        p.P = (int32_t)(((p.Lc + 16 + 15) / 16) * 16);
        p.first = 1;
        dt_fill_axes(p, q, a, 1);
    }
    // groups from the inner side outwards: axes [j1-b, j1)
    int j1 = n - a;
    for (int g = 0; g < ngroups; g++) {
        const int left = ngroups - g;
        const int b = (j1 + left - 1) / left;   // spread the remaining axes evenly
        const int j0 = j1 - b;
        DtPass& p = out[np++];
        p = DtPass{};
        p.N = N; p.q = q; p.hn = n;
        int64_t rows = 1;
        for (int i = 0; i < b; i++) rows *= q;
        int64_t S_in = 1;
        for (int i = 0; i < n - j1; i++) S_in *= q;
        int64_t P = 16;
        while (rows * P * 2 <= budget && P < 4096) P *= 2;
        if (rows > 1024 || P < 64) return 0;
        int64_t Lc = P - 32;
        if (Lc > S_in) Lc = S_in;
        int ps = 0;
        while ((16 << ps) < P) ps++;
        int xs = 0;
        while (((int64_t)1 << xs) < Lc) xs++;
        int64_t outer = 1;
        for (int i = 0; i < j0; i++) outer *= q;
        p.rows = rows; p.S_in = S_in; p.Lc = Lc; p.P = (int32_t)P; p.pshift = ps; p.xs_shift = xs;
        p.nch = (S_in + Lc - 1) / Lc;
        p.ntiles = outer * p.nch;
        dt_fill_axes(p, q, b, 1);
        j1 = j0;
    }
    out[np - 1].last = 1;
    return np;
}

}  // namespace cov
