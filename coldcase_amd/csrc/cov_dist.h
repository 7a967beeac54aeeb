// cov_dist.h -- the uint8 axis-DP Hamming distance transform (kernel c).
//
// dist[w] = min over codewords c of d_H(w, c).  Start from D = 0 on codewords
// and "unreached" elsewhere; for each axis j, every fiber (the q words that
// differ only in digit j) becomes D[x] <- min(D[x], 1 + min over the fiber).
// After all n axes, in any order, D is the exact minimum distance (the R-fold
// dilation of cov/verify_independent.py:128-141 reads the same information as
// #{D <= r}).
//
// The n axes are grouped into a few passes over HBM instead of n.  A pass
// loads tiles of rows into LDS, applies the axes of its group there and
// writes the tiles back:
//   inner pass  rows are the innermost fibers (q consecutive words), each
//               padded to 4, 8 or 16 bytes; the innermost axis is applied
//               inside each row, the next a-1 axes across rows; the tile is
//               built from the codeword bitmap (one bit per word), so no
//               pass initialises D
//   outer pass  a group of b consecutive axes [j0, j0+b): a tile is q^b rows
//               of Lc contiguous words, the rows q^(n-j0-b) words apart
// and the last pass also histograms the distances it stores.
//
// Values are bytes <= 127 ("unreached" = 127; distances never exceed 16), so
// four or sixteen of them are handled per lane with carry-free integer
// arithmetic (dt_pmin, dt_pinc) on dwords.  Rows sit in LDS dword-aligned
// whatever their global alignment: loads and stores realign with alignbyte.
//
// The per-thread phases are __host__ __device__ functions so that a host build
// can run them thread by thread over plain arrays.
#pragma once
#include "cov_common.h"

namespace cov {

constexpr int DT_HB = MAXN + 2;      // histogram bins: 0..n, and unreached in bin n+1
constexpr uint32_t DT_INF = 0x7Fu;   // "not reached yet"
constexpr uint32_t DT_INF4 = 0x7F7F7F7Fu;

// one pass of the transform, planned on the host
struct DtPass {
    int64_t N;
    int64_t nrow;        // rows per tile
    int64_t S_row;       // words between consecutive rows of a tile
    int64_t S_tile;      // words between consecutive row blocks (nrow * S_row)
    int64_t Lc;          // words per row chunk (inner: q)
    int64_t nch;         // chunks per row (inner: 1)
    int64_t ntiles;
    int32_t Pr;          // LDS bytes per row: inner 4 * 2^ds, outer Lc + 16
    int32_t ds;          // log2 of the dword slots per row (inner: 1, 2, 4; outer: Lc / 4)
    int32_t q, nax, first, last, hn, inner;
    uint32_t u[MAXN];    // per row axis: stride in rows
    uint32_t mg[MAXN];   // floor(2^32 / u) + 1, for u > 1
    uint32_t mq;         // floor(2^32 / q) + 1
    uint32_t nk;         // outer store: dword slots per row, Lc / 4 + 1
    uint32_t mk;         // floor(2^32 / nk) + 1
};

// exact floor(f / d) for f * d < 2^32, with m = floor(2^32 / d) + 1 (d > 1)
__host__ __device__ inline uint32_t dt_div(uint32_t f, uint32_t d, uint32_t m)
{
#ifdef __HIP_DEVICE_COMPILE__
    return d == 1 ? f : __umulhi(f, m);
#else
    return d == 1 ? f : (uint32_t)(((uint64_t)f * m) >> 32);
#endif
}

// bytes s..s+3 of the little-endian pair (hi:lo)
__host__ __device__ inline uint32_t dt_align(uint32_t hi, uint32_t lo, uint32_t s)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_alignbyte(hi, lo, s);
#else
    return (uint32_t)(((((uint64_t)hi) << 32) | lo) >> (8 * s));
#endif
}

// per-byte min of bytes <= 127: byte i of (a | 0x80) - b is 128 + a_i - b_i,
// which never borrows, and its top bit says a_i >= b_i
__host__ __device__ inline uint32_t dt_pmin(uint32_t a, uint32_t b)
{
    const uint32_t d = (a | 0x80808080u) - b;
    const uint32_t ge = ((d >> 7) & 0x01010101u) * 0xFFu;
    return (b & ge) | (a & ~ge);
}
// per-byte min(a + 1, 127) of bytes <= 127
__host__ __device__ inline uint32_t dt_pinc(uint32_t a)
{
    const uint32_t y = a + 0x01010101u;
    return y - ((y >> 7) & 0x01010101u);
}
__host__ __device__ inline uint4 dt_pmin4(uint4 a, uint4 b)
{
    return make_uint4(dt_pmin(a.x, b.x), dt_pmin(a.y, b.y), dt_pmin(a.z, b.z), dt_pmin(a.w, b.w));
}
__host__ __device__ inline uint4 dt_pinc4(uint4 a)
{
    return make_uint4(dt_pinc(a.x), dt_pinc(a.y), dt_pinc(a.z), dt_pinc(a.w));
}

// packed 16-bit histogram counters: bin x in word x >> 2, field x & 3
struct DtHist {
    uint64_t w[(DT_HB + 3) / 4];
};
__host__ __device__ inline void dt_hist_add(DtHist& h, uint32_t x, int n)
{
    const int b = (int)x <= n ? (int)x : n + 1;
    const uint64_t one = (uint64_t)1 << (16 * (b & 3));
#pragma unroll
    for (int k = 0; k < (DT_HB + 3) / 4; k++) h.w[k] += (k == (b >> 2)) ? one : 0;
}
__host__ __device__ inline void dt_hist_add4(DtHist& h, uint32_t v, int n)
{
#pragma unroll
    for (int k = 0; k < 4; k++) dt_hist_add(h, (v >> (8 * k)) & 0xFFu, n);
}

// tile t: row r holds the words [gs0 + r * S_row, + len)
struct DtTile {
    int64_t gs0, len;
};
__host__ __device__ inline DtTile dt_tile(const DtPass& p, int64_t t)
{
    const int64_t o = t / p.nch, c = t - o * p.nch;
    const int64_t x0 = c * p.Lc;
    DtTile T;
    T.gs0 = o * p.S_tile + x0;
    T.len = (p.S_row - x0 < p.Lc) ? p.S_row - x0 : p.Lc;
    return T;
}

// ------------------------------------------------------------------ load
// LDS dword j of row r = global bytes [gs + 4j, +4); bytes past the row's
// end, and dwords past them up to the row's slots, are unreached
__host__ __device__ inline void dt_load(const DtPass& p, const DtTile& T, int tid, int nth, uint32_t* sm,
                                        const uint32_t* D32, const uint32_t* cwb)
{
    const int64_t ndp = (int64_t)1 << p.ds;
    const int64_t nd = (T.len + 3) >> 2;
    const int64_t nslot = p.nrow << p.ds;
    for (int64_t e = tid; e < nslot; e += nth) {
        const int64_t r = e >> p.ds, j = e & (ndp - 1);
        uint32_t v = DT_INF4;
        if (j < nd) {
            const int64_t a = T.gs0 + r * p.S_row + 4 * j;
            if (p.first) {
                const uint64_t w = (uint64_t)cwb[a >> 5] | ((uint64_t)cwb[(a >> 5) + 1] << 32);
                const uint32_t bits = (uint32_t)(w >> (a & 31)) & 0xFu;
                v = 0;
                for (int k = 0; k < 4; k++) v |= (((bits >> k) & 1u) ? 0u : DT_INF) << (8 * k);
            } else {
                v = dt_align(D32[(a >> 2) + 1], D32[a >> 2], (uint32_t)(a & 3));
            }
            const int64_t keep = T.len - 4 * j;
            if (keep < 4) {
                const uint32_t mask = (1u << (8 * keep)) - 1u;
                v = (v & mask) | (DT_INF4 & ~mask);
            }
        }
        sm[((r * p.Pr) >> 2) + j] = v;
    }
}

// ----------------------------------------------------- innermost axis
// inner pass: each row is one fiber of the innermost axis (q bytes, the rest
// of its 4, 8 or 16 bytes unreached)
__host__ __device__ inline void dt_inrow(const DtPass& p, int tid, int nth, uint32_t* sm)
{
    const int nd = 1 << p.ds;
    for (int64_t r = tid; r < p.nrow; r += nth) {
        uint32_t* row = sm + ((r * p.Pr) >> 2);
        uint32_t d[4];
        uint32_t m = DT_INF4;
        for (int j = 0; j < 4; j++) {
            d[j] = j < nd ? row[j] : DT_INF4;
            m = dt_pmin(m, d[j]);
        }
        m = dt_pmin(m, m >> 16);
        m = dt_pmin(m, m >> 8);
        const uint32_t mb = m & 0xFFu;
        const uint32_t m1 = (mb < DT_INF ? mb + 1 : DT_INF) * 0x01010101u;
        for (int j = 0; j < 4; j++)
            if (j < nd) row[j] = dt_pmin(d[j], m1);
    }
}

// ------------------------------------------------------------ row axes
// fibers across rows: q rows u apart, the same column; inner rows by dwords,
// outer rows by 16-byte vectors
__host__ __device__ inline void dt_axis(const DtPass& p, const DtTile& T, int ax, int tid, int nth, uint32_t* sm)
{
    const uint32_t u = p.u[ax], mg = p.mg[ax];
    const int64_t nf = p.nrow / p.q;
    if (p.inner) {
        const int64_t ndp = (int64_t)1 << p.ds;
        const int64_t nslot = nf << p.ds;
        for (int64_t e = tid; e < nslot; e += nth) {
            const int64_t j = e & (ndp - 1);
            const uint32_t f = (uint32_t)(e >> p.ds);
            const uint32_t hi = dt_div(f, u, mg), lo = f - hi * u;
            const int64_t r0 = (int64_t)hi * u * p.q + lo;
            uint32_t m = DT_INF4;
            for (int v = 0; v < p.q; v++) m = dt_pmin(m, sm[(((r0 + (int64_t)v * u) * p.Pr) >> 2) + j]);
            const uint32_t m1 = dt_pinc(m);
            for (int v = 0; v < p.q; v++) {
                uint32_t* x = sm + (((r0 + (int64_t)v * u) * p.Pr) >> 2) + j;
                *x = dt_pmin(*x, m1);
            }
        }
    } else {
        const int vs = p.ds - 2;                      // 16-byte vectors per row: 1 << vs
        const int64_t nvec = (T.len + 15) >> 4;
        const int64_t nslot = nf << vs;
        uint4* s4 = reinterpret_cast<uint4*>(sm);
        for (int64_t e = tid; e < nslot; e += nth) {
            const int64_t c = e & (((int64_t)1 << vs) - 1);
            if (c >= nvec) continue;
            const uint32_t f = (uint32_t)(e >> vs);
            const uint32_t hi = dt_div(f, u, mg), lo = f - hi * u;
            const int64_t r0 = (int64_t)hi * u * p.q + lo;
            uint4 m = make_uint4(DT_INF4, DT_INF4, DT_INF4, DT_INF4);
            for (int v = 0; v < p.q; v++) m = dt_pmin4(m, s4[(((r0 + (int64_t)v * u) * p.Pr) >> 4) + c]);
            const uint4 m1 = dt_pinc4(m);
            for (int v = 0; v < p.q; v++) {
                uint4* x = s4 + (((r0 + (int64_t)v * u) * p.Pr) >> 4) + c;
                *x = dt_pmin4(*x, m1);
            }
        }
    }
}

// ----------------------------------------------------------------- store
__host__ __device__ inline void dt_store(const DtPass& p, const DtTile& T, int tid, int nth, const uint32_t* sm,
                                         uint8_t* D, DtHist& h)
{
    const uint8_t* s8 = reinterpret_cast<const uint8_t*>(sm);
    uint32_t* D32 = reinterpret_cast<uint32_t*>(D);
    if (p.inner) {
        // the tile is the contiguous range [gs0, gs0 + S_tile); word x sits at
        // row (x - gs0) / q, column (x - gs0) % q
        const int64_t b0 = T.gs0, e0 = T.gs0 + p.S_tile;
        const int64_t g0 = b0 & ~(int64_t)3;
        const int64_t nw = (e0 - g0 + 3) >> 2;
        for (int64_t k = tid; k < nw; k += nth) {
            const int64_t g = g0 + 4 * k;
            uint32_t v = 0;
            for (int b = 0; b < 4; b++) {
                const int64_t x = g + b;
                if (x < b0 || x >= e0) continue;
                const uint32_t off = (uint32_t)(x - b0);
                const uint32_t row = dt_div(off, (uint32_t)p.q, p.mq), col = off - row * (uint32_t)p.q;
                v |= (uint32_t)s8[(int64_t)row * p.Pr + col] << (8 * b);
            }
            if (g >= b0 && g + 4 <= e0) {
                D32[g >> 2] = v;
                if (p.last) dt_hist_add4(h, v, p.hn);
            } else {
                for (int b = 0; b < 4; b++) {
                    const int64_t x = g + b;
                    if (x < b0 || x >= e0) continue;
                    const uint32_t y = (v >> (8 * b)) & 0xFFu;
                    D[x] = (uint8_t)y;
                    if (p.last) dt_hist_add(h, y, p.hn);
                }
            }
        }
    } else {
        // row r: global dwords from (gs & ~3) covering [gs, gs + len)
        const int64_t nslot = p.nrow * (int64_t)p.nk;
        for (int64_t e = tid; e < nslot; e += nth) {
            const uint32_t r = dt_div((uint32_t)e, p.nk, p.mk);
            const int64_t k = e - (int64_t)r * p.nk;
            const int64_t gs = T.gs0 + (int64_t)r * p.S_row, ge = gs + T.len;
            const int64_t g = (gs & ~(int64_t)3) + 4 * k;
            if (g >= ge) continue;
            const uint32_t* row = sm + (((int64_t)r * p.Pr) >> 2);
            if (g >= gs && g + 4 <= ge) {
                const int64_t i0 = g - gs;
                const uint32_t s = (uint32_t)(i0 & 3);
                const uint32_t v = dt_align(s ? row[(i0 >> 2) + 1] : 0u, row[i0 >> 2], s);
                D32[g >> 2] = v;
                if (p.last) dt_hist_add4(h, v, p.hn);
            } else {
                const uint8_t* rb = s8 + (int64_t)r * p.Pr;
                for (int b = 0; b < 4; b++) {
                    const int64_t x = g + b;
                    if (x < gs || x >= ge) continue;
                    const uint32_t y = rb[x - gs];
                    D[x] = (uint8_t)y;
                    if (p.last) dt_hist_add(h, y, p.hn);
                }
            }
        }
    }
}

// ---------------------------------------------------------------- kernels
__device__ inline void dt_hist_flush(const DtHist& h, int hn, unsigned int* lh)
{
#pragma unroll
    for (int b = 0; b < DT_HB; b++) {
        const uint32_t c = (uint32_t)((h.w[b >> 2] >> (16 * (b & 3))) & 0xFFFFu);
        if (b <= hn + 1 && c) atomicAdd(&lh[b], c);
    }
}

__global__ void __launch_bounds__(WG) k_dt_pass(const DtPass p, uint8_t* D, const uint32_t* cwb,
                                                unsigned long long* hist)
{
    extern __shared__ __align__(16) uint32_t sm[];
    __shared__ unsigned int lh[DT_HB];
    if (threadIdx.x < DT_HB) lh[threadIdx.x] = 0;
    const uint32_t* D32 = reinterpret_cast<const uint32_t*>(D);
    for (int64_t t = blockIdx.x; t < p.ntiles; t += gridDim.x) {
        const DtTile T = dt_tile(p, t);
        DtHist h{};
        __syncthreads();
        dt_load(p, T, threadIdx.x, WG, sm, D32, cwb);
        __syncthreads();
        if (p.inner) {
            dt_inrow(p, threadIdx.x, WG, sm);
            __syncthreads();
        }
        for (int ax = 0; ax < p.nax; ax++) {
            dt_axis(p, T, ax, threadIdx.x, WG, sm);
            __syncthreads();
        }
        dt_store(p, T, threadIdx.x, WG, sm, D, h);
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
constexpr int DT_LDS_MAX = 65536 - 1024;

inline uint32_t dt_magic(uint64_t d)
{
    return d > 1 ? (uint32_t)((((uint64_t)1 << 32) / d) + 1) : 0u;
}

// row axes of a group of b axes (k = 0 outermost): stride q^(b-1-k) rows
inline void dt_fill_axes(DtPass& p, int q, int b)
{
    p.nax = b;
    for (int k = 0; k < b; k++) {
        uint64_t u = 1;
        for (int i = 0; i < b - 1 - k; i++) u *= (uint64_t)q;
        p.u[k] = (uint32_t)u;
        p.mg[k] = dt_magic(u);
    }
}

inline int64_t dt_ipow(int q, int e)
{
    int64_t r = 1;
    for (int i = 0; i < e; i++) r *= q;
    return r;
}

// the passes for a cell; returns their number (0 on a bad budget)
inline int dt_plan(int q, int n, int64_t N, int budget, DtPass* out)
{
    if (budget < 4096 || budget > DT_LDS_MAX || q < 2 || q > 16 || n < 1 || n > MAXN) return 0;
    int np = 0;
    // inner pass: rows of one innermost fiber, padded to 4, 8 or 16 bytes;
    // a axes (the innermost + a-1 across rows) in blocks of q^(a-1) rows,
    // keeping >= 512 tiles on large spaces
    const int ndp = q <= 4 ? 1 : (q <= 8 ? 2 : 4);
    const int pr_in = 4 * ndp;
    int a = 1;
    while (a < n && dt_ipow(q, a) * pr_in <= budget &&
           (N / dt_ipow(q, a + 1) >= 512 || N <= 64 * (int64_t)budget))
        a++;
    {
        DtPass& p = out[np++];
        p = DtPass{};
        const int64_t blk = dt_ipow(q, a - 1);       // rows per block of q^a words
        int64_t per = budget / (blk * pr_in);
        const int64_t per_par = N / (blk * q * 512);
        if (per > per_par) per = per_par;
        if (per < 1) per = 1;
        while (per > 1 && (N / q) % (per * blk)) per--;
        p.N = N; p.q = q; p.hn = n; p.inner = 1; p.first = 1;
        p.nrow = per * blk;
        p.S_row = q; p.Lc = q; p.nch = 1;
        p.S_tile = p.nrow * q;
        p.ntiles = N / p.S_tile;
        p.Pr = pr_in;
        p.ds = ndp == 1 ? 0 : (ndp == 2 ? 1 : 2);
        p.mq = dt_magic((uint64_t)q);
        dt_fill_axes(p, q, a - 1);
        if (N % p.S_tile) return 0;
    }
    // outer groups: at most bmax axes each, rows of >= 64 bytes
    const int m = n - a;
    int bmax = 0;
    for (int64_t B = q; B * (64 + 16) <= budget && bmax < m; B *= q) bmax++;
    if (m > 0 && bmax == 0) return 0;
    const int ngroups = m > 0 ? (m + bmax - 1) / bmax : 0;
    int j1 = n - a;
    for (int g = 0; g < ngroups; g++) {
        const int left = ngroups - g;
        const int b = (j1 + left - 1) / left;   // spread the remaining axes evenly
        const int j0 = j1 - b;
        DtPass& p = out[np++];
        p = DtPass{};
        const int64_t rows = dt_ipow(q, b);
        const int64_t S_row = dt_ipow(q, n - j1);
        if (rows > 4096) return 0;
        // Lc = 16 * 2^k: the largest that fits, no longer than the row needs
        int64_t Lc = 64;
        while (rows * (2 * Lc + 16) <= budget && Lc < S_row && Lc < 4096) Lc *= 2;
        int lg = 0;
        while (((int64_t)4 << lg) < Lc) lg++;     // Lc / 4 = 2^lg dword slots
        p.N = N; p.q = q; p.hn = n;
        p.nrow = rows; p.S_row = S_row; p.S_tile = rows * S_row; p.Lc = Lc;
        p.nch = (S_row + Lc - 1) / Lc;
        p.ntiles = dt_ipow(q, j0) * p.nch;
        p.Pr = (int32_t)(Lc + 16);
        p.ds = lg;
        p.nk = (uint32_t)(Lc / 4 + 1);
        p.mk = dt_magic(p.nk);
        if ((uint64_t)rows * p.nk * p.nk >= ((uint64_t)1 << 32)) return 0;   // dt_div exactness
        dt_fill_axes(p, q, b);
        j1 = j0;
    }
    out[np - 1].last = 1;
    return np;
}

// LDS bytes a pass's tile needs
inline int64_t dt_lds_bytes(const DtPass& p)
{
    return p.nrow * p.Pr;
}

}  // namespace cov
