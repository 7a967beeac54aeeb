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
#include <algorithm>
#include "cov_common.h"

namespace cov {

constexpr int DT_HB = MAXN + 2;      // histogram bins: 0..n, and unreached in bin n+1
// the last pass's workgroups add into DT_COPIES copies of the bins, copy
// blockIdx % DT_COPIES, each copy on its own 256-byte line (atomics to one
// address from every workgroup would serialise); the host folds the copies
constexpr int DT_COPIES = 64;
constexpr int DT_CSTRIDE = 32;       // uint64 per copy
constexpr int DT_SMAX = 6;           // axes a group of at most 64 words can span (q = 2)
constexpr uint32_t DT_INF = 0x7Fu;   // "not reached yet"
constexpr uint32_t DT_INF4 = 0x7F7F7F7Fu;

// one pass of the transform, planned on the host.  LDS holds, per tile, the
// rows realigned to 16 bytes ("aligned", at offset 0) and the 16-byte global
// chunks they were read from ("raw", at raw_off).
struct DtPass {
    int64_t N;
    int64_t nrow;        // rows per tile
    int64_t S_row;       // words between consecutive rows (inner: q)
    int64_t S_tile;      // words between consecutive row blocks = nrow * S_row
    int64_t Lc;          // words per row chunk (inner: q)
    int64_t nch;         // chunks per row (inner: 1)
    int64_t ntiles;
    int64_t lds;         // LDS bytes of a tile
    int32_t PA;          // aligned bytes per row: inner W = 4, 8 or 16; outer Lc + 16
    int32_t PR;          // outer: raw bytes per row, Lc + 32
    int32_t raw_off;     // LDS offset of the raw chunks
    int32_t nkr;         // outer: raw chunk slots per row, Lc / 16 + 2
    uint32_t mkr;        // magic of nkr
    int32_t nvp;         // outer: 16-byte aligned chunks per row, Lc / 16
    uint32_t mvp;        // magic of nvp
    uint32_t nda;        // dwords per aligned row, PA / 4
    uint32_t mda;        // magic of nda
    int32_t q, nax, first, last, hn, inner;
    int32_t full;        // last pass: 1 = full histogram, 0 = #{values <= rcut} and the maximum
    int32_t rcut;
    uint32_t u[MAXN];    // per row axis: stride in rows
    uint32_t mg[MAXN];   // magic of u
    uint32_t mq;         // magic of q
    // column passes (no LDS): one thread per dword column of a row block
    int32_t col;         // 1: a column pass of nax axes, nrow = q^nax rows
    uint32_t ncol;       // dword columns per row, ceil(S_row / 4)
    uint32_t nthr;       // threads: row blocks * ncol
    uint32_t nbytes;     // bytes the buffer descriptor covers (N + 64)
    // the seed pass (first pass, from the codeword bitmap; rows are groups)
    int32_t seed;        // 1: the seed pass of ns + nax axes
    int32_t ns;          // innermost axes decided inside a group from its bits
    int32_t G;           // words per group, q^ns <= 64
    int32_t P;           // LDS bytes per group row: G rounded up to 4
    uint32_t mG;         // magic of G
    uint64_t sbase[DT_SMAX]; // digit j < ns of a group: the words whose digit j is 0
    // the seed pass's closed form for tiles with few codewords
    int32_t sk;          // most codewords a tile may hold for it (0: never)
    int32_t Ls;          // LDS bytes per codeword's column table
    int32_t hoff;        // LDS offset of the row tables
    int32_t Hs;          // LDS bytes per codeword's row table, nrow + 1 rounded up to 4
    int32_t Pd, Rp;      // the store's dword period G / gcd(G, 4) and the rows it spans, 4 / gcd(G, 4)
    uint32_t rb;         // group rows per block of q^a words, q^nax
    uint32_t mrb;        // magic of rb
    // two-stage column passes (rows through LDS): nrow = q^(b1 + b2) rows,
    // tiles of Wd dword columns
    int32_t tcol;        // 1: a two-stage column pass of nax = b1 + b2 axes
    int32_t b1;          // the stage-A axes (the row index's low digits)
    int32_t Wd;          // dword columns per tile
    int32_t nchunk;      // tiles per row block, ceil(S_row / (4 Wd))
    int32_t nt;          // threads per workgroup
    int32_t grid;        // workgroups
    uint32_t mR;         // magic of nrow
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

// 16 bytes starting at byte s (0..15) of the 32 bytes a:b
__host__ __device__ inline uint4 dt_funnel(uint4 a, uint4 b, uint32_t s)
{
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    const uint32_t sb = s & 3, sd = s >> 2;
    uint32_t y[7];
#pragma unroll
    for (int i = 0; i < 7; i++) y[i] = dt_align(w[i + 1], w[i], sb);
    uint32_t o[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
        o[j] = sd == 0 ? y[j] : (sd == 1 ? y[j + 1] : (sd == 2 ? y[j + 2] : y[j + 3]));
    return make_uint4(o[0], o[1], o[2], o[3]);
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
// bytes [k, 16) of x set to unreached
__host__ __device__ inline uint4 dt_keep4(uint4 x, int32_t k)
{
    uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int32_t kk = k - 4 * j;
        if (kk <= 0) w[j] = DT_INF4;
        else if (kk < 4) {
            const uint32_t m = (1u << (8 * kk)) - 1u;
            w[j] = (w[j] & m) | (DT_INF4 & ~m);
        }
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
}

__host__ __device__ inline uint4 dt_bits_to_bytes(uint32_t bits16)
{
    // bit k set (a codeword) -> byte k = 0, else unreached
    uint32_t w[4];
    for (int k = 0; k < 4; k++) {
        uint32_t x = 0;
        for (int b = 0; b < 4; b++)
            x |= (((bits16 >> (4 * k + b)) & 1u) ? 0u : DT_INF) << (8 * b);
        w[k] = x;
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
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
// Index arithmetic inside a tile is 32-bit; only global addresses are 64-bit.
// Row r of an outer tile starts at global word gs0 + r S_row, which is
// (gs0 + r S_row) & 15 = (gs0 + r (S_row & 15)) & 15 bytes past a 16-byte
// boundary.
__host__ __device__ inline uint32_t dt_row_shift(const DtPass& p, const DtTile& T, uint32_t r)
{
    return ((uint32_t)T.gs0 + r * (uint32_t)p.S_row) & 15u;
}

// the 16-byte global chunks under the tile, as they are, into the raw area:
// inner: the contiguous range [gs0, gs0 + S_tile); outer: each row's range
__host__ __device__ inline void dt_load(const DtPass& p, const DtTile& T, int tid, int nth, uint8_t* sm,
                                        const uint4* D4, const uint32_t* cwb)
{
    uint4* raw = reinterpret_cast<uint4*>(sm + p.raw_off);
    if (p.inner) {
        const uint32_t s = (uint32_t)T.gs0 & 15u;
        const int64_t c0 = (T.gs0 - s) >> 4;               // first chunk
        const uint32_t nk = (s + (uint32_t)p.S_tile + 15) >> 4;
        for (uint32_t k = tid; k < nk; k += nth) {
            if (p.first) {
                const int64_t A = (c0 + k) << 4;
                raw[k] = dt_bits_to_bytes((cwb[A >> 5] >> (A & 31)) & 0xFFFFu);
            } else {
                raw[k] = D4[c0 + k];
            }
        }
    } else {
        const uint32_t nslot = (uint32_t)p.nrow * (uint32_t)p.nkr;
        const uint32_t len = (uint32_t)T.len, prc = (uint32_t)p.PR >> 4;
        for (uint32_t e = tid; e < nslot; e += nth) {
            const uint32_t r = dt_div(e, (uint32_t)p.nkr, p.mkr);
            const uint32_t k = e - r * (uint32_t)p.nkr;
            const uint32_t s = dt_row_shift(p, T, r);
            if (k < ((s + len + 15) >> 4)) {
                const int64_t c0 = (T.gs0 + (int64_t)r * p.S_row - s) >> 4;
                raw[r * prc + k] = D4[c0 + k];
            }
        }
    }
}

// ---------------------------------------------------------------- realign
// raw chunks -> rows starting at 16-byte boundaries; bytes past a row's end
// are unreached.  inner: row r = tile bytes [r q, r q + q) padded to PA
__host__ __device__ inline void dt_realign(const DtPass& p, const DtTile& T, int tid, int nth, uint8_t* sm)
{
    if (p.inner) {
        const uint32_t* raw = reinterpret_cast<const uint32_t*>(sm + p.raw_off);
        uint32_t* al = reinterpret_cast<uint32_t*>(sm);
        const uint32_t s = (uint32_t)T.gs0 & 15u;
        const uint32_t nd = p.nda, sh = nd == 1 ? 0 : (nd == 2 ? 1 : 2), q = (uint32_t)p.q;
        const uint32_t nslot = (uint32_t)p.nrow << sh;
        for (uint32_t e = tid; e < nslot; e += nth) {
            const uint32_t r = e >> sh, j = e & (nd - 1);
            const uint32_t o = s + r * q + 4 * j;             // raw byte of row byte 4j
            uint32_t v = dt_align(raw[(o >> 2) + 1], raw[o >> 2], o & 3);
            const int32_t keep = (int32_t)q - (int32_t)(4 * j);
            if (keep < 4) {
                const uint32_t m = keep > 0 ? (1u << (8 * keep)) - 1u : 0u;
                v = (v & m) | (DT_INF4 & ~m);
            }
            al[e] = v;
        }
    } else {
        const uint4* raw = reinterpret_cast<const uint4*>(sm + p.raw_off);
        uint4* al = reinterpret_cast<uint4*>(sm);
        const uint32_t na = (uint32_t)p.nvp + 1, prc = (uint32_t)p.PR >> 4, pac = (uint32_t)p.PA >> 4;
        const uint32_t nslot = (uint32_t)p.nrow * na;
        const uint32_t ma = (uint32_t)((((uint64_t)1 << 32) / na) + 1);
        const int32_t len = (int32_t)T.len;
        for (uint32_t e = tid; e < nslot; e += nth) {
            const uint32_t r = dt_div(e, na, na > 1 ? ma : 0u);
            const uint32_t c = e - r * na;
            const uint4* rr = raw + r * prc;
            const uint4 x = dt_funnel(rr[c], rr[c + 1], dt_row_shift(p, T, r));
            al[r * pac + c] = dt_keep4(x, len - 16 * (int32_t)c);
        }
    }
}

// ----------------------------------------------------- innermost axis
// inner pass: each row is one fiber of the innermost axis (q bytes, the rest
// of its 4, 8 or 16 bytes unreached)
__host__ __device__ inline void dt_inrow(const DtPass& p, int tid, int nth, uint8_t* sm)
{
    const uint32_t nd = p.nda;
    uint32_t* al = reinterpret_cast<uint32_t*>(sm);
    for (uint32_t r = tid; r < (uint32_t)p.nrow; r += nth) {
        uint32_t* row = al + r * nd;
        uint32_t d[4];
        uint32_t m = DT_INF4;
        for (uint32_t j = 0; j < 4; j++) {
            d[j] = j < nd ? row[j] : DT_INF4;
            m = dt_pmin(m, d[j]);
        }
        m = dt_pmin(m, m >> 16);
        m = dt_pmin(m, m >> 8);
        const uint32_t mb = m & 0xFFu;
        const uint32_t m1 = (mb < DT_INF ? mb + 1 : DT_INF) * 0x01010101u;
        for (uint32_t j = 0; j < 4; j++)
            if (j < nd) row[j] = dt_pmin(d[j], m1);
    }
}

// ------------------------------------------------------------ row axes
// fibers across rows: q rows u apart, the same column; inner rows by dwords,
// outer rows by 16-byte chunks
__host__ __device__ inline void dt_axis(const DtPass& p, const DtTile& T, int ax, int tid, int nth, uint8_t* sm)
{
    const uint32_t u = p.u[ax], mg = p.mg[ax], q = (uint32_t)p.q;
    const uint32_t nf = (uint32_t)p.nrow / q;
    if (p.inner) {
        uint32_t* al = reinterpret_cast<uint32_t*>(sm);
        const uint32_t nd = p.nda, sh = nd == 1 ? 0 : (nd == 2 ? 1 : 2);
        const uint32_t nslot = nf << sh, step = u * nd;
        for (uint32_t e = tid; e < nslot; e += nth) {
            const uint32_t j = e & (nd - 1), f = e >> sh;
            const uint32_t hi = dt_div(f, u, mg), lo = f - hi * u;
            uint32_t* x = al + (hi * u * q + lo) * nd + j;
            uint32_t m = DT_INF4;
            for (uint32_t v = 0; v < q; v++) m = dt_pmin(m, x[v * step]);
            const uint32_t m1 = dt_pinc(m);
            for (uint32_t v = 0; v < q; v++) x[v * step] = dt_pmin(x[v * step], m1);
        }
    } else {
        uint4* al = reinterpret_cast<uint4*>(sm);
        const uint32_t pac = (uint32_t)p.PA >> 4, nvp = (uint32_t)p.nvp;
        const uint32_t nvec = ((uint32_t)T.len + 15) >> 4;
        const uint32_t nslot = nf * nvp, step = u * pac;
        for (uint32_t e = tid; e < nslot; e += nth) {
            const uint32_t f = dt_div(e, nvp, p.mvp);
            const uint32_t c = e - f * nvp;
            if (c >= nvec) continue;
            const uint32_t hi = dt_div(f, u, mg), lo = f - hi * u;
            uint4* x = al + (hi * u * q + lo) * pac + c;
            uint4 m = make_uint4(DT_INF4, DT_INF4, DT_INF4, DT_INF4);
            for (uint32_t v = 0; v < q; v++) m = dt_pmin4(m, x[v * step]);
            const uint4 m1 = dt_pinc4(m);
            for (uint32_t v = 0; v < q; v++) x[v * step] = dt_pmin4(x[v * step], m1);
        }
    }
}

// ----------------------------------------------------------------- store
// inner: tile byte b sits at row b / q, column b % q of the padded rows
__host__ __device__ inline uint32_t dt_inner_byte(const DtPass& p, const uint8_t* sm, uint32_t b)
{
    const uint32_t r = dt_div(b, (uint32_t)p.q, p.mq);
    return sm[r * (uint32_t)p.PA + (b - r * (uint32_t)p.q)];
}
// the 4 tile bytes [b, b + 4) (q >= 4: at most two rows)
__host__ __device__ inline uint32_t dt_inner_dword(const DtPass& p, const uint8_t* sm, uint32_t b)
{
    const uint32_t* al = reinterpret_cast<const uint32_t*>(sm);
    const uint32_t r = dt_div(b, (uint32_t)p.q, p.mq), c = b - r * (uint32_t)p.q;
    const uint32_t w0 = r * p.nda + (c >> 2);
    uint32_t v = dt_align(al[w0 + 1], al[w0], c & 3);
    const uint32_t k = (uint32_t)p.q - c;                  // bytes left in row r
    if (k < 4) {
        const uint32_t m = (1u << (8 * k)) - 1u;
        v = (v & m) | (al[(r + 1) * p.nda] << (8 * k));
    }
    return v;
}

__host__ __device__ inline void dt_store(const DtPass& p, const DtTile& T, int tid, int nth, const uint8_t* sm,
                                         uint8_t* D)
{
    uint4* D4 = reinterpret_cast<uint4*>(D);
    if (p.inner) {
        const int64_t b0 = T.gs0, e0 = T.gs0 + p.S_tile;
        const uint32_t s = (uint32_t)b0 & 15u;
        const uint32_t head = (16 - s) & 15u;                   // bytes before the first whole chunk
        const uint32_t tot = (uint32_t)p.S_tile;
        const uint32_t nk = tot > head ? (tot - head) >> 4 : 0u; // whole chunks
        const int64_t c0 = (b0 + head) >> 4;
        for (uint32_t k = tid; k < nk; k += nth) {
            const uint32_t b = head + 16 * k;
            uint32_t w[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (p.q >= 4) {
                    w[j] = dt_inner_dword(p, sm, b + 4 * j);
                } else {
                    w[j] = 0;
                    for (uint32_t i = 0; i < 4; i++) w[j] |= dt_inner_byte(p, sm, b + 4 * j + i) << (8 * i);
                }
            }
            D4[c0 + k] = make_uint4(w[0], w[1], w[2], w[3]);
        }
        // the bytes outside whole chunks, at both ends
        if (tid < 2) {
            const uint32_t tail0 = head + 16 * nk;
            const uint32_t x0 = tid == 0 ? 0u : (tail0 > head ? tail0 : head);
            const uint32_t x1 = tid == 0 ? (head < tot ? head : tot) : tot;
            for (uint32_t x = x0; x < x1; x++) D[b0 + x] = (uint8_t)dt_inner_byte(p, sm, x);
        }
        (void)e0;
    } else {
        const uint4* al = reinterpret_cast<const uint4*>(sm);
        const uint32_t pac = (uint32_t)p.PA >> 4, len = (uint32_t)T.len;
        const uint32_t nslot = (uint32_t)p.nrow * (uint32_t)p.nkr;
        for (uint32_t e = tid; e < nslot; e += nth) {
            const uint32_t r = dt_div(e, (uint32_t)p.nkr, p.mkr);
            const uint32_t k = e - r * (uint32_t)p.nkr;
            const uint32_t s = dt_row_shift(p, T, r);
            const uint32_t o = ((16 - s) & 15u) + 16 * k;       // row byte of the k-th whole chunk
            if (o + 16 > len) continue;
            const uint4* rr = al + r * pac;
            const int64_t g = (T.gs0 + (int64_t)r * p.S_row + o) >> 4;
            D4[g] = dt_funnel(rr[o >> 4], rr[(o >> 4) + 1], o & 15);
        }
        // each row's bytes outside whole chunks: up to 15 at the head and 15
        // at the tail, one (row, byte) per thread
        const uint32_t nb = (uint32_t)p.nrow * 32;
        for (uint32_t e = tid; e < nb; e += nth) {
            const uint32_t r = e >> 5, i = e & 31;
            const uint32_t s = dt_row_shift(p, T, r);
            const uint32_t head = (16 - s) & 15u;
            const uint32_t h1 = head < len ? head : len;
            const uint32_t whole = len > head ? (len - head) >> 4 : 0u;
            const uint32_t t0 = head + 16 * whole > h1 ? head + 16 * whole : h1;
            const uint32_t x = i < 16 ? i : t0 + (i - 16);      // head byte i, or tail byte i - 16
            if ((i < 16 && x < h1) || (i >= 16 && x < len))
                D[T.gs0 + (int64_t)r * p.S_row + x] = sm[r * (uint32_t)p.PA + x];
        }
    }
}

// histogram of the last pass: per thread, c[r] = #{values <= r} for r < 8
// and the number of values; a byte x <= 127 is <= r exactly when the top bit
// of (0x80 | r) - x is set, four bytes at a time.  Values of 8 and more
// (only far words of long codes, and "unreached") are rare and counted one
// by one in the workgroup's bins.
constexpr int DT_HL = 8;
struct DtHist {
    uint32_t le[DT_HL];
    uint32_t all;
};
__host__ __device__ inline void dt_hist_add4(DtHist& h, uint32_t v, uint32_t nvalid, int n, unsigned int* lh)
{
    // bytes [nvalid, 4) of v are not counted: make them 0 and discount them
    const uint32_t keep = nvalid >= 4 ? 0xFFFFFFFFu : ((1u << (8 * nvalid)) - 1u);
    v &= keep;
    h.all += 4;
#pragma unroll
    for (int r = 0; r < DT_HL; r++)
        h.le[r] += (uint32_t)__builtin_popcount(((0x80808080u | ((uint32_t)r * 0x01010101u)) - v) & 0x80808080u);
    if (nvalid < 4) {                       // the zeroed bytes were counted as 0
        h.le[0] -= 4 - nvalid;
#pragma unroll
        for (int r = 1; r < DT_HL; r++) h.le[r] -= 4 - nvalid;
        h.all -= 4 - nvalid;
    }
    const uint32_t big = ((v | 0x80808080u) - 0x08080808u) & 0x80808080u;   // bytes >= 8
    if (big) {
        for (int k = 0; k < 4; k++)
            if ((big >> (8 * k + 7)) & 1u) {
                const uint32_t x = (v >> (8 * k)) & 0xFFu;
                const int bin = (int)x <= n ? (int)x : n + 1;
#ifdef __HIP_DEVICE_COMPILE__
                atomicAdd(&lh[bin], 1u);
#else
                lh[bin] += 1;
#endif
                h.all -= 1;                 // binned here, not from le[]
            }
    }
}

// the last pass without the full histogram: #{values <= rcut} and the
// maximum, per thread.  A byte x <= 127 is <= rcut exactly when the top bit
// of (0x80 | rcut) - x is set (one subtract, mask and bit count for four
// bytes).  The maximum is kept as packed 16-bit maxima: mo over the dwords
// themselves (lane k's high byte is byte 2k + 1) and me over the dwords
// shifted up a byte (byte 2k); dt_stat_max reads the largest high byte.
struct DtStat {
    uint32_t le, mo, me;
};
__host__ __device__ inline uint32_t dt_max16(uint32_t a, uint32_t b)
{
#ifdef __HIP_DEVICE_COMPILE__
    typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
#else
    const uint32_t l = (a & 0xFFFFu) > (b & 0xFFFFu) ? (a & 0xFFFFu) : (b & 0xFFFFu);
    const uint32_t h = (a >> 16) > (b >> 16) ? (a >> 16) : (b >> 16);
    return l | (h << 16);
#endif
}
__host__ __device__ inline void dt_stat_add4(DtStat& st, uint32_t v, uint32_t nvalid, int rcut)
{
    if (nvalid < 4) {                       // bytes [nvalid, 4) count as 0 (<= rcut), discounted below
        v &= (1u << (8 * nvalid)) - 1u;
        st.le -= 4 - nvalid;
    }
    st.le += (uint32_t)__builtin_popcount(((0x80808080u | ((uint32_t)rcut * 0x01010101u)) - v) & 0x80808080u);
    st.mo = dt_max16(st.mo, v);
    st.me = dt_max16(st.me, v << 8);
}
__host__ __device__ inline uint32_t dt_stat_max(const DtStat& st)
{
    const uint32_t a = st.mo >> 24, b = (st.mo >> 8) & 0xFFu, c = st.me >> 24, d = (st.me >> 8) & 0xFFu;
    const uint32_t x = a > b ? a : b, y = c > d ? c : d;
    return x > y ? x : y;
}

// the last pass's values, read back from LDS after the store
__host__ __device__ inline void dt_hist_tile(const DtPass& p, const DtTile& T, int tid, int nth, const uint8_t* sm,
                                             DtHist& h, DtStat& st, unsigned int* lh)
{
    const uint32_t* al = reinterpret_cast<const uint32_t*>(sm);
    const uint32_t nd = p.nda;                                          // dwords per aligned row
    const int32_t valid = p.inner ? p.q : (int32_t)T.len;               // bytes per row
    const uint32_t nslot = (uint32_t)p.nrow * nd;
    for (uint32_t e = tid; e < nslot; e += nth) {
        const uint32_t r = dt_div(e, nd, p.mda), j = e - r * nd;
        const int32_t k = valid - 4 * (int32_t)j;
        if (k <= 0) continue;
        if (p.full) dt_hist_add4(h, al[e], (uint32_t)(k < 4 ? k : 4), p.hn, lh);
        else dt_stat_add4(st, al[e], (uint32_t)(k < 4 ? k : 4), p.rcut);
    }
}

// the last pass's statistics of a workgroup into its copy of the global bins.
// Light: [0] #{values <= rcut}, [1] the maximum; full: bin d = #{values == d},
// d <= n, and bin n + 1 the unreached.  Wave sums first, then one LDS atomic
// per wave and bin (dt_stats_lds, which may run several times), then one
// global atomic per bin (dt_stats_global).  All threads.
__device__ inline void dt_stats_lds(const DtPass& p, const DtHist& h, const DtStat& st, unsigned int* lh)
{
    const bool lead = (threadIdx.x & 63) == 0;
    if (!p.full) {
        const uint32_t mb = dt_stat_max(st);
        const int64_t le = wave_sum64((int64_t)st.le);
        const int64_t mx = wave_max64((int64_t)mb);
        if (lead && le) atomicAdd(&lh[0], (unsigned int)le);
        if (lead && mx) atomicMax(&lh[1], (unsigned int)mx);
    } else {
        // values < 8 (all of them below bin n + 1 unless n < 7, where no
        // distance exceeds n): bin r = le[r] - le[r-1]
        uint32_t prev = 0;
#pragma unroll
        for (int r = 0; r < DT_HL; r++) {
            const int64_t c = wave_sum64((int64_t)(h.le[r] - prev));
            prev = h.le[r];
            if (lead && c) atomicAdd(&lh[r <= p.hn ? r : p.hn + 1], (unsigned int)c);
        }
    }
}
__device__ inline void dt_stats_global(const DtPass& p, unsigned int* lh, unsigned long long* hist)
{
    __syncthreads();
    const int nbin = p.full ? p.hn + 2 : 2;
    unsigned long long* hc = hist + (size_t)(blockIdx.x % DT_COPIES) * DT_CSTRIDE;
    if ((int)threadIdx.x < nbin && lh[threadIdx.x]) {
        if (p.full || threadIdx.x == 0) atomicAdd(&hc[threadIdx.x], (unsigned long long)lh[threadIdx.x]);
        else atomicMax(&hc[1], (unsigned long long)lh[1]);
    }
}
__device__ inline void dt_stats_flush(const DtPass& p, const DtHist& h, const DtStat& st, unsigned int* lh,
                                      unsigned long long* hist)
{
    __syncthreads();                            // the rare-value bins are complete
    dt_stats_lds(p, h, st, lh);
    dt_stats_global(p, lh, hist);
}

// the copies folded into hh[DT_HB] (host)
inline void dt_fold_copies(const int64_t* copies, bool full, int64_t* hh)
{
    for (int b = 0; b < DT_HB; b++) hh[b] = 0;
    for (int c = 0; c < DT_COPIES; c++) {
        const int64_t* x = copies + (size_t)c * DT_CSTRIDE;
        if (full) {
            for (int b = 0; b < DT_HB; b++) hh[b] += x[b];
        } else {
            hh[0] += x[0];
            if (x[1] > hh[1]) hh[1] = x[1];
        }
    }
}

// ---------------------------------------------------------------- kernels
template <int NT>
__global__ void __launch_bounds__(NT) k_dt_pass(const DtPass p, uint8_t* D, const uint32_t* cwb,
                                                unsigned long long* hist)
{
    extern __shared__ __align__(16) uint8_t sm[];
    __shared__ unsigned int lh[DT_HB];
    if (threadIdx.x < DT_HB) lh[threadIdx.x] = 0;
    const uint4* D4 = reinterpret_cast<const uint4*>(D);
    DtHist h{};
    DtStat st{};
    for (int64_t t = blockIdx.x; t < p.ntiles; t += gridDim.x) {
        const DtTile T = dt_tile(p, t);
        __syncthreads();
        dt_load(p, T, threadIdx.x, NT, sm, D4, cwb);
        __syncthreads();
        dt_realign(p, T, threadIdx.x, NT, sm);
        __syncthreads();
        if (p.inner) {
            dt_inrow(p, threadIdx.x, NT, sm);
            __syncthreads();
        }
        for (int ax = 0; ax < p.nax; ax++) {
            dt_axis(p, T, ax, threadIdx.x, NT, sm);
            __syncthreads();
        }
        dt_store(p, T, threadIdx.x, NT, sm, D);
        if (p.last) dt_hist_tile(p, T, threadIdx.x, NT, sm, h, st, lh);
    }
    if (p.last) dt_stats_flush(p, h, st, lh, hist);
}

// ------------------------------------------------------------ column pass
// An outer group of b axes whose q^b rows fit in a thread's registers
// (q^b <= DT_COL_MAX): thread t owns dword column c of row block a, the 4
// words [c4, c4 + 4) of each of the q^b rows (c4 = 4 c).  It loads them,
// applies the b axes in registers and stores them back: no LDS and no
// barrier, and the 64 lanes of a wave read 256 consecutive bytes of each row.
// Rows start at any byte, so the dwords are unaligned (the buffer loads and
// stores take byte offsets).  A row's last column can hold fewer than 4 of
// its words: its loads read on into the next row (or the allocation's 64
// bytes of padding) and its stores write only the row's own bytes.  Offsets
// are 32-bit: the pass needs N + 64 < 2^32.
constexpr int DT_COL_MAX = 64;
constexpr int DT_COL_NT = 256;

constexpr int dt_cpow(int q, int b) { return b == 0 ? 1 : q * dt_cpow(q, b - 1); }

// axes per column pass: as many as keep a thread under ~100 VGPRs
// (2^4, 3^3 and 8^2 rows do; 2^5 and 4^3 rows do not)
inline int dt_col_bmax(int q)
{
    return q == 2 ? 4 : (q == 3 ? 3 : (q <= 8 ? 2 : 1));
}

// the B axes of Q^B rows held by one thread, four words per row: axis k has
// fibers of Q rows u = Q^k apart
template <int Q, int B>
__host__ __device__ __forceinline__ void dt_col_axes(uint32_t* v)
{
    constexpr int NR = dt_cpow(Q, B);
    int u = 1;
#pragma unroll
    for (int k = 0; k < B; k++) {
#pragma unroll
        for (int f = 0; f < NR / Q; f++) {
            const int r0 = (f / u) * u * Q + f % u;
            uint32_t mn = v[r0];
#pragma unroll
            for (int j = 1; j < Q; j++) mn = dt_pmin(mn, v[r0 + j * u]);
            mn = dt_pinc(mn);
#pragma unroll
            for (int j = 0; j < Q; j++) v[r0 + j * u] = dt_pmin(v[r0 + j * u], mn);
        }
        u *= Q;
    }
}

template <int Q, int B, class Mem>
__host__ __device__ __forceinline__ void dt_col_thread(const DtPass& p, uint32_t t, const Mem& m, DtHist& h,
                                                       DtStat& st, unsigned int* lh)
{
    constexpr int NR = dt_cpow(Q, B);
    static_assert(NR <= DT_COL_MAX, "a column must fit in registers");
    const uint32_t S = (uint32_t)p.S_row;
    const uint32_t a = t / p.ncol;
    const uint32_t c4 = 4 * (t - a * p.ncol);
    const uint32_t off = a * (uint32_t)p.S_tile + c4;
    uint32_t v[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) v[r] = m.ld(off, (uint32_t)r * S);
    dt_col_axes<Q, B>(v);
    const uint32_t nb = S - c4;                 // the row's words in this column (>= 1)
    if (nb >= 4) {
#pragma unroll
        for (int r = 0; r < NR; r++) m.st(off, (uint32_t)r * S, v[r]);
    } else {
#pragma unroll
        for (int r = 0; r < NR; r++)
            for (uint32_t i = 0; i < nb; i++) m.st8(off + i, (uint32_t)r * S, v[r] >> (8 * i));
    }
    if (p.last) {
        const uint32_t nv = nb < 4 ? nb : 4u;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (p.full) dt_hist_add4(h, v[r], nv, p.hn, lh);
            else dt_stat_add4(st, v[r], nv, p.rcut);
        }
    }
}

#ifdef __HIP_DEVICE_COMPILE__
// the distance array through a buffer descriptor: byte offsets, range-checked
// against nbytes (a load past it reads 0, a store past it is dropped)
struct DtBufMem {
    __amdgpu_buffer_rsrc_t rs;
    __device__ uint32_t ld(uint32_t off, uint32_t soff) const
    {
        return __builtin_amdgcn_raw_buffer_load_b32(rs, off, soff, 0);
    }
    __device__ void st(uint32_t off, uint32_t soff, uint32_t v) const
    {
        __builtin_amdgcn_raw_buffer_store_b32(v, rs, off, soff, 0);
    }
    __device__ void st8(uint32_t off, uint32_t soff, uint32_t v) const
    {
        __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v, rs, off, soff, 0);
    }
};
#endif

template <int Q, int B>
__global__ void __launch_bounds__(DT_COL_NT) k_dt_col(const DtPass p, uint8_t* D, unsigned long long* hist)
{
    __shared__ unsigned int lh[DT_HB];
    if (p.last) {
        if (threadIdx.x < DT_HB) lh[threadIdx.x] = 0;
        __syncthreads();
    }
    DtHist h{};
    DtStat st{};
    const uint32_t t = blockIdx.x * DT_COL_NT + threadIdx.x;
#ifdef __HIP_DEVICE_COMPILE__
    if (t < p.nthr) {
        DtBufMem m;
        m.rs = __builtin_amdgcn_make_buffer_rsrc(D, (short)0, (int)p.nbytes, 0x00020000);
        dt_col_thread<Q, B>(p, t, m, h, st, lh);
    }
#endif
    if (p.last) dt_stats_flush(p, h, st, lh, hist);
}

// the (q, b) pairs of column passes, each its own kernel
template <int Q, int B>
struct DtQB {
    static constexpr int q = Q, b = B;
};
template <class F>
inline bool dt_col_dispatch(int q, int b, F&& f)
{
#define DT_QB(Q, B) \
    case Q * 8 + B: f(DtQB<Q, B>{}); return true;
    switch (q * 8 + b) {
        DT_QB(2, 1) DT_QB(2, 2) DT_QB(2, 3) DT_QB(2, 4)
        DT_QB(3, 1) DT_QB(3, 2) DT_QB(3, 3)
        DT_QB(4, 1) DT_QB(4, 2)
        DT_QB(5, 1) DT_QB(5, 2) DT_QB(6, 1) DT_QB(6, 2) DT_QB(7, 1) DT_QB(7, 2) DT_QB(8, 1) DT_QB(8, 2)
        DT_QB(9, 1) DT_QB(10, 1) DT_QB(11, 1) DT_QB(12, 1) DT_QB(13, 1) DT_QB(14, 1) DT_QB(15, 1) DT_QB(16, 1)
    default: break;
    }
#undef DT_QB
    return false;
}

// --------------------------------------------------- two-stage column pass
// A group of b = b1 + b2 axes with q^b1, q^b2 <= 64: a tile is the R = q^b
// rows of a row block, Wd dword columns (4 Wd words) of each, in LDS column
// by column (dword r of column c at c R + r, so that both stages address
// their rows at compile-time offsets).
// Stage A: thread (hh, c), hh < R2 = q^b2, c < Wd, loads column c of the R1
// = q^b1 rows hh R1 + i, applies the b1 axes of the row index's low digits in
// registers and writes the rows to LDS.  Stage B: thread (l, c), l < R1, reads
// column c of rows i R1 + l, i < R2, applies the b2 high-digit axes in
// registers and stores them.  One read and one write of the words for b
// axes, where register column passes take ceil(b / b1) of each.  Columns at
// or past the row's end are skipped (a fiber never mixes columns); a row's
// last column with 1-3 of its words is stored byte by byte.  Offsets are
// 32-bit: N + 64 < 2^32.
constexpr int DT_TC_LDS = 163840 - 512;      // LDS of a tile: 160 KiB less the static bins

// x, opaque to the optimiser: the per-row offsets i x of a stage are then
// computed where they are used, not hoisted out of the tile loop (the two
// stages' offsets together would hold 2 R1 scalar registers for the whole
// kernel and spill)
__host__ __device__ inline uint32_t dt_opaque(uint32_t x)
{
#ifdef __HIP_DEVICE_COMPILE__
    asm volatile("" : "+s"(x));
#endif
    return x;
}

template <int Q, int B1, int B2, class Mem>
__host__ __device__ __forceinline__ void dt_tc_stage_a(const DtPass& p, uint32_t base, uint32_t x0, uint32_t tid,
                                                       const Mem& m, uint32_t* lds)
{
    constexpr int R1 = dt_cpow(Q, B1), R2 = dt_cpow(Q, B2);
    const uint32_t Wd = (uint32_t)p.Wd, S = (uint32_t)p.S_row;
    const uint32_t hh = tid / Wd, c = tid - hh * Wd;
    if (hh >= (uint32_t)R2) return;
    const uint32_t col = x0 + 4 * c;
    if (col >= S) return;
    const uint32_t off = base + col + hh * (uint32_t)R1 * S;
    const uint32_t Su = dt_opaque(S);
    uint32_t v[R1];
#pragma unroll
    for (int i = 0; i < R1; i++) v[i] = m.ld(off, (uint32_t)i * Su);
    dt_col_axes<Q, B1>(v);
    uint32_t* col_l = lds + c * (uint32_t)(R1 * R2) + hh * (uint32_t)R1;
#pragma unroll
    for (int i = 0; i < R1; i++) col_l[i] = v[i];
}

// stage B; on the last pass the values also go back to LDS for dt_tc_stats
template <int Q, int B1, int B2, class Mem>
__host__ __device__ __forceinline__ void dt_tc_stage_b(const DtPass& p, uint32_t base, uint32_t x0, uint32_t tid,
                                                       const Mem& m, uint32_t* lds)
{
    constexpr int R1 = dt_cpow(Q, B1), R2 = dt_cpow(Q, B2);
    const uint32_t Wd = (uint32_t)p.Wd, S = (uint32_t)p.S_row;
    const uint32_t l = tid / Wd, c = tid - l * Wd;
    if (l >= (uint32_t)R1) return;
    const uint32_t col = x0 + 4 * c;
    if (col >= S) return;
    uint32_t u[R2];
    uint32_t* col_l = lds + c * (uint32_t)(R1 * R2) + l;
#pragma unroll
    for (int i = 0; i < R2; i++) u[i] = col_l[i * R1];
    dt_col_axes<Q, B2>(u);
    const uint32_t off = base + col + l * S;
    const uint32_t nb = S - col;                   // the row's words in this column (>= 1)
    const uint32_t step = dt_opaque((uint32_t)R1 * S);
    if (nb >= 4) {
#pragma unroll
        for (int i = 0; i < R2; i++) m.st(off, (uint32_t)i * step, u[i]);
    } else {
#pragma unroll
        for (int i = 0; i < R2; i++)
            for (uint32_t k = 0; k < nb; k++) m.st8(off + k, (uint32_t)i * step, u[i] >> (8 * k));
    }
    if (p.last) {
#pragma unroll
        for (int i = 0; i < R2; i++) col_l[i * R1] = u[i];
    }
}

// the last pass's statistics of a tile from LDS (after a barrier), one
// dword at a time: dword e = column e / R, row e % R
__host__ __device__ inline void dt_tc_stats(const DtPass& p, uint32_t x0, int tid, int nth, const uint32_t* lds,
                                            DtHist& h, DtStat& st, unsigned int* lh)
{
    const uint32_t R = (uint32_t)p.nrow, S = (uint32_t)p.S_row, ne = R * (uint32_t)p.Wd;
    for (uint32_t e = tid; e < ne; e += nth) {
        const uint32_t c = dt_div(e, R, p.mR);
        const uint32_t col = x0 + 4 * c;
        if (col >= S) break;                       // e only grows: every later column is past the end too
        const uint32_t nb = S - col < 4 ? S - col : 4u;
        if (p.full) dt_hist_add4(h, lds[e], nb, p.hn, lh);
        else dt_stat_add4(st, lds[e], nb, p.rcut);
    }
}

// the tile of workgroup slot k (k = blockIdx + i gridDim): workgroups are
// dealt to the 8 XCDs in turn (slot k on XCD k % 8), and each XCD takes a
// contiguous run of tiles, so that neighbouring tiles, which share the cache
// lines at their edges, meet in one L2
__host__ __device__ inline uint32_t dt_tc_slot_tile(uint32_t k, uint32_t ntiles)
{
    const uint32_t per = (ntiles + 7) / 8;
    const uint32_t x = k & 7u, i = k >> 3;
    const uint32_t full = ntiles - 8 * (per - 1);       // XCDs 0 .. full-1 take per tiles, the rest per - 1
    return x < full ? x * per + i : full * per + (x - full) * (per - 1) + i;
}

// tile t of a two-stage pass: its row block's first word and its first column
__host__ __device__ inline void dt_tc_tile(const DtPass& p, uint32_t t, uint32_t& base, uint32_t& x0)
{
    const uint32_t blk = t / (uint32_t)p.nchunk, ch = t - blk * (uint32_t)p.nchunk;
    base = blk * (uint32_t)p.S_tile;
    x0 = ch * 4u * (uint32_t)p.Wd;
}

// workgroup threads at most: 1024 (four waves per SIMD, 128 registers each)
// unless a thread holds 64 rows, which needs more registers than that
constexpr int dt_tc_ntmax(int q, int b1, int b2)
{
    return (dt_cpow(q, b1) >= 64 || dt_cpow(q, b2) >= 64) ? 512 : 1024;
}

template <int Q, int B1, int B2>
__global__ void __launch_bounds__(dt_tc_ntmax(Q, B1, B2)) k_dt_tcol(const DtPass p, uint8_t* D,
                                                                    unsigned long long* hist)
{
    extern __shared__ __align__(16) uint8_t sm[];
    __shared__ unsigned int lh[DT_HB];
    if (threadIdx.x < DT_HB) lh[threadIdx.x] = 0;
#ifdef __HIP_DEVICE_COMPILE__
    DtBufMem m;
    m.rs = __builtin_amdgcn_make_buffer_rsrc(D, (short)0, (int)p.nbytes, 0x00020000);
    uint32_t* lds = reinterpret_cast<uint32_t*>(sm);
    for (uint32_t k = blockIdx.x; k < (uint32_t)p.ntiles; k += gridDim.x) {
        uint32_t base, x0;
        dt_tc_tile(p, dt_tc_slot_tile(k, (uint32_t)p.ntiles), base, x0);
        __syncthreads();                            // the previous tile's stage B is done with LDS
        dt_tc_stage_a<Q, B1, B2>(p, base, x0, threadIdx.x, m, lds);
        __syncthreads();
        dt_tc_stage_b<Q, B1, B2>(p, base, x0, threadIdx.x, m, lds);
        if (p.last) {
            // the tile's statistics into the LDS bins (the per-thread counts
            // live only here, not across the loads of the next tile)
            __syncthreads();
            DtHist h{};
            DtStat st{};
            dt_tc_stats(p, x0, threadIdx.x, blockDim.x, lds, h, st, lh);
            dt_stats_lds(p, h, st, lh);
        }
    }
#endif
    if (p.last) dt_stats_global(p, lh, hist);
}

// the (q, b1, b2) of two-stage passes, each its own kernel: b1 = ceil(b / 2)
// of b in (bmax, 2 bmax] axes, bmax = dt_col_bmax(q)
template <int Q, int B1, int B2>
struct DtQBB {
    static constexpr int q = Q, b1 = B1, b2 = B2;
};
template <class F>
inline bool dt_tcol_dispatch(int q, int b1, int b2, F&& f)
{
#define DT_QBB(Q, B1, B2) \
    case (Q * 8 + B1) * 8 + B2: f(DtQBB<Q, B1, B2>{}); return true;
    switch ((q * 8 + b1) * 8 + b2) {
        DT_QBB(2, 3, 2) DT_QBB(2, 3, 3) DT_QBB(2, 4, 3) DT_QBB(2, 4, 4)
        DT_QBB(3, 2, 2) DT_QBB(3, 3, 2) DT_QBB(3, 3, 3)
        DT_QBB(4, 2, 1) DT_QBB(4, 2, 2) DT_QBB(5, 2, 1) DT_QBB(5, 2, 2) DT_QBB(6, 2, 1) DT_QBB(6, 2, 2)
        DT_QBB(7, 2, 1) DT_QBB(7, 2, 2) DT_QBB(8, 2, 1) DT_QBB(8, 2, 2)
        DT_QBB(9, 1, 1) DT_QBB(10, 1, 1) DT_QBB(11, 1, 1) DT_QBB(12, 1, 1) DT_QBB(13, 1, 1) DT_QBB(14, 1, 1)
        DT_QBB(15, 1, 1) DT_QBB(16, 1, 1)
    default: break;
    }
#undef DT_QBB
    return false;
}

// -------------------------------------------------------------- seed pass
// The first pass, built from the codeword bitmap (one bit per word) instead
// of a distance array.  Its a innermost axes are split in two:
//   the ns innermost axes inside a group of G = q^ns <= 64 consecutive words,
//   decided from the group's G bits alone: with L_0 the codewords and
//   L_{k+1} = L_k | spread_0(L_k) | ... | spread_{ns-1}(L_k) (spread_j sets a
//   whole fiber of digit j where any of its words is set), a word's distance
//   inside the group is the first k with the word in L_k, or "unreached" for a
//   group without codewords;
//   the other a - ns axes across the rows of the LDS tile, one row per group
//   (G bytes padded to P), each axis one phase over q rows u = q^k apart.
// A tile without any codeword (common on sparse codes) skips the row axes and
// stores "unreached" everywhere.  The store reads the tile's bytes back in
// order and writes them with unaligned dword stores; the last pass's
// statistics are taken from the stored values.

// bytes 0x00 / 0xFF from the low 4 bits of b
__host__ __device__ inline uint32_t dt_e4(uint32_t b)
{
    return ((b * 0x00204081u) & 0x01010101u) * 0xFFu;
}

// spread along digit j (fibers of Q words u apart) of a group's bit mask
template <int Q>
__host__ __device__ inline uint64_t dt_spread(uint64_t x, uint32_t u, uint64_t base)
{
    uint64_t y = x;
#pragma unroll
    for (int v = 1; v < Q; v++) y |= x >> (v * u);
    y &= base;
    uint64_t z = y;
#pragma unroll
    for (int v = 1; v < Q; v++) z |= y << (v * u);
    return z;
}

// group gi of tile T into its LDS row; returns whether it holds a codeword
template <int Q>
__host__ __device__ inline bool dt_seed_group(const DtPass& p, const DtTile& T, uint32_t gi, uint8_t* sm,
                                              const uint32_t* cwb)
{
    const uint32_t G = (uint32_t)p.G;
    const int64_t W0 = T.gs0 + (int64_t)gi * G;           // first word of the group
    const int64_t wi = W0 >> 5;
    const uint32_t sh = (uint32_t)W0 & 31u;
    const uint64_t lo = (uint64_t)cwb[wi] | ((uint64_t)cwb[wi + 1] << 32);   // wi + 1 <= (N - 1) / 32 + 1
    uint64_t m = lo >> sh;
    if (sh + G > 64) m |= (uint64_t)cwb[wi + 2] << (64 - sh);
    const uint64_t full = G == 64 ? ~(uint64_t)0 : (((uint64_t)1 << G) - 1);
    m &= full;
    uint32_t* row = reinterpret_cast<uint32_t*>(sm + (size_t)gi * (uint32_t)p.P);
    const uint32_t nd = (uint32_t)p.P >> 2;
    if (!m) {
        for (uint32_t c = 0; c < nd; c++) row[c] = DT_INF4;
        return false;
    }
    // levels L_0 .. L_{ns-1}; L_ns is the whole group
    uint64_t L[DT_SMAX];
    L[0] = m;
#pragma unroll
    for (int k = 1; k < DT_SMAX; k++) {
        if (k < p.ns) {
            uint64_t x = L[k - 1];
            uint32_t u = 1;
            for (int j = 0; j < p.ns; j++, u *= Q) x |= dt_spread<Q>(L[k - 1], u, p.sbase[j]);
            L[k] = x;
        }
    }
    const uint32_t ks = (uint32_t)p.ns * 0x01010101u;
    for (uint32_t c = 0; c < nd; c++) {
        const uint32_t ef = dt_e4((uint32_t)(full >> (4 * c)) & 15u);   // words of the group (not padding)
        uint32_t v = (ef & ks) | (~ef & DT_INF4);
#pragma unroll
        for (int k = DT_SMAX - 1; k >= 0; k--) {
            if (k < p.ns) {
                const uint32_t e = dt_e4((uint32_t)(L[k] >> (4 * c)) & 15u);
                v = (e & ((uint32_t)k * 0x01010101u)) | (~e & v);
            }
        }
        row[c] = v;
    }
    return true;
}

// per-u16 min and sum of two dwords each holding two u16 lanes
__host__ __device__ inline uint32_t dt_min16(uint32_t a, uint32_t b)
{
#ifdef __HIP_DEVICE_COMPILE__
    typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
#else
    const uint32_t l = (a & 0xFFFFu) < (b & 0xFFFFu) ? (a & 0xFFFFu) : (b & 0xFFFFu);
    const uint32_t h = (a >> 16) < (b >> 16) ? (a >> 16) : (b >> 16);
    return l | (h << 16);
#endif
}

// row axis k of the tile (fibers of Q rows u = q^k apart): each dword is split
// into its even and odd bytes, two u16 lanes each, so that one packed
// 16-bit min handles two values; 1 + the fiber minimum is clamped to
// "unreached", so every value stays a byte <= 127
template <int Q>
__host__ __device__ inline void dt_seed_axis(const DtPass& p, int ax, int tid, int nth, uint8_t* sm)
{
    const uint32_t u = p.u[ax], mg = p.mg[ax];
    const uint32_t nd = (uint32_t)p.P >> 2;
    const uint32_t nf = (uint32_t)p.nrow / Q;
    const uint32_t nslot = nf * nd, step = u * nd;
    uint32_t* al = reinterpret_cast<uint32_t*>(sm);
    for (uint32_t e = tid; e < nslot; e += nth) {
        const uint32_t f = dt_div(e, nd, p.mda);
        const uint32_t c = e - f * nd;
        const uint32_t hi = dt_div(f, u, mg), lo = f - hi * u;
        uint32_t* x = al + (hi * u * Q + lo) * nd + c;
        uint32_t ev[Q], od[Q];
#pragma unroll
        for (int v = 0; v < Q; v++) {
            const uint32_t w = x[v * step];
            ev[v] = w & 0x00FF00FFu;
            od[v] = (w >> 8) & 0x00FF00FFu;
        }
        uint32_t me = ev[0], mo = od[0];
#pragma unroll
        for (int v = 1; v < Q; v++) {
            me = dt_min16(me, ev[v]);
            mo = dt_min16(mo, od[v]);
        }
        me = dt_min16(me + 0x00010001u, 0x007F007Fu);
        mo = dt_min16(mo + 0x00010001u, 0x007F007Fu);
#pragma unroll
        for (int v = 0; v < Q; v++) x[v * step] = dt_min16(ev[v], me) | (dt_min16(od[v], mo) << 8);
    }
}

// the tile's bytes in order: byte b sits in group b / G, column b % G
__host__ __device__ inline uint32_t dt_seed_dword(const DtPass& p, const uint8_t* sm, uint32_t b)
{
    const uint32_t G = (uint32_t)p.G, P = (uint32_t)p.P;
    const uint32_t gi = dt_div(b, G, p.mG), col = b - gi * G;
    const uint32_t at = gi * P + col;
    const uint32_t* al = reinterpret_cast<const uint32_t*>(sm);
    if (col + 4 <= G) {
        if ((at & 3) == 0) return al[at >> 2];
        return dt_align(al[(at >> 2) + 1], al[at >> 2], at & 3);
    }
    uint32_t v = 0;
    for (uint32_t i = 0; i < 4; i++) {
        const uint32_t c = col + i;
        const uint32_t a2 = c < G ? at + i : (gi + 1) * P + (c - G);
        v |= (uint32_t)sm[a2] << (8 * i);
    }
    return v;
}

// the store of a seed tile (all of it "unreached" when it has no codeword),
// and the last pass's statistics
template <class Mem>
__host__ __device__ inline void dt_seed_store(const DtPass& p, const DtTile& T, int tid, int nth, const uint8_t* sm,
                                              bool any, const Mem& m, DtHist& h, DtStat& st, unsigned int* lh)
{
    const uint32_t tot = (uint32_t)p.S_tile, nk = tot >> 2;
    const uint32_t g0 = (uint32_t)T.gs0;
    for (uint32_t k = tid; k < nk; k += nth) {
        const uint32_t v = any ? dt_seed_dword(p, sm, 4 * k) : DT_INF4;
        m.st(4 * k, g0, v);
        if (p.last) {
            if (p.full) dt_hist_add4(h, v, 4, p.hn, lh);
            else dt_stat_add4(st, v, 4, p.rcut);
        }
    }
    const uint32_t tail = tot & 3u;
    if (tail && tid == nth - 1) {
        uint32_t v = 0;
        for (uint32_t i = 0; i < tail; i++) {
            const uint32_t b = 4 * nk + i;
            const uint32_t gi = dt_div(b, (uint32_t)p.G, p.mG);
            const uint32_t x = any ? sm[gi * (uint32_t)p.P + (b - gi * (uint32_t)p.G)] : DT_INF;
            m.st8(g0 + b, 0, x);
            v |= x << (8 * i);
        }
        if (p.last) {
            if (p.full) dt_hist_add4(h, v, tail, p.hn, lh);
            else dt_stat_add4(st, v, tail, p.rcut);
        }
    }
}

// ------------------------------------------------------- sparse seed tiles
// After the a inner axes a word x holds the minimum, over the codewords c of
// its block (the q^a words that share x's n - a outer digits), of the Hamming
// distance on the a inner digits.  A tile with at most sk codewords takes
// that closed form instead of the level sets and row axes.  With x = hi G +
// lo (group row hi, column lo), d(x, c) = H_c(hi) + L_c(lo): the distance on
// the nax row digits plus the one on the ns group digits.  LDS holds, per
// codeword, L_c over Ls >= G + 8 columns (wrapping into the next row) and
// H_c over the rows (127 for a row of another block, and for row nrow).
//
// The store writes aligned dwords.  Tile byte b = head + 4 j of dword j; its
// four bytes lie in at most two rows (G >= 4).  The dwords repeat their
// column offsets with period Pd = G / gcd(G, 4), which spans Rp = 4 / gcd(G,
// 4) rows: dword j + Pd is dword j moved Rp rows down.  So thread (a, p), p <
// Pd, keeps the column part of its dwords p + Pd Q (Q = a, a + A, ...) per
// codeword in registers and per dword only reads H_c of two rows: a
// multiply, a bit-field insert, an add and a per-byte minimum per codeword.
// The tile's 0-3 head and tail bytes are byte stores.  LDS: positions (DT_SK
// int32) at 0, column tables of Ls bytes at DT_SPL, row tables of Hs bytes at
// hoff.
constexpr int DT_SK = 4;
constexpr int DT_SPL = 4 * DT_SK;    // LDS offset of the column tables

// bitmap dword d restricted to the words [lo, hi)
__host__ __device__ inline uint32_t dt_sp_bits(const uint32_t* cwb, uint32_t d, uint32_t lo, uint32_t hi)
{
    uint32_t v = cwb[d];
    const uint32_t w0 = d * 32;
    if (w0 < lo) v &= ~0u << (lo - w0);
    if (w0 + 32 > hi) v &= (hi - w0 >= 32) ? ~0u : ((1u << (hi - w0)) - 1u);
    return v;
}

// the tile's codewords: each thread counts the bits of its bitmap dwords,
// reserves that many slots with one add to *cnt, and writes the tile offsets
// of those slots below sk.  After a barrier *cnt is the tile's count.
__host__ __device__ inline void dt_sp_collect(const DtPass& p, const DtTile& T, int tid, int nth, const uint32_t* cwb,
                                              int32_t* pos, int* cnt)
{
    const uint32_t g0 = (uint32_t)T.gs0, g1 = g0 + (uint32_t)p.S_tile;
    const uint32_t d0 = g0 >> 5, nd = ((g1 - 1) >> 5) - d0 + 1;
    uint32_t c = 0;
    for (uint32_t j = tid; j < nd; j += nth) c += (uint32_t)__builtin_popcount(dt_sp_bits(cwb, d0 + j, g0, g1));
    if (!c) return;
#ifdef __HIP_DEVICE_COMPILE__
    int slot = atomicAdd(cnt, (int)c);
#else
    int slot = *cnt;
    *cnt += (int)c;
#endif
    for (uint32_t j = tid; j < nd && slot < p.sk; j += nth) {
        uint32_t b = dt_sp_bits(cwb, d0 + j, g0, g1);
        while (b && slot < p.sk) {
            pos[slot++] = (int32_t)((d0 + j) * 32 + (uint32_t)__builtin_ctz(b) - g0);
            b &= b - 1;
        }
    }
}

// Hamming distance of the nd low base-q digits of a and c
__host__ __device__ inline uint32_t dt_sp_ham(const DtPass& p, uint32_t a, uint32_t c, int nd)
{
    uint32_t d = 0;
    for (int j = 0; j < nd; j++) {
        const uint32_t a1 = dt_div(a, (uint32_t)p.q, p.mq), c1 = dt_div(c, (uint32_t)p.q, p.mq);
        d += (a - a1 * (uint32_t)p.q) != (c - c1 * (uint32_t)p.q);
        a = a1;
        c = c1;
    }
    return d;
}

// the column and row tables of the k <= sk codewords at pos[]
__host__ __device__ inline void dt_sp_tables(const DtPass& p, int k, const int32_t* pos, int tid, int nth, uint8_t* sm)
{
    const uint32_t G = (uint32_t)p.G, Ls = (uint32_t)p.Ls, Hs = (uint32_t)p.Hs, nrow = (uint32_t)p.nrow;
    for (int i = 0; i < k; i++) {
        const uint32_t o = (uint32_t)pos[i];
        const uint32_t cr = dt_div(o, G, p.mG), cl = o - cr * G;
        const uint32_t cb = dt_div(cr, p.rb, p.mrb), crl = cr - cb * p.rb;
        for (uint32_t x = tid; x < Ls; x += nth) {
            const uint32_t r = dt_div(x, G, p.mG);
            sm[DT_SPL + i * Ls + x] = (uint8_t)dt_sp_ham(p, x - r * G, cl, p.ns);
        }
        for (uint32_t hi = tid; hi < Hs; hi += nth) {
            const uint32_t hb = dt_div(hi, p.rb, p.mrb);
            sm[p.hoff + i * Hs + hi] = (uint8_t)(hi >= nrow || hb != cb ? DT_INF
                                                                         : dt_sp_ham(p, hi - hb * p.rb, crl, p.nax));
        }
    }
}

// bytes over 127 (127 + a column distance, rows of another block) to 127
__host__ __device__ inline uint32_t dt_sp_clamp(uint32_t d)
{
    const uint32_t t = d & 0x80808080u;
    return (d | (t - (t >> 7))) & ~t;
}

// bytes of {s0, s1} by selector (v_perm_b32): selector byte i picks byte
// sel_i of s1 (0-3) or byte sel_i - 4 of s0 (4-7)
__host__ __device__ inline uint32_t dt_perm(uint32_t s0, uint32_t s1, uint32_t sel)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_perm(s0, s1, sel);
#else
    const uint64_t x = ((uint64_t)s0 << 32) | s1;
    uint32_t r = 0;
    for (int i = 0; i < 4; i++) r |= (uint32_t)((x >> (8 * ((sel >> (8 * i)) & 7u))) & 0xFFu) << (8 * i);
    return r;
#endif
}

// tile byte b from the tables of its k codewords
__host__ __device__ inline uint32_t dt_sp_byte(const DtPass& p, const uint8_t* sm, int k, uint32_t b)
{
    const uint32_t G = (uint32_t)p.G;
    const uint32_t hi = dt_div(b, G, p.mG), lo = b - hi * G;
    uint32_t v = DT_INF;
    for (int i = 0; i < k; i++) {
        const uint32_t d = sm[DT_SPL + i * (uint32_t)p.Ls + lo] + (uint32_t)sm[p.hoff + i * (uint32_t)p.Hs + hi];
        v = d < v ? d : v;
    }
    return v;
}

// the closed-form dwords of thread (aa, pp): dwords j = Q Pd + pp for Q = aa,
// aa + A, ...; K >= k codewords (the ones past k skipped), CLAMP when rows
// of another block (127 + a column distance) can occur
template <int K, bool CLAMP, class Mem>
__host__ __device__ inline void dt_sp_loop(const DtPass& p, const uint8_t* sm, int k, uint32_t aa, uint32_t pp,
                                           uint32_t A, uint32_t head, uint32_t nk, uint32_t g0, const Mem& m,
                                           DtHist& h, DtStat& st, unsigned int* lh)
{
    const uint32_t G = (uint32_t)p.G, Pd = (uint32_t)p.Pd, Rp = (uint32_t)p.Rp;
    const uint32_t Ls = (uint32_t)p.Ls, Hs = (uint32_t)p.Hs;
    const uint32_t b0 = head + 4 * pp;
    const uint32_t r0 = dt_div(b0, G, p.mG), lo0 = b0 - r0 * G;
    const uint32_t left = G - lo0;                       // bytes of row r0 in the dword
    // byte i from row r (selector 0: byte 0 of s1) or r + 1 (selector 4: byte 0 of s0)
    const uint32_t sel = left >= 4 ? 0u : (0x04040404u & ~((1u << (8 * left)) - 1u));
    const uint32_t* L = reinterpret_cast<const uint32_t*>(sm);
    const uint8_t* H = sm + p.hoff;
    uint32_t lq[K];
#pragma unroll
    for (int i = 0; i < K; i++) {
        const uint32_t at = DT_SPL + (uint32_t)i * Ls + lo0;
        lq[i] = (K == 1 || i < k) ? dt_align(L[(at >> 2) + 1], L[at >> 2], at & 3u) : 0u;
    }
    for (uint32_t Q = aa;; Q += A) {
        const uint32_t j = Q * Pd + pp;
        if (j >= nk) break;
        const uint32_t r = r0 + Rp * Q;
        uint32_t h0[K], h1[K];
#pragma unroll
        for (int i = 0; i < K; i++) {
            if (K <= 2 || i < k) {
                h0[i] = H[(uint32_t)i * Hs + r];
                h1[i] = H[(uint32_t)i * Hs + r + 1];
            }
        }
        uint32_t v = DT_INF4;
#pragma unroll
        for (int i = 0; i < K; i++) {
            if (K <= 2 || i < k) {
                uint32_t d = lq[i] + dt_perm(h1[i], h0[i], sel);
                if (CLAMP) d = dt_sp_clamp(d);
                v = i == 0 ? d : dt_pmin(v, d);
            }
        }
        m.st(head + 4 * j, g0, v);
        if (p.last) {
            if (p.full) dt_hist_add4(h, v, 4, p.hn, lh);
            else dt_stat_add4(st, v, 4, p.rcut);
        }
    }
}

// the store of a tile in the closed form (k = 0: all unreached) and the last
// pass's statistics
template <class Mem>
__host__ __device__ inline void dt_sp_store(const DtPass& p, const DtTile& T, int tid, int nth, const uint8_t* sm,
                                            int k, const Mem& m, DtHist& h, DtStat& st, unsigned int* lh)
{
    const uint32_t tot = (uint32_t)p.S_tile, g0 = (uint32_t)T.gs0, G = (uint32_t)p.G;
    const uint32_t head = (4u - (g0 & 3u)) & 3u;
    const uint32_t nk = (tot - head) >> 2;
    if (k == 0) {
        for (uint32_t j = tid; j < nk; j += nth) {
            m.st(head + 4 * j, g0, DT_INF4);
            if (p.last) {
                if (p.full) dt_hist_add4(h, DT_INF4, 4, p.hn, lh);
                else dt_stat_add4(st, DT_INF4, 4, p.rcut);
            }
        }
    } else {
        const uint32_t Pd = (uint32_t)p.Pd, A = (uint32_t)nth / Pd;
        const uint32_t aa = (uint32_t)tid / Pd, pp = (uint32_t)tid - aa * Pd;
        if (aa < A) {
            const bool blocks = (uint32_t)p.nrow > p.rb;
            if (k == 1) {
                if (blocks) dt_sp_loop<1, true>(p, sm, k, aa, pp, A, head, nk, g0, m, h, st, lh);
                else dt_sp_loop<1, false>(p, sm, k, aa, pp, A, head, nk, g0, m, h, st, lh);
            } else if (k == 2) {
                if (blocks) dt_sp_loop<2, true>(p, sm, k, aa, pp, A, head, nk, g0, m, h, st, lh);
                else dt_sp_loop<2, false>(p, sm, k, aa, pp, A, head, nk, g0, m, h, st, lh);
            } else {
                dt_sp_loop<DT_SK, true>(p, sm, k, aa, pp, A, head, nk, g0, m, h, st, lh);
            }
        }
    }
    if (tid == nth - 1) {
        const uint32_t t0 = head + 4 * nk, ne = head + (tot - t0);
        for (uint32_t e = 0; e < ne; e++) {
            const uint32_t b = e < head ? e : t0 + (e - head);
            const uint32_t x = k ? dt_sp_byte(p, sm, k, b) : DT_INF;
            m.st8(b, g0, x);
            if (p.last) {
                if (p.full) dt_hist_add4(h, x, 1, p.hn, lh);
                else dt_stat_add4(st, x, 1, p.rcut);
            }
        }
    }
}

template <int Q>
__global__ void __launch_bounds__(DT_COL_NT) k_dt_seed(const DtPass p, uint8_t* D, const uint32_t* cwb,
                                                       unsigned long long* hist)
{
    extern __shared__ __align__(16) uint8_t sm[];
    __shared__ unsigned int lh[DT_HB];
    __shared__ int ncw;
    if (threadIdx.x < DT_HB) lh[threadIdx.x] = 0;
    DtHist h{};
    DtStat st{};
#ifdef __HIP_DEVICE_COMPILE__
    DtBufMem m;
    m.rs = __builtin_amdgcn_make_buffer_rsrc(D, (short)0, (int)p.nbytes, 0x00020000);
    for (int64_t t = blockIdx.x; t < p.ntiles; t += gridDim.x) {
        const DtTile T = dt_tile(p, t);
        int k = p.sk + 1;
        if (p.sk) {
            __syncthreads();                        // the previous tile is done with LDS and ncw
            if (threadIdx.x == 0) ncw = 0;
            __syncthreads();
            dt_sp_collect(p, T, threadIdx.x, DT_COL_NT, cwb, reinterpret_cast<int32_t*>(sm), &ncw);
            __syncthreads();
            k = ncw;
        }
        if (k <= p.sk) {
            if (k) {
                dt_sp_tables(p, k, reinterpret_cast<const int32_t*>(sm), threadIdx.x, DT_COL_NT, sm);
                __syncthreads();
            }
            dt_sp_store(p, T, threadIdx.x, DT_COL_NT, sm, k, m, h, st, lh);
            continue;
        }
        __syncthreads();
        int any = 0;
        for (uint32_t gi = threadIdx.x; gi < (uint32_t)p.nrow; gi += DT_COL_NT)
            any |= dt_seed_group<Q>(p, T, gi, sm, cwb) ? 1 : 0;
        any = __syncthreads_or(any);
        if (any) {
            for (int ax = 0; ax < p.nax; ax++) {
                dt_seed_axis<Q>(p, ax, threadIdx.x, DT_COL_NT, sm);
                __syncthreads();
            }
        }
        dt_seed_store(p, T, threadIdx.x, DT_COL_NT, sm, any != 0, m, h, st, lh);
    }
#endif
    if (p.last) dt_stats_flush(p, h, st, lh, hist);
}

template <class F>
inline bool dt_seed_dispatch(int q, F&& f)
{
#define DT_SQ(Q) \
    case Q: f(DtQB<Q, 0>{}); return true;
    switch (q) {
        DT_SQ(2) DT_SQ(3) DT_SQ(4) DT_SQ(5) DT_SQ(6) DT_SQ(7) DT_SQ(8) DT_SQ(9)
        DT_SQ(10) DT_SQ(11) DT_SQ(12) DT_SQ(13) DT_SQ(14) DT_SQ(15) DT_SQ(16)
    default: break;
    }
#undef DT_SQ
    return false;
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
constexpr int DT_LDS_DEFAULT = 49152;
constexpr int DT_THREADS_DEFAULT = 512;
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

// axes a seed group spans: q^ns <= 64 words
inline int dt_seed_smax(int q)
{
    return q == 2 ? 6 : (q <= 4 ? 3 : (q <= 8 ? 2 : 1));
}

// the columns and threads of a two-stage pass (nrow, S_row, S_tile set): the
// tile width that finishes its tiles in the fewest resident rounds of ncu
// CUs, each round costing the bytes a CU holds in it (ties: the wider tile)
inline bool dt_plan_tcol(DtPass& p, int q, int64_t nblk, int ncu)
{
    const int64_t R = p.nrow, R1 = dt_ipow(q, p.b1), R2 = R / R1;
    const int64_t rmax = R1 > R2 ? R1 : R2;
    const int64_t ntmax = dt_tc_ntmax(q, p.b1, p.nax - p.b1);
    int64_t best = -1, wbest = 0;
    for (int64_t Wd = std::min<int64_t>(ntmax / rmax, DT_TC_LDS / (4 * R)); Wd >= 4; Wd--) {
        const int64_t nt = (rmax * Wd + 63) / 64 * 64;
        int64_t per = DT_TC_LDS / (4 * R * Wd);
        if (per > 2048 / nt) per = 2048 / nt;
        if (per < 1) continue;
        const int64_t ntiles = nblk * ((p.S_row + 4 * Wd - 1) / (4 * Wd));
        const int64_t rounds = (ntiles + ncu * per - 1) / (ncu * per);
        const int64_t cost = rounds * per * R * Wd;
        if (best < 0 || cost < best) {
            best = cost;
            wbest = Wd;
        }
    }
    if (best < 0) return false;
    p.Wd = (int32_t)wbest;
    p.nchunk = (int32_t)((p.S_row + 4 * wbest - 1) / (4 * wbest));
    p.ntiles = nblk * p.nchunk;
    p.nt = (int32_t)((rmax * wbest + 63) / 64 * 64);
    p.lds = 4 * R * wbest;
    p.mR = dt_magic((uint64_t)R);
    int64_t per = DT_TC_LDS / p.lds;
    if (per > 2048 / p.nt) per = 2048 / p.nt;
    p.grid = (int32_t)std::min<int64_t>(p.ntiles, per * ncu);
    return true;
}

// column passes for the outer axes [0, m), appended to out[np..]; returns the
// new count.  Groups of at most bmax = dt_col_bmax(q) axes are register
// passes; with tcol, groups of up to 2 bmax axes are two-stage passes.
inline int dt_plan_columns(int q, int n, int64_t N, int m, DtPass* out, int np, bool tcol = true, int ncu = 256)
{
    const int bmax = dt_col_bmax(q);
    const int gmax = tcol ? 2 * bmax : bmax;
    const int ngroups = (m + gmax - 1) / gmax;
    int j1 = m;
    for (int g = 0; g < ngroups; g++) {
        const int left = ngroups - g;
        const int b = (j1 + left - 1) / left;
        const int j0 = j1 - b;
        DtPass& p = out[np++];
        p = DtPass{};
        p.N = N; p.q = q; p.hn = n; p.col = 1; p.nax = b;
        p.nrow = dt_ipow(q, b);
        p.S_row = dt_ipow(q, n - j1);
        p.S_tile = p.nrow * p.S_row;
        p.ncol = (uint32_t)((p.S_row + 3) / 4);
        p.ntiles = dt_ipow(q, j0) * (int64_t)p.ncol;
        p.nthr = (uint32_t)p.ntiles;
        p.nbytes = (uint32_t)(N + 64);
        if (b > bmax) {
            p.col = 0;
            p.tcol = 1;
            p.b1 = (b + 1) / 2;
            if (!dt_plan_tcol(p, q, dt_ipow(q, j0), ncu)) return 0;
        }
        j1 = j0;
    }
    return np;
}

// the seed pass over the a innermost axes (tiles of `per` blocks of q^a
// words, >= 512 tiles on large spaces), then column passes
inline int dt_plan_seed(int q, int n, int64_t N, int budget, DtPass* out, bool tcol, int ncu)
{
    const int smax = dt_seed_smax(q);
    auto lds_of = [&](int a) {
        const int s = a < smax ? a : smax;
        const int64_t G = dt_ipow(q, s);
        return dt_ipow(q, a - s) * ((G + 3) & ~(int64_t)3);
    };
    int a = 1;
    while (a < n && lds_of(a + 1) <= budget && (N / dt_ipow(q, a + 1) >= 512 || N <= 64 * (int64_t)budget)) a++;
    if (lds_of(a) > budget) return 0;
    const int s = a < smax ? a : smax;
    const int64_t blk = dt_ipow(q, a);
    int64_t per = budget / lds_of(a);
    const int64_t per_par = N / (blk * 512);
    if (per > per_par) per = per_par;
    if (per < 1) per = 1;
    while (per > 1 && N % (per * blk)) per--;
    DtPass& p = out[0];
    p = DtPass{};
    p.N = N; p.q = q; p.hn = n; p.first = 1; p.seed = 1;
    p.ns = s;
    p.G = (int32_t)dt_ipow(q, s);
    p.P = (p.G + 3) & ~3;
    p.mG = dt_magic((uint64_t)p.G);
    p.nrow = per * dt_ipow(q, a - s);
    p.S_row = p.G; p.Lc = p.G; p.nch = 1;
    p.S_tile = per * blk;
    p.ntiles = N / p.S_tile;
    p.nda = (uint32_t)(p.P / 4);
    p.mda = dt_magic(p.nda);
    p.lds = p.nrow * p.P;
    p.nbytes = (uint32_t)(N + 64);
    dt_fill_axes(p, q, a - s);
    for (int j = 0; j < s; j++) {
        const int64_t uj = dt_ipow(q, j);
        uint64_t b = 0;
        for (int64_t w = 0; w < p.G; w++)
            if ((w / uj) % q == 0) b |= (uint64_t)1 << w;
        p.sbase[j] = b;
    }
    if (N % p.S_tile || p.lds > budget) return 0;
    // the closed form for tiles with few codewords: as many codewords as the
    // LDS budget holds tables for, up to DT_SK; groups of 4 words or more
    p.mq = dt_magic((uint64_t)q);
    p.rb = (uint32_t)dt_ipow(q, a - s);
    p.mrb = dt_magic(p.rb);
    p.Ls = (p.G + 8 + 3) & ~3;
    p.Hs = (int32_t)((p.nrow + 1 + 3) & ~(int64_t)3);
    {
        const int g4 = (p.G % 4 == 0) ? 4 : (p.G % 2 == 0 ? 2 : 1);
        p.Pd = p.G / g4;
        p.Rp = 4 / g4;
    }
    p.sk = 0;
    if (p.G >= 4) {
        for (int sk = DT_SK; sk >= 1; sk--) {
            const int64_t hoff = DT_SPL + (int64_t)sk * p.Ls;
            const int64_t need = hoff + (int64_t)sk * p.Hs;
            if (need <= budget) {
                p.sk = sk;
                p.hoff = (int32_t)hoff;
                if (need > p.lds) p.lds = need;
                break;
            }
        }
    }
    const int np = dt_plan_columns(q, n, N, n - a, out, 1, tcol, ncu);
    if (np <= 0) return 0;
    out[np - 1].last = 1;
    return np;
}

// the passes for a cell; returns their number (0 on a bad budget): a seed
// pass and column passes, unless `columns` is false or N is too large for
// 32-bit offsets; then an inner pass and outer passes in LDS tiles.
inline int dt_plan(int q, int n, int64_t N, int budget, DtPass* out, bool columns = true, bool tcol = true,
                   int ncu = 256)
{
    if (budget < 4096 || budget > DT_LDS_MAX || q < 2 || q > 16 || n < 1 || n > MAXN || ncu < 1) return 0;
    if (columns && N + 64 <= (int64_t)0xFFFFFFFFu) return dt_plan_seed(q, n, N, budget, out, tcol, ncu);
    int np = 0;
    // inner pass: rows of one innermost fiber, padded to W = 4, 8 or 16 bytes;
    // a axes (the innermost + a-1 across rows) in blocks of q^a words,
    // keeping >= 512 tiles on large spaces.  LDS: rows * W + raw q^a + 48.
    const int W = q <= 4 ? 4 : (q <= 8 ? 8 : 16);
    auto inner_lds = [&](int64_t words) {
        return ((words / q * W + 15) & ~(int64_t)15) + ((words + 47) & ~(int64_t)15);
    };
    int a = 1;
    while (a < n && inner_lds(dt_ipow(q, a + 1)) <= budget &&
           (N / dt_ipow(q, a + 1) >= 512 || N <= 64 * (int64_t)budget))
        a++;
    {
        DtPass& p = out[np++];
        p = DtPass{};
        const int64_t blk = dt_ipow(q, a);          // words per block
        int64_t per = budget / inner_lds(blk);
        const int64_t per_par = N / (blk * 512);
        if (per > per_par) per = per_par;
        if (per < 1) per = 1;
        while (per > 1 && N % (per * blk)) per--;
        p.N = N; p.q = q; p.hn = n; p.inner = 1; p.first = 1;
        p.nrow = per * blk / q;
        p.S_row = q; p.Lc = q; p.nch = 1;
        p.S_tile = per * blk;
        p.ntiles = N / p.S_tile;
        p.PA = W;
        p.nda = (uint32_t)(W / 4);
        p.mda = dt_magic(p.nda);
        p.raw_off = (int32_t)(((p.nrow * W) + 15) & ~(int64_t)15);
        p.lds = p.raw_off + ((p.S_tile + 47) & ~(int64_t)15);
        p.mq = dt_magic((uint64_t)q);
        dt_fill_axes(p, q, a - 1);
        if (N % p.S_tile || p.lds > budget) return 0;
    }
    const int m = n - a;
    // outer groups in LDS tiles: at most bmax axes each, rows of >= 64 words
    int bmax = 0;
    for (int64_t B = q; B * (2 * 64 + 48) <= budget && bmax < m; B *= q) bmax++;
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
        // Lc: a multiple of 16 as large as fits, no longer than the row needs
        int64_t Lc = ((budget / rows - 48) / 2) & ~(int64_t)15;
        const int64_t need = (S_row + 15) & ~(int64_t)15;
        if (Lc > need) Lc = need;
        if (Lc > 4096) Lc = 4096;
        if (Lc < 64) return 0;
        p.N = N; p.q = q; p.hn = n;
        p.nrow = rows; p.S_row = S_row; p.S_tile = rows * S_row; p.Lc = Lc;
        p.nch = (S_row + Lc - 1) / Lc;
        p.ntiles = dt_ipow(q, j0) * p.nch;
        p.PA = (int32_t)(Lc + 16);
        p.nda = (uint32_t)(p.PA / 4);
        p.mda = dt_magic(p.nda);
        p.PR = (int32_t)(Lc + 32);
        p.raw_off = (int32_t)(rows * p.PA);
        p.lds = p.raw_off + rows * p.PR;
        p.nkr = (int32_t)(Lc / 16 + 2);
        p.mkr = dt_magic((uint64_t)p.nkr);
        p.nvp = (int32_t)(Lc / 16);
        p.mvp = dt_magic((uint64_t)p.nvp);
        if (p.lds > budget) return 0;
        if ((uint64_t)rows * p.nkr * p.nkr >= ((uint64_t)1 << 32)) return 0;   // dt_div exactness
        dt_fill_axes(p, q, b);
        j1 = j0;
    }
    out[np - 1].last = 1;
    return np;
}

// LDS bytes a pass's tile needs
inline int64_t dt_lds_bytes(const DtPass& p)
{
    return p.lds;
}

}  // namespace cov
