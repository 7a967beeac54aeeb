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
// maximum, per thread; per-byte max(a, b) = a + b - min(a, b) (bytes <= 127)
struct DtStat {
    uint32_t le, mx;
};
__host__ __device__ inline void dt_stat_add4(DtStat& st, uint32_t v, uint32_t nvalid, int rcut)
{
    const uint32_t keep = nvalid >= 4 ? 0xFFFFFFFFu : ((1u << (8 * nvalid)) - 1u);
    v &= keep;
    st.le += (uint32_t)__builtin_popcount(((0x80808080u | ((uint32_t)rcut * 0x01010101u)) - v) & 0x80808080u) -
             (4 - (nvalid >= 4 ? 4 : nvalid));
    st.mx = st.mx + v - dt_pmin(st.mx, v);
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
    if (p.last && !p.full) {
        // bins: [0] #{values <= rcut}, [1] the maximum
        __syncthreads();
        uint32_t mb = 0;
        for (int k = 0; k < 4; k++) {
            const uint32_t b = (st.mx >> (8 * k)) & 0xFFu;
            mb = b > mb ? b : mb;
        }
        if (st.le) atomicAdd(&lh[0], st.le);
        if (mb) atomicMax(&lh[1], mb);
        __syncthreads();
        if (threadIdx.x < 2 && lh[threadIdx.x]) {
            if (threadIdx.x == 0) atomicAdd(&hist[0], (unsigned long long)lh[0]);
            else atomicMax(&hist[1], (unsigned long long)lh[1]);
        }
    }
    if (p.last && p.full) {
        __syncthreads();
        // values < 8 (all of them below bin n + 1 unless n < 7, where no
        // distance exceeds n): bin r = le[r] - le[r-1]
        uint32_t prev = 0;
#pragma unroll
        for (int r = 0; r < DT_HL; r++) {
            const uint32_t c = h.le[r] - prev;
            prev = h.le[r];
            if (c) atomicAdd(&lh[r <= p.hn ? r : p.hn + 1], c);
        }
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

// the passes for a cell; returns their number (0 on a bad budget)
inline int dt_plan(int q, int n, int64_t N, int budget, DtPass* out)
{
    if (budget < 4096 || budget > DT_LDS_MAX || q < 2 || q > 16 || n < 1 || n > MAXN) return 0;
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
    // outer groups: at most bmax axes each, rows of >= 64 words
    const int m = n - a;
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
