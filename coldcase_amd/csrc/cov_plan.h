// cov_plan.h -- host side: the shape of a cell and the tables its kernels
// index (position subsets, powers, the sphere-walk split, division magics).
// Used by cov_create and by the host emulation of the step kernels' word
// index arithmetic, so both see the same plan.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

#include "covhip.h"
#include "cov_common.h"

namespace cov {

inline int64_t binom(int a, int b)
{
    if (b < 0 || b > a) return 0;
    if (b > a - b) b = a - b;
    int64_t r = 1;
    for (int i = 0; i < b; i++) r = r * (a - i) / (i + 1);
    return r;
}

// all r-subsets of `pos`, lexicographic, packed 4 bits per position
inline std::vector<uint32_t> subsets(const std::vector<int>& pos, int r)
{
    std::vector<uint32_t> out;
    const int np = (int)pos.size();
    if (r > np) return out;
    std::vector<int> c(r);
    for (int i = 0; i < r; i++) c[i] = i;
    for (;;) {
        uint32_t s = 0;
        for (int i = 0; i < r; i++) s |= (uint32_t)pos[c[i]] << (4 * i);
        out.push_back(s);
        int i = r - 1;
        while (i >= 0 && c[i] == np - r + i) i--;
        if (i < 0) break;
        c[i]++;
        for (int j = i + 1; j < r; j++) c[j] = c[j - 1] + 1;
    }
    return out;
}

struct Shape {
    int q, n, R, M, q1, F, E;
    int64_t P, N, V, nzw, nblk, ulcap;
};

// the accepted shapes: q, n <= 16, R <= 8 and below n, M <= 65535 (16-bit
// counts), q^n <= 4e10, (q-1)^R below 2^24
inline int shape_of(int q, int n, int R, int M, Shape* s)
{
    if (q < 2 || q > MAXQ || n < 1 || n > MAXN || R < 0 || R > MAXR || R >= n || M < 1 || M > 65535)
        return COV_EPARAM;
    int64_t N = 1;
    for (int i = 0; i < n; i++) {
        if (N > (int64_t)4e10 / q) return COV_ESPACE;
        N *= q;
    }
    s->q = q; s->n = n; s->R = R; s->M = M; s->q1 = q - 1;
    s->F = (int)binom(n, R);
    s->E = (int)binom(n - 1, R);
    s->P = 1;
    for (int i = 0; i < R; i++) s->P *= (q - 1);
    if (s->P >= (1 << 24)) return COV_ELIMIT;
    s->N = N;
    s->V = 0;
    for (int r = 0; r <= R; r++) {
        int64_t t = binom(n, r);
        for (int i = 0; i < r; i++) t *= (q - 1);
        s->V += t;
    }
    s->nzw = (N + 31) / 32;
    s->nblk = (N + BLK_WORDS - 1) / BLK_WORDS;
    s->ulcap = std::min<int64_t>(N, (int64_t)1 << 20);
    return COV_OK;
}

// the position-subset tables the descriptor points to
struct SubsetTables {
    std::vector<uint32_t> fs;   // C(n,R) R-subsets of all positions
    std::vector<uint32_t> ps;   // per position p, the C(n-1,R) R-subsets avoiding p
    std::vector<uint32_t> bs;   // subsets of size 0..R, grouped by size
};

// The descriptor's scalar fields (everything but the device pointers) and
// the subset tables, for shape s.
inline void plan_cell(const Shape& s, Cell& C, SubsetTables& t)
{
    const int q = s.q, n = s.n, R = s.R;
    C.q = q; C.n = n; C.R = R; C.M = s.M; C.q1 = q - 1; C.F = s.F; C.E = s.E; C.P = s.P;
    C.N = s.N; C.nzw = s.nzw; C.nblk = s.nblk; C.V = s.V; C.ncap = (int64_t)s.M * n; C.ulcap = s.ulcap;
    C.ul_use = (int64_t)(R + 1) * s.E * s.P / n;
    for (int j = 0; j < MAXN; j++) C.pw[j] = 0;
    for (int j = 0; j < n; j++) {
        C.pw[j] = 1;
        for (int e = 0; e < n - 1 - j; e++) C.pw[j] *= q;
    }
    // t / (q-1) = umulhi(t, ceil(2^32 / (q-1))) for the pattern numbers t < (q-1)^R
    C.magic = (q - 1 == 1) ? 0u : (uint32_t)((((uint64_t)1 << 32) + (uint64_t)(q - 2)) / (uint64_t)(q - 1));
    C.nmag = magic64((uint64_t)s.N);
    C.mmag = magic64((uint64_t)s.M);
    {
        // sphere-walk tables: the innermost ceil(R/2) positions in lo
        const int a = (R + 1) / 2;
        int64_t nl = 1, nh = 1;
        for (int i = 0; i < a; i++) nl *= (q - 1);
        for (int i = a; i < R; i++) nh *= (q - 1);
        C.wa = a;
        C.wok = (R >= 1 && nl <= WT_MAX && nh <= WT_MAX) ? 1 : 0;
        C.wnl = C.wok ? (int32_t)nl : 1;
        C.wnh = C.wok ? (int32_t)nh : 1;
        C.wmag = C.wnl > 1 ? (uint32_t)((((uint64_t)1 << 32) / (uint64_t)C.wnl) + 1) : 0u;
    }
    std::vector<int> all(n);
    for (int j = 0; j < n; j++) all[j] = j;
    t.fs = subsets(all, R);
    t.ps.clear();
    for (int p = 0; p < n; p++) {
        std::vector<int> rest;
        for (int j = 0; j < n; j++) if (j != p) rest.push_back(j);
        const std::vector<uint32_t> u = subsets(rest, R);
        t.ps.insert(t.ps.end(), u.begin(), u.end());
    }
    t.bs.clear();
    for (int r = 0; r <= MAXR + 1; r++) { C.bsub_off[r] = 0; C.bpat_off[r] = 0; }
    int64_t patacc = 0;
    for (int r = 0; r <= R; r++) {
        C.bsub_off[r] = (int64_t)t.bs.size();
        C.bpat_off[r] = patacc;
        const std::vector<uint32_t> u = subsets(all, r);
        t.bs.insert(t.bs.end(), u.begin(), u.end());
        int64_t pr = 1;
        for (int i = 0; i < r; i++) pr *= (q - 1);
        patacc += (int64_t)u.size() * pr;
    }
    for (int r = R + 1; r <= MAXR + 1; r++) { C.bsub_off[r] = (int64_t)t.bs.size(); C.bpat_off[r] = patacc; }
}

}  // namespace cov
