// Host emulation of the distance-transform kernels: runs the pass plan and the
// per-thread phases of cov_dist.h thread by thread (phases separated where the
// kernels place a barrier; column passes one thread after another, through
// range-checked accessors) over exact-size heap buffers, and compares the
// result with a brute-force minimum distance.  Built with AddressSanitizer by
// tests/test_dt_emul.py, so any access outside a buffer aborts.
//
//   dt_emul q n M seed budget [threads [light_rcut [columns]]]
//     -> prints "ok passes=P hist=..." or "FAIL ..."; light_rcut < 0 is the full
//        histogram, columns = 0 keeps every outer pass in LDS tiles
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <vector>
#include "cov_dist.h"

using namespace cov;

// the distance array as the column passes' buffer descriptor sees it: byte
// offsets, every access checked against the descriptor's range
struct HostMem {
    uint8_t* D;
    uint64_t nbytes;
    void chk(uint64_t at, int len) const
    {
        if (at + len > nbytes) { printf("FAIL column access %llu+%d past %llu\n", (unsigned long long)at, len, (unsigned long long)nbytes); exit(1); }
    }
    uint32_t ld(uint32_t off, uint32_t soff) const
    {
        const uint64_t at = (uint64_t)off + soff;
        chk(at, 4);
        uint32_t v;
        memcpy(&v, D + at, 4);
        return v;
    }
    void st(uint32_t off, uint32_t soff, uint32_t v) const
    {
        const uint64_t at = (uint64_t)off + soff;
        chk(at, 4);
        memcpy(D + at, &v, 4);
    }
    void st8(uint32_t off, uint32_t soff, uint32_t v) const
    {
        const uint64_t at = (uint64_t)off + soff;
        chk(at, 1);
        D[at] = (uint8_t)v;
    }
};

static uint64_t sm64(uint64_t& x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int main(int argc, char** argv)
{
    if (argc < 6) return 2;
    const int q = atoi(argv[1]), n = atoi(argv[2]), M = atoi(argv[3]);
    uint64_t seed = strtoull(argv[4], 0, 10);
    const int budget = atoi(argv[5]);
    int64_t N = 1;
    for (int j = 0; j < n; j++) N *= q;
    std::vector<int64_t> cw(M);
    std::vector<std::vector<int>> dig(M, std::vector<int>(n));
    for (int i = 0; i < M; i++) {
        int64_t w = 0;
        for (int j = 0; j < n; j++) {
            dig[i][j] = (int)(sm64(seed) % (uint64_t)q);
            w = w * q + dig[i][j];
        }
        cw[i] = w;
    }
    DtPass pass[MAXN + 1];
    const bool columns = argc > 8 ? atoi(argv[8]) != 0 : true;
    const bool tcol = argc > 9 ? atoi(argv[9]) != 0 : true;
    const int ncu = argc > 10 ? atoi(argv[10]) : 256;
    const int np = dt_plan(q, n, N, budget, pass, columns, tcol, ncu);
    if (np <= 0) { printf("FAIL plan\n"); return 1; }

    int ntc = 0;                        // two-stage column passes
    int axes = 0;                       // the inner pass's rows hold the innermost axis, a seed pass's groups ns axes
    for (int i = 0; i < np; i++) axes += pass[i].nax + (pass[i].seed ? pass[i].ns : (pass[i].inner ? 1 : 0));
    if (axes != n) { printf("FAIL axes %d\n", axes); return 1; }
    // the device allocation: N + 64 bytes, 16-aligned
    uint4* Dw = new uint4[(N + 64) / 16];
    uint8_t* D = reinterpret_cast<uint8_t*>(Dw);
    memset(D, 0xAB, (N + 64) / 16 * 16);
    const int64_t nw = (N >> 5) + 2;   // as the device allocation
    uint32_t* cwb = new uint32_t[nw];
    memset(cwb, 0, nw * 4);
    for (int i = 0; i < M; i++) cwb[cw[i] >> 5] |= 1u << (cw[i] & 31);
    uint64_t hist[DT_HB] = {0};
    const int nth = argc > 6 ? atoi(argv[6]) : WG;
    const int light = argc > 7 ? atoi(argv[7]) : -1;     // >= 0: the light statistics at rcut = light
    pass[np - 1].full = light < 0;
    pass[np - 1].rcut = light < 0 ? 0 : light;
    uint64_t lt_le = 0, lt_max = 0;
    int64_t sparse_tiles = 0, empty_tiles = 0, dense_tiles = 0;     // seed tiles by the path they took
    auto flush = [&](const DtPass& p, const DtHist& h, const DtStat& st, const unsigned int* lh) -> bool {
        if (!p.full) {
            lt_le += st.le;
            lt_max = std::max<uint64_t>(lt_max, dt_stat_max(st));
            return true;
        }
        uint32_t prev = 0;
        for (int b = 0; b < DT_HL; b++) {
            hist[b <= n ? b : n + 1] += h.le[b] - prev;
            prev = h.le[b];
        }
        for (int b = 0; b < DT_HB; b++) hist[b] += lh[b];
        if (h.all != prev) { printf("FAIL hist all %u le7 %u\n", h.all, prev); return false; }
        return true;
    };
    for (int pi = 0; pi < np; pi++) {
        const DtPass& p = pass[pi];
        if (p.tcol) {
            // the two-stage pass: per tile, stage A for every thread, then stage B
            const int64_t R1 = dt_ipow(p.q, p.b1), R2 = p.nrow / R1;
            if (pi == 0 || p.col || R1 > 64 || R2 > 64 || R1 * R2 != p.nrow || p.nbytes != (uint64_t)N + 64 ||
                p.lds != 4 * p.nrow * p.Wd || p.lds > DT_TC_LDS || p.nt > dt_tc_ntmax(p.q, p.b1, p.nax - p.b1) || p.nt % 64 ||
                p.nt < std::max(R1, R2) * p.Wd || p.Wd < 4 || (int64_t)p.nchunk * 4 * p.Wd < p.S_row ||
                (int64_t)(p.nchunk - 1) * 4 * p.Wd >= p.S_row || p.S_tile != p.nrow * p.S_row ||
                p.ntiles * p.S_tile != N * p.nchunk || p.grid < 1 || p.grid > p.ntiles) {
                printf("FAIL two-stage plan\n");
                return 1;
            }
            ntc++;
            const HostMem m{D, (uint64_t)p.nbytes};
            std::vector<uint32_t> lds((size_t)p.lds / 4, 0xDEADBEEFu);
            bool ok = true;
            std::vector<uint8_t> seen(p.ntiles, 0);     // every tile exactly once
            const bool known = dt_tcol_dispatch(p.q, p.b1, p.nax - p.b1, [&](auto qb) {
                using QB = decltype(qb);
                // workgroup g takes tiles g, g + grid, ...: emulated in that order
                for (int g = 0; g < p.grid && ok; g++) {
                    DtHist h{};
                    DtStat st{};
                    unsigned int lh[DT_HB] = {0};
                    std::vector<DtHist> hs(p.nt);
                    std::vector<DtStat> sts(p.nt);
                    for (auto& x : hs) memset(&x, 0, sizeof x);
                    for (auto& x : sts) memset(&x, 0, sizeof x);
                    for (int64_t k = g; k < p.ntiles; k += p.grid) {
                        const uint32_t t = dt_tc_slot_tile((uint32_t)k, (uint32_t)p.ntiles);
                        if (t >= (uint32_t)p.ntiles || seen[t]++) { printf("FAIL slot %lld -> tile %u\n", (long long)k, t); ok = false; break; }
                        uint32_t base, x0;
                        dt_tc_tile(p, t, base, x0);
                        for (int tid = 0; tid < p.nt; tid++)
                            dt_tc_stage_a<QB::q, QB::b1, QB::b2>(p, base, x0, tid, m, lds.data());
                        for (int tid = 0; tid < p.nt; tid++)
                            dt_tc_stage_b<QB::q, QB::b1, QB::b2>(p, base, x0, tid, m, lds.data());
                        if (p.last)
                            for (int tid = 0; tid < p.nt; tid++) dt_tc_stats(p, x0, tid, p.nt, lds.data(), hs[tid], sts[tid], lh);
                    }
                    if (p.last)
                        for (int tid = 0; tid < p.nt && ok; tid++) {
                            unsigned int none[DT_HB] = {0};
                            ok = flush(p, hs[tid], sts[tid], tid == 0 ? lh : none);
                        }
                    (void)h;
                    (void)st;
                }
            });
            if (!known) { printf("FAIL no two-stage kernel q=%d b1=%d b2=%d\n", p.q, p.b1, p.nax - p.b1); return 1; }
            if (!ok) return 1;
            for (int64_t t = 0; t < p.ntiles; t++)
                if (seen[t] != 1) { printf("FAIL tile %lld taken %d times\n", (long long)t, seen[t]); return 1; }
            continue;
        }
        if (p.col) {
            if (pi == 0 || p.nthr != (uint32_t)p.ntiles || p.nrow > DT_COL_MAX || p.nbytes != (uint64_t)N + 64 ||
                (int64_t)p.ncol * 4 < p.S_row || p.S_tile != p.nrow * p.S_row) {
                printf("FAIL column plan\n");
                return 1;
            }
            const HostMem m{D, (uint64_t)p.nbytes};
            bool ok = true;
            const bool known = dt_col_dispatch(p.q, p.nax, [&](auto qb) {
                using QB = decltype(qb);
                for (uint32_t t = 0; t < p.nthr && ok; t++) {
                    DtHist h{};
                    DtStat st{};
                    unsigned int lh[DT_HB] = {0};
                    dt_col_thread<QB::q, QB::b>(p, t, m, h, st, lh);
                    if (p.last) ok = flush(p, h, st, lh);
                }
            });
            if (!known) { printf("FAIL no column kernel q=%d b=%d\n", p.q, p.nax); return 1; }
            if (!ok) return 1;
            continue;
        }
        const int64_t lds = dt_lds_bytes(p);
        if (lds + 4 * DT_HB > 65536 || lds > budget) { printf("FAIL lds %lld\n", (long long)lds); return 1; }
        if ((p.inner != 0 || p.seed != 0) != (pi == 0) || (p.first != 0) != (pi == 0)) { printf("FAIL flags\n"); return 1; }
        if (p.seed) {
            if (p.G > 64 || p.P % 4 || p.P < p.G || p.nbytes != (uint64_t)N + 64 || p.nrow * p.P > p.lds ||
                p.sk < 0 || p.sk > DT_SK || (p.sk && (p.G < 4 || p.Ls < p.G + 8 || p.Ls % 4 || p.hoff % 4 ||
                                                      p.hoff < DT_SPL + p.sk * p.Ls || p.Hs < p.nrow + 1 || p.hoff + p.sk * p.Hs > p.lds ||
                                                      p.Pd * 4 != p.Rp * p.G || p.Pd > 64 ||
                                                      p.rb * (int64_t)p.G * (p.S_tile / (p.rb * (int64_t)p.G)) != p.S_tile))) {
                printf("FAIL seed plan\n");
                return 1;
            }
            const HostMem m{D, (uint64_t)p.nbytes};
            std::vector<uint32_t> smv((size_t)lds / 4);
            uint8_t* sm = reinterpret_cast<uint8_t*>(smv.data());
            bool ok = true;
            const bool known = dt_seed_dispatch(p.q, [&](auto qb) {
                using QB = decltype(qb);
                for (int64_t t = 0; t < p.ntiles && ok; t++) {
                    const DtTile T = dt_tile(p, t);
                    // the closed form for tiles with at most sk codewords, as k_dt_seed
                    int k = p.sk + 1;
                    if (p.sk) {
                        int ncw = 0;
                        for (int tid = 0; tid < nth; tid++)
                            dt_sp_collect(p, T, tid, nth, cwb, reinterpret_cast<int32_t*>(sm), &ncw);
                        k = ncw;
                    }
                    if (k <= p.sk) {
                        (k ? sparse_tiles : empty_tiles)++;
                        if (k)
                            for (int tid = 0; tid < nth; tid++)
                                dt_sp_tables(p, k, reinterpret_cast<const int32_t*>(sm), tid, nth, sm);
                        for (int tid = 0; tid < nth && ok; tid++) {
                            DtHist h{};
                            DtStat st{};
                            unsigned int lh[DT_HB] = {0};
                            dt_sp_store(p, T, tid, nth, sm, k, m, h, st, lh);
                            if (p.last) ok = flush(p, h, st, lh);
                        }
                        continue;
                    }
                    dense_tiles++;
                    bool any = false;
                    for (uint32_t gi = 0; gi < (uint32_t)p.nrow; gi++) any |= dt_seed_group<QB::q>(p, T, gi, sm, cwb);
                    if (any)
                        for (int ax = 0; ax < p.nax; ax++)
                            for (int tid = 0; tid < nth; tid++) dt_seed_axis<QB::q>(p, ax, tid, nth, sm);
                    for (int tid = 0; tid < nth && ok; tid++) {
                        DtHist h{};
                        DtStat st{};
                        unsigned int lh[DT_HB] = {0};
                        dt_seed_store(p, T, tid, nth, sm, any, m, h, st, lh);
                        if (p.last) ok = flush(p, h, st, lh);
                    }
                }
            });
            if (!known) { printf("FAIL no seed kernel q=%d\n", p.q); return 1; }
            if (!ok) return 1;
            continue;
        }
        uint4* smv = new uint4[lds / 16];
        uint8_t* sm = reinterpret_cast<uint8_t*>(smv);
        for (int64_t t = 0; t < p.ntiles; t++) {
            const DtTile T = dt_tile(p, t);
            std::vector<DtHist> h(nth);
            for (auto& x : h) memset(&x, 0, sizeof x);     // per tile here; the kernel keeps them across tiles
            for (int tid = 0; tid < nth; tid++) dt_load(p, T, tid, nth, sm, reinterpret_cast<const uint4*>(D), cwb);
            for (int tid = 0; tid < nth; tid++) dt_realign(p, T, tid, nth, sm);
            if (p.inner)
                for (int tid = 0; tid < nth; tid++) dt_inrow(p, tid, nth, sm);
            for (int ax = 0; ax < p.nax; ax++)
                for (int tid = 0; tid < nth; tid++) dt_axis(p, T, ax, tid, nth, sm);
            for (int tid = 0; tid < nth; tid++) dt_store(p, T, tid, nth, sm, D);
            if (p.last)
                for (int tid = 0; tid < nth; tid++) {
                    unsigned int lh[DT_HB] = {0};
                    DtStat st{};
                    dt_hist_tile(p, T, tid, nth, sm, h[tid], st, lh);
                    if (!flush(p, h[tid], st, lh)) return 1;
                }
        }
        delete[] smv;
    }
    // brute force on small spaces; on large ones the textbook axis DP, one
    // full-array pass per axis
    std::vector<uint8_t> ref;
    if (N * (int64_t)M > 400000000LL) {
        ref.assign(N, 255);
        for (int i = 0; i < M; i++) ref[cw[i]] = 0;
        int64_t s = 1;
        for (int j = n - 1; j >= 0; j--, s *= q)
            for (int64_t hi = 0; hi < N; hi += s * q)
                for (int64_t lo = 0; lo < s; lo++) {
                    int m = 255;
                    for (int v = 0; v < q; v++) m = std::min<int>(m, ref[hi + lo + v * s]);
                    const int m1 = m < 255 ? m + 1 : 255;
                    for (int v = 0; v < q; v++) ref[hi + lo + v * s] = (uint8_t)std::min<int>(ref[hi + lo + v * s], m1);
                }
    }
    std::vector<int> d(n);
    uint64_t bh[DT_HB] = {0};
    int64_t bad = 0;
    for (int64_t w = 0; w < N; w++) {
        int best = 255;
        if (!ref.empty()) {
            best = ref[w];
        } else {
            int64_t x = w;
            for (int j = n - 1; j >= 0; j--) { d[j] = (int)(x % q); x /= q; }
            for (int i = 0; i < M; i++) {
                int dd = 0;
                for (int j = 0; j < n; j++) dd += d[j] != dig[i][j];
                if (dd < best) best = dd;
            }
        }
        if (best == 255) best = 127;    // unreached (no codeword), never with M >= 1
        if (D[w] != best) { if (bad < 5) printf("mismatch w=%lld got %d want %d\n", (long long)w, D[w], best); bad++; }
        bh[best <= n ? best : n + 1]++;
    }
    if (light >= 0) {
        uint64_t want_le = 0, want_max = 0;
        for (int b = 0; b <= n + 1; b++) {
            if (b <= light) want_le += bh[b];
            if (bh[b]) want_max = b;
        }
        if (want_le != lt_le || want_max != lt_max) {
            printf("FAIL light le %llu/%llu max %llu/%llu\n", (unsigned long long)lt_le, (unsigned long long)want_le,
                   (unsigned long long)lt_max, (unsigned long long)want_max);
            return 1;
        }
        for (int b = 0; b <= n + 1; b++) hist[b] = bh[b];
    }
    for (int b = 0; b <= n + 1; b++)
        if (bh[b] != hist[b]) { printf("FAIL hist bin %d %llu vs %llu\n", b, (unsigned long long)hist[b], (unsigned long long)bh[b]); return 1; }
    for (int64_t k = N; k < (N + 64) / 16 * 16; k++)
        if (D[k] != 0xAB) { printf("FAIL wrote past N at %lld\n", (long long)k); return 1; }
    if (bad) { printf("FAIL %lld mismatches\n", (long long)bad); return 1; }
    printf("ok passes=%d tiles=", np);
    for (int i = 0; i < np; i++) printf("%lld%s", (long long)pass[i].ntiles, i + 1 < np ? "," : "");
    printf(" seed=%lld/%lld/%lld tcol=%d", (long long)empty_tiles, (long long)sparse_tiles, (long long)dense_tiles, ntc);
    if (pass[0].seed && pass[0].rb) printf(" blocks/tile=%lld", (long long)(pass[0].nrow / pass[0].rb));
    printf(" hist=");
    for (int b = 0; b <= n; b++) printf("%llu ", (unsigned long long)hist[b]);
    printf("\n");
    delete[] Dw;
    delete[] cwb;
    return 0;
}
