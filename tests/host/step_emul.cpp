// Host emulation of the step kernels' word-index arithmetic: builds a cell's
// plan with the library's own host code (cov_plan.h), puts its subset tables
// in exact-size heap arrays, and runs the index functions the kernels use --
// pattern_word, the sphere-walk tables and walk_word, ball_word_at,
// pack_index, the division magics, the commit's offset table and word loop,
// and the kept-count event condition -- against a brute-force digit
// arithmetic.  Built with AddressSanitizer by tests/test_step_emul.py, so a
// table read out of its array aborts.
//
//   step_emul q n R seed centres [cap]
//     -> prints "ok checked=<words>" or "FAIL ..."; at most `cap` patterns per
//        subset family and centre are checked (all of them when they fit)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cov_plan.h"
#include "cov_cover.h"
#include "cov_step.h"

using namespace cov;

static uint64_t sm64(uint64_t& x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

#define FAIL(...) do { printf("FAIL "); printf(__VA_ARGS__); printf("\n"); exit(1); } while (0)

struct Brute {
    int q, n;
    int64_t N;
    std::vector<int> digits(int64_t w) const
    {
        std::vector<int> d(n);
        for (int j = n - 1; j >= 0; j--) { d[j] = (int)(w % q); w /= q; }
        return d;
    }
    int64_t index(const std::vector<int>& d) const
    {
        int64_t w = 0;
        for (int j = 0; j < n; j++) w = w * q + d[j];
        return w;
    }
    uint64_t packed(const std::vector<int>& d) const
    {
        uint64_t pk = 0;
        for (int j = 0; j < n; j++) pk |= (uint64_t)d[j] << (4 * j);
        return pk;
    }
};

// exact-size heap copy: an index past `size` is an out-of-bounds access
template <typename T>
static T* heap_copy(const std::vector<T>& v)
{
    T* p = (T*)malloc(std::max<size_t>(v.size(), 1) * sizeof(T));
    std::copy(v.begin(), v.end(), p);
    return p;
}

static void check_subset(const Cell& C, uint32_t sub, int r, const char* what)
{
    int prev = -1;
    for (int i = 0; i < r; i++) {
        const int j = (sub >> (4 * i)) & 15;
        if (j <= prev || j >= C.n) FAIL("%s subset %x: position %d out of order or >= n", what, sub, j);
        prev = j;
    }
}

int main(int argc, char** argv)
{
    if (argc < 6) { printf("usage: step_emul q n R seed centres [cap]\n"); return 2; }
    const int q = atoi(argv[1]), n = atoi(argv[2]), R = atoi(argv[3]);
    uint64_t rs = (uint64_t)atoll(argv[4]);
    const int centres = atoi(argv[5]);
    const int64_t cap = argc > 6 ? atoll(argv[6]) : 1000000;
    Shape s;
    const int rc = shape_of(q, n, R, 1, &s);
    if (rc != COV_OK) FAIL("shape_of -> %d", rc);
    Cell C{};
    SubsetTables tabs;
    plan_cell(s, C, tabs);
    if ((int64_t)tabs.fs.size() != C.F) FAIL("fsub holds %zu subsets, F = %d", tabs.fs.size(), C.F);
    if ((int64_t)tabs.ps.size() != (int64_t)C.n * C.E) FAIL("psub holds %zu subsets, n E = %lld", tabs.ps.size(), (long long)n * C.E);
    if (C.bsub_off[C.R + 1] != (int64_t)tabs.bs.size()) FAIL("bsub offsets");
    if (C.bpat_off[C.R + 1] != C.V) FAIL("ball pattern offsets end at %lld, V = %lld", (long long)C.bpat_off[C.R + 1], (long long)C.V);
    uint32_t* fsub = heap_copy(tabs.fs);
    uint32_t* psub = heap_copy(tabs.ps);
    uint32_t* bsub = heap_copy(tabs.bs);
    C.fsub = (decltype(C.fsub))fsub;
    C.psub = (decltype(C.psub))psub;
    C.bsub = (decltype(C.bsub))bsub;
    for (int f = 0; f < C.F; f++) check_subset(C, fsub[f], R, "fsub");
    for (int p = 0; p < n; p++)
        for (int e = 0; e < C.E; e++) {
            const uint32_t sub = psub[(int64_t)p * C.E + e];
            check_subset(C, sub, R, "psub");
            for (int i = 0; i < R; i++)
                if ((int)((sub >> (4 * i)) & 15) == p) FAIL("psub row %d holds position %d", p, p);
            // the commit derives row p from the row avoiding n-1
            const uint32_t derived = cov::sub_avoiding(psub[(int64_t)(n - 1) * C.E + e], R, p);
            if (derived != sub) FAIL("row %d subset %d: derived %x, table %x", p, e, derived, sub);
        }
    for (int r = 0; r <= R; r++)
        for (int64_t x = C.bsub_off[r]; x < C.bsub_off[r + 1]; x++) check_subset(C, bsub[x], r, "bsub");
    const Brute B{q, n, C.N};
    int64_t checked = 0;

    // the division magics over every pattern number
    for (uint32_t t = 0; t < (uint32_t)C.P; t++) {
        if (div_q1(C, t) != t / (uint32_t)C.q1) FAIL("div_q1(%u) = %u", t, div_q1(C, t));
        if (C.wok) {
            const uint32_t h = C.wnl == 1 ? t : mulhi32(t, C.wmag);
            if (h != t / (uint32_t)C.wnl || h >= (uint32_t)C.wnh) FAIL("walk split of %u: %u", t, h);
        }
    }
    if (C.wok && (int64_t)C.wnl * C.wnh != C.P) FAIL("walk tables %d x %d for P = %lld", C.wnl, C.wnh, (long long)C.P);

    // the draws' reduction mod N and mod M by multiply-high: edge operands
    // and random ones against the division, for this cell's N and a spread
    // of code sizes
    {
        const uint64_t ds[] = {(uint64_t)C.N, 2, 3, 7, 65535, 65536, 0xFFFFFFFFull, 0x100000001ull,
                               (uint64_t)4e10, (uint64_t)(sm64(rs) % 65535) + 1, (uint64_t)(sm64(rs) % (uint64_t)C.N) + 1};
        for (const uint64_t d : ds) {
            const uint64_t mag = magic64(d);
            const uint64_t xs[] = {0, 1, d - 1, d, d + 1, 2 * d - 1, ~0ull, ~0ull - 1, ~0ull / d * d, ~0ull / d * d - 1};
            for (const uint64_t x : xs)
                if (mod_magic(x, d, mag) != x % d) FAIL("mod_magic(%llu, %llu)", (unsigned long long)x, (unsigned long long)d);
            for (int i = 0; i < 200000; i++) {
                const uint64_t x = sm64(rs);
                if (mod_magic(x, d, mag) != x % d) FAIL("mod_magic(%llu, %llu)", (unsigned long long)x, (unsigned long long)d);
            }
        }
    }

    // the short remainders: every divisor 1 .. 65536 on edge and random
    // operands, and the fallback just past it
    for (uint32_t d = 1; d <= 65536; d += (d < 1100 ? 1 : 97)) {
        const uint64_t xs[] = {0, 1, d - 1u, d, ~0ull, ~0ull - 1, (uint64_t)d << 32, ((uint64_t)d << 32) - 1, sm64(rs), sm64(rs)};
        for (const uint64_t x : xs) {
            if (mod_u16(x, d) != x % d) FAIL("mod_u16(%llu, %u)", (unsigned long long)x, d);
            if (mod_any(x, d) != x % d) FAIL("mod_any(%llu, %u)", (unsigned long long)x, d);
        }
    }
    for (const uint64_t d : {65536ull, 65537ull, 1ull << 40, 999999999999ull})
        for (int i = 0; i < 1000; i++) {
            const uint64_t x = sm64(rs);
            if (mod_any(x, d) != x % d) FAIL("mod_any(%llu, %llu)", (unsigned long long)x, (unsigned long long)d);
        }

    // pack_index on the extremes and random words
    for (int i = 0; i < 1000; i++) {
        const int64_t w = i == 0 ? 0 : i == 1 ? C.N - 1 : (int64_t)(sm64(rs) % (uint64_t)C.N);
        if (pack_index(C, w) != B.packed(B.digits(w))) FAIL("pack_index(%lld)", (long long)w);
    }

    // slot_diff: the t-th position (increasing) where two words' digits
    // differ, against a list of the differing positions made digit by digit
    for (int i = 0; i < 2000; i++) {
        const int64_t a = (int64_t)(sm64(rs) % (uint64_t)C.N);
        const int64_t b = i == 0 ? a : (int64_t)(sm64(rs) % (uint64_t)C.N);
        const std::vector<int> da = B.digits(a), db = B.digits(b);
        std::vector<int> diff;
        for (int j = 0; j < n; j++)
            if (da[j] != db[j]) diff.push_back(j);
        if (nib_dist(B.packed(da), B.packed(db)) != (int)diff.size()) FAIL("nib_dist of %lld, %lld", (long long)a, (long long)b);
        for (int t = 0; t < (int)diff.size(); t++)
            if (slot_diff(B.packed(da), B.packed(db), t) != diff[t])
                FAIL("slot_diff(%lld, %lld, %d) = %d, not %d", (long long)a, (long long)b, t,
                     slot_diff(B.packed(da), B.packed(db), t), diff[t]);
    }

    // pattern t of subset `sub` around the centre: the brute-force word
    // differs from the centre exactly on the subset, there taking the k-th
    // value other than the centre's (the last position fastest)
    auto brute_pattern = [&](const std::vector<int>& cd, uint32_t sub, int r, uint32_t t) {
        std::vector<int> d = cd;
        for (int i = r - 1; i >= 0; i--) {
            const int k = (int)(t % (uint32_t)C.q1);
            t /= (uint32_t)C.q1;
            const int j = (sub >> (4 * i)) & 15;
            d[j] = k + (k >= cd[j]);
        }
        return d;
    };
    // every pattern (or `cap` random ones) of one subset family around one centre
    std::vector<int64_t> seen;
    int64_t* lo = (int64_t*)malloc(sizeof(int64_t) * (size_t)C.wnl);
    int64_t* hi = (int64_t*)malloc(sizeof(int64_t) * (size_t)C.wnh);
    auto family = [&](const std::vector<int>& cd, const uint32_t* subs, int nsub, const char* what) {
        const int64_t base = B.index(cd);
        const uint64_t pk = B.packed(cd);
        const bool all = (int64_t)nsub * C.P <= cap;
        const int64_t total = all ? (int64_t)nsub * C.P : cap;
        int cur = -1;
        for (int64_t x = 0; x < total; x++) {
            int f;
            uint32_t t;
            if (all) { f = (int)(x / C.P); t = (uint32_t)(x - (int64_t)f * C.P); }
            else { f = (int)(sm64(rs) % (uint64_t)nsub); t = (uint32_t)(sm64(rs) % (uint64_t)C.P); }
            const uint32_t sub = subs[f];
            if (f != cur && C.wok) { walk_fill(C, pk, sub, lo, hi, 0, 1); }
            if (all && f != cur) {
                if (cur >= 0) {
                    std::sort(seen.begin(), seen.end());
                    if (std::unique(seen.begin(), seen.end()) != seen.end()) FAIL("%s: repeated word in subset %d", what, cur);
                }
                seen.clear();
            }
            cur = f;
            uint64_t wpk;
            const int64_t w = pattern_word(C, base, pk, sub, R, t, &wpk);
            if (w < 0 || w >= C.N) FAIL("%s: pattern_word %lld out of [0, N)", what, (long long)w);
            const std::vector<int> want = brute_pattern(cd, sub, R, t);
            if (w != B.index(want) || wpk != B.packed(want)) FAIL("%s: pattern %u of subset %x", what, t, sub);
            if (pattern_digits(C, pk, sub, R, t) != wpk) FAIL("%s: pattern_digits(%u) of subset %x", what, t, sub);
            if (nib_dist(wpk, pk) != R) FAIL("%s: distance %d", what, nib_dist(wpk, pk));
            if (C.wok && walk_word(C, lo, hi, base, t) != w) FAIL("%s: walk_word(%u) differs", what, t);
            if (all) seen.push_back(w);
            checked++;
        }
        if (all && !seen.empty()) {
            std::sort(seen.begin(), seen.end());
            if (std::unique(seen.begin(), seen.end()) != seen.end()) FAIL("%s: repeated word", what);
        }
    };

    int64_t* T = (int64_t*)malloc(sizeof(int64_t) * (size_t)n * 16);
    for (int ci = 0; ci < centres; ci++) {
        const int64_t cw = ci == 0 ? 0 : ci == 1 ? C.N - 1 : (int64_t)(sm64(rs) % (uint64_t)C.N);
        const std::vector<int> cd = B.digits(cw);
        const uint64_t pk = B.packed(cd);
        // leaving side: the distance-R sphere of the codeword, subset by subset
        family(cd, fsub, C.F, "leave");
        // entering side: the moved codeword c(p := v), subsets avoiding p
        for (int p = 0; p < n; p++) {
            int v = (int)(sm64(rs) % (uint64_t)q);
            if (v == cd[p]) v = (v + 1) % q;
            std::vector<int> md = cd;
            md[p] = v;
            const int64_t mb = cw + (int64_t)(v - cd[p]) * C.pw[p];
            if (mb != B.index(md) || nib_put(pk, p, v) != B.packed(md)) FAIL("moved centre of position %d", p);
            if (C.E) family(md, psub + (int64_t)p * C.E, C.E, "enter");
        }
        // the commit of move (bp, bv): a through the offset table, b = a + dv
        for (int x = 0; x < n * 16; x++) T[x] = 0;
        for (int j = 0; j < n; j++)
            for (int k = 0; k < C.q1; k++) T[j * 16 + k] = commit_offset(C, pk, j, k);
        for (int bp = 0; bp < n && C.E; bp++) {
            int bv = (int)(sm64(rs) % (uint64_t)q);
            if (bv == cd[bp]) bv = (bv + 1) % q;
            const int64_t dv = (int64_t)(bv - cd[bp]) * C.pw[bp];
            const uint64_t mpk = nib_put(pk, bp, bv);
            const int64_t S = (int64_t)C.E * C.P;
            const bool all = S <= cap;
            for (int64_t y = 0; y < (all ? S : cap); y++) {
                const int64_t g = all ? y : (int64_t)(sm64(rs) % (uint64_t)S);
                const int64_t e = g / C.P;
                const uint32_t t = (uint32_t)(g - e * C.P);
                const uint32_t sub = psub[(int64_t)bp * C.E + e];
                uint64_t apk;
                const int64_t a = commit_word(C, T, cw, sub, t);
                if (a != pattern_word(C, cw, pk, sub, R, t, &apk)) FAIL("commit word %lld of move (%d,%d)", (long long)g, bp, bv);
                const int64_t b = a + dv;
                std::vector<int> bd = B.digits(a);
                bd[bp] = bv;
                if (a < 0 || a >= C.N || b < 0 || b >= C.N || b != B.index(bd)) FAIL("commit pair (%lld, %lld)", (long long)a, (long long)b);
                // the kept-count events: a at distance R from the old codeword,
                // b at distance R from the moved one, the slots 1 + position
                const uint64_t bpk = nib_put(apk, bp, bv);
                if (nib_dist(apk, pk) != R || nib_dist(bpk, mpk) != R) FAIL("commit words off the spheres");
                for (uint64_t m = nib_ne(apk, pk); m; m &= m - 1) {
                    const int slot = 1 + (__builtin_ctzll(m) >> 2);
                    if (slot < 1 || slot > n || slot >= THW) FAIL("kept-count slot %d", slot);
                }
                checked++;
            }
        }
        // the ball: V distinct words within distance R
        {
            std::vector<int64_t> ball;
            ball.reserve((size_t)std::min<int64_t>(C.V, cap));
            const bool all = C.V <= cap;
            for (int64_t y = 0; y < (all ? C.V : cap); y++) {
                const int64_t g = all ? y : (int64_t)(sm64(rs) % (uint64_t)C.V);
                uint64_t wpk;
                const int64_t w = ball_word_at(C, cw, pk, g, &wpk);
                if (w < 0 || w >= C.N || wpk != B.packed(B.digits(w)) || nib_dist(wpk, pk) > R)
                    FAIL("ball word %lld -> %lld", (long long)g, (long long)w);
                if (all) ball.push_back(w);
                checked++;
            }
            if (all) {
                std::sort(ball.begin(), ball.end());
                if (std::unique(ball.begin(), ball.end()) != ball.end()) FAIL("repeated ball word");
            }
        }
    }
    free(T);
    free(lo);
    free(hi);
    free(fsub);
    free(psub);
    free(bsub);
    printf("ok checked=%lld P=%lld wok=%d\n", (long long)checked, (long long)C.P, C.wok);
    return 0;
}
