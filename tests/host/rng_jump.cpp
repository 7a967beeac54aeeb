// Host check of the rng jump matrices used by the four-wave ring fill: for
// random states s, J_w s (applied as the fill's waves apply it, one parity of
// row & s per result bit) must equal the state after RNG_PART * w sequential
// steps, and the RNG_SPLIT parts of draws laid end to end must be the
// RNG_CHUNK sequential draws, ending in the sequential state.
//
//   rng_jump -> prints "ok" or "FAIL ..."
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "cov_common.h"

using namespace cov;

static void apply(const uint64_t* M, const uint64_t* s, uint64_t* o)
{
    for (int k = 0; k < 4; k++) {
        uint64_t word = 0;
        for (int lane = 0; lane < 64; lane++) {
            const uint64_t* row = M + (size_t)(64 * k + lane) * 4;
            const uint64_t x = (row[0] & s[0]) ^ (row[1] & s[1]) ^ (row[2] & s[2]) ^ (row[3] & s[3]);
            word |= (uint64_t)(__builtin_popcountll(x) & 1) << lane;
        }
        o[k] = word;
    }
}

int main()
{
    std::vector<uint64_t> J(RNG_JUMP_WORDS);
    rng_jump_tables(J.data());
    uint64_t seed[4] = {0x9E3779B97F4A7C15ull, 0xD1B54A32D192ED03ull, 0x8CB92BA72F3D8DD7ull, 1};
    for (int trial = 0; trial < 64; trial++) {
        uint64_t s0[4];
        for (int k = 0; k < 4; k++) s0[k] = xo_next(seed);
        if (trial == 0) { s0[0] = 1; s0[1] = s0[2] = s0[3] = 0; }   // a single bit
        // sequential reference
        std::vector<uint64_t> seq(RNG_CHUNK);
        uint64_t st[4] = {s0[0], s0[1], s0[2], s0[3]};
        uint64_t at[RNG_SPLIT][4];
        for (int g = 0; g < RNG_CHUNK; g++) {
            if (g % RNG_PART == 0)
                for (int k = 0; k < 4; k++) at[g / RNG_PART][k] = st[k];
            seq[g] = xo_next(st);
        }
        // the parts as the fill's waves make them
        uint64_t last[4] = {0, 0, 0, 0};
        for (int w = 0; w < RNG_SPLIT; w++) {
            uint64_t s[4] = {s0[0], s0[1], s0[2], s0[3]};
            if (w > 0) {
                uint64_t o[4];
                apply(J.data() + (size_t)(w - 1) * 256 * 4, s0, o);
                for (int k = 0; k < 4; k++) {
                    if (o[k] != at[w][k]) { printf("FAIL trial %d wave %d state word %d\n", trial, w, k); return 1; }
                    s[k] = o[k];
                }
            }
            for (int i = 0; i < RNG_PART; i++) {
                const uint64_t r = xo_next(s);
                if (r != seq[(size_t)w * RNG_PART + i]) { printf("FAIL trial %d draw %d\n", trial, w * RNG_PART + i); return 1; }
            }
            if (w == RNG_SPLIT - 1)
                for (int k = 0; k < 4; k++) last[k] = s[k];
        }
        for (int k = 0; k < 4; k++)
            if (last[k] != st[k]) { printf("FAIL trial %d final state\n", trial); return 1; }
    }
    printf("ok\n");
    return 0;
}
