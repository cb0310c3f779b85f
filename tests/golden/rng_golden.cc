// rng_golden.cc -- writes probe-sign fixtures from the C++ standard library.
//
// The reference draws its probes with std::mt19937 and
// std::uniform_real_distribution<double>(0,1), +1 iff the draw is > 0.5
// (DistrRngStd.hpp:27-40).  This program calls exactly those library routines
// (libstdc++ of the compiler that builds it) on one mt19937(seed) stream per
// run, cell c of probe k taking outputs 2Nk+2c and 2Nk+2c+1, and prints one
// JSON object per (seed, N, probes) case:
//   plus[k]   number of +1 cells in probe k
//   wsum[k]   sum_i w_i*(2i+1) mod 2^64 over the probe's sign words, where bit
//             (c mod 32) of word c/32 is 1 iff cell c is +1
//   head      the first min(N,4096) signs of probe 0 as 32-bit words (hex)
// plus the standard's known answer for mt19937 (the 10000th output of a
// default-constructed engine is 4123659995).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

static void run_case(uint32_t seed, long n, int probes, bool last) {
    std::mt19937 eng(seed);
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    printf("  {\"seed\": %u, \"ncells\": %ld, \"probes\": %d, \"plus\": [", seed, n, probes);
    std::vector<uint64_t> wsums;
    std::vector<uint32_t> head;
    for (int k = 0; k < probes; k++) {
        long plus = 0;
        uint64_t wsum = 0;
        uint32_t word = 0;
        for (long c = 0; c < n; c++) {
            const bool pos = dist(eng) > 0.5;
            plus += pos;
            if (pos) word |= 1u << (c & 31);
            if ((c & 31) == 31 || c == n - 1) {
                const uint64_t i = (uint64_t)(c >> 5);
                wsum += (uint64_t)word * (2 * i + 1);
                if (k == 0 && c < 4096) head.push_back(word);
                word = 0;
            }
        }
        printf("%s%ld", k ? ", " : "", plus);
        wsums.push_back(wsum);
    }
    printf("], \"wsum\": [");
    for (size_t k = 0; k < wsums.size(); k++) printf("%s\"%llu\"", k ? ", " : "", (unsigned long long)wsums[k]);
    printf("], \"head\": [");
    for (size_t k = 0; k < head.size(); k++) printf("%s\"%08x\"", k ? ", " : "", head[k]);
    printf("]}%s\n", last ? "" : ",");
}

int main() {
    std::mt19937 def;
    uint32_t v = 0;
    for (int i = 0; i < 10000; i++) v = def();
    printf("{\"mt19937_default_10000th\": %u,\n \"cases\": [\n", v);
    struct C { uint32_t seed; long n; int probes; } cases[] = {
        {1234, 32L * 32 * 32, 8},          // 32^3, 8 samples
        {5, 13L * 7 * 5, 3},               // odd grid
        {7, 256L * 256 * 256, 16},         // 256^3, 16 probes
        {11, 256L * 256 * 256, 64},        // 256^3, 64 probes
        {23, 512L * 512 * 512, 32},        // 512^3, 32 probes
        {0, 64L * 64 * 64, 8}, {1, 96L * 96 * 96, 8},
        {2, 128L * 128 * 128, 8}, {3, 192L * 192 * 192, 8},
    };
    const int nc = sizeof(cases) / sizeof(cases[0]);
    for (int i = 0; i < nc; i++) run_case(cases[i].seed, cases[i].n, cases[i].probes, i == nc - 1);
    printf("]}\n");
    return 0;
}
