// Host-side check of the device neighbour walk (sphfluidsimulation_amd/csrc/sph_math.hpp):
// for random cells and bucket counts M, the buckets walk_runs visits (with
// multiplicity) are exactly the buckets of the reference's 27-cell loop
// (SPHFunctions.cpp:397-404), fastmod equals %, and the dilated-arithmetic
// neighbour codes equal morton3 of the neighbour cells. Exit status 0 on success.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "../../sphfluidsimulation_amd/csrc/sph_math.hpp"

using namespace sph;

int main() {
    std::mt19937_64 rng(7);
    long bad = 0, checked = 0;
    // fastmod against the hardware remainder
    for (int t = 0; t < 200000; ++t) {
        uint32_t d = (uint32_t)(rng() % 3 == 0 ? rng() % 64 + 1 : rng() | 1u);
        if (t % 5 == 0) d = (uint32_t)(rng() >> 33) + 1u;
        const uint64_t m = fastmod_magic(d);
        for (int k = 0; k < 8; ++k) {
            const uint32_t a = (uint32_t)rng();
            bad += fastmod(a, m, d) != a % d;
            ++checked;
        }
    }
    // The cell-wave passes' neighbour codes: dilated arithmetic on the cell's
    // own code against morton3 of the offset coordinates, in the 27-cell order.
    for (int t = 0; t < 200000; ++t) {
        uint32_t c[3];
        for (int a = 0; a < 3; ++a) {
            const int kind = (int)(rng() % 3);
            c[a] = kind == 0 ? (uint32_t)(rng() % 4) - 2u : kind == 1 ? 1022u + (uint32_t)(rng() % 4)
                                                                      : (uint32_t)rng();
        }
        const uint32_t self = morton3(c[0], c[1], c[2]);
        for (uint32_t n = 0; n < 27; ++n) {
            const uint32_t want = morton3(c[0] + n / 9u - 1u, c[1] + (n / 3u) % 3u - 1u, c[2] + n % 3u - 1u);
            bad += neighbour_code(self, neighbour_offsets(n)) != want;
            ++checked;
        }
    }
    std::vector<uint32_t> Ms = {1, 2, 3, 5, 6, 7, 8, 9, 15, 16, 17, 31, 32000, 32002, 2097152, 16777216, 1000003};
    for (int t = 0; t < 40; ++t) Ms.push_back((uint32_t)(rng() % 5000) + 1u);
    for (int t = 0; t < 10; ++t) Ms.push_back((uint32_t)(rng() >> 34) + 1u);
    std::vector<uint32_t> table;
    for (uint32_t M : Ms) {
        const uint64_t magic = fastmod_magic(M);
        const uint32_t tabN = M < 100000 ? M : 0;  // identity-ish table only when small
        for (int t = 0; t < 3000; ++t) {
            uint32_t c[3];
            for (int a = 0; a < 3; ++a) {
                const int kind = (int)(rng() % 4);
                c[a] = kind == 0 ? (uint32_t)(rng() % 4) - 2u           // around 0 (wraps)
                       : kind == 1 ? 1022u + (uint32_t)(rng() % 4)       // around the 1024 alias
                       : kind == 2 ? (uint32_t)rng()                     // anywhere in uint32
                                   : (uint32_t)(rng() % 300);
            }
            std::vector<uint32_t> want;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz)
                        want.push_back(morton3(c[0] + (uint32_t)dx, c[1] + (uint32_t)dy, c[2] + (uint32_t)dz) % M);
            // table[b] = 4 b (scaled), so a range [4 b0, 4 b1) names buckets b0 .. b1-1
            std::vector<uint32_t> got;
            std::vector<uint32_t> tab;
            if (tabN) {
                tab.resize((size_t)M + 1);
                for (uint32_t b = 0; b <= M; ++b) tab[b] = 4u * b;
            }
            const uint32_t* tp = tabN ? tab.data() : nullptr;
            if (!tp) {
                // large M: evaluate the walk's bucket arithmetic with a sparse table
                const CellPair X = axis_pair(c[0], 0), Y = axis_pair(c[1], 1), Z = axis_pair(c[2], 2);
                for (uint32_t r = 0; r < 14; ++r) {
                    uint32_t code, len;
                    run_of(r, X, Y, Z, code, len);
                    const uint32_t b0 = fastmod(code, magic, M);
                    const bool wraps = b0 + len > M;
                    for (uint32_t part = 0; part < (wraps ? len : 1u); ++part) {
                        const uint32_t b = wraps ? fastmod(b0 + part, magic, M) : b0;
                        const uint32_t bl = wraps ? 1u : len;
                        for (uint32_t q = 0; q < bl; ++q) got.push_back(b + q);
                    }
                }
            } else {
                walk_runs(c[0], c[1], c[2], M, magic, tp, 0u, 0xFFFFFFFFu, [&](uint32_t s, uint32_t e) {
                    for (uint32_t q = s; q < e; q += 4u) got.push_back(q / 4u);
                });
            }
            std::sort(want.begin(), want.end());
            std::sort(got.begin(), got.end());
            if (want != got) {
                if (bad < 5)
                    fprintf(stderr, "mismatch M=%u cell=(%u,%u,%u): %zu vs %zu buckets\n", M, c[0], c[1], c[2],
                            want.size(), got.size());
                ++bad;
            }
            ++checked;
        }
    }
    printf("walk_check: %ld checks, %ld mismatches\n", checked, bad);
    return bad ? 1 : 0;
}
