// Host check of sgld::FastDiv / NodeDiv (lattice_kernels.hpp) against integer
// division: every divisor 1..5000 with edge and pseudo-random numerators below
// 2^31, and node (row, col, layer) decompositions of a few lattices.
#include <cstdio>
#include <cstdint>

#include "../sgldbench_amd/csrc/lattice_kernels.hpp"

int main() {
    uint64_t state = 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        state ^= state << 13;
        state ^= state >> 7;
        state ^= state << 17;
        return (uint32_t)(state & 0x7FFFFFFFu);
    };
    long checks = 0;
    {
        const sgld::FastDiv a = sgld::FastDiv::make(3), b = sgld::FastDiv::make3();
        if (a.d != b.d || a.m != b.m || a.s != b.s) { std::printf("FAIL make3\n"); return 1; }
    }
    for (uint32_t d = 1; d <= 5000; ++d) {
        const sgld::FastDiv f = sgld::FastDiv::make(d);
        const uint32_t edge[] = {0u, 1u, d - 1, d, d + 1, 2 * d - 1, 2 * d, 0x7FFFFFFFu, 0x7FFFFFFFu - d,
                                 (0x7FFFFFFFu / d) * d, (0x7FFFFFFFu / d) * d - 1};
        for (uint32_t n : edge) {
            if (n > 0x7FFFFFFFu) continue;
            ++checks;
            if (f.div(n) != n / d) { std::printf("FAIL d=%u n=%u got %u\n", d, n, f.div(n)); return 1; }
        }
        for (int i = 0; i < 2000; ++i) {
            const uint32_t n = rnd();
            ++checks;
            if (f.div(n) != n / d) { std::printf("FAIL d=%u n=%u got %u\n", d, n, f.div(n)); return 1; }
        }
    }
    const int dims[][3] = {{256, 128, 128}, {512, 512, 512}, {7, 3, 5}, {1, 1, 1}};
    for (auto &dm : dims) {
        sgld::Grid g{dm[0], dm[1], dm[2]};
        const sgld::NodeDiv nd = sgld::NodeDiv::make(g);
        const int64_t N = g.nodes();
        const int64_t step = N > 4000000 ? 997 : 1;
        for (int64_t n = 0; n < N; n += step) {
            int r, c, l;
            nd.rcl((uint32_t)n, r, c, l);
            ++checks;
            if (r + g.NY() * c + g.NY() * g.NX() * l != n || r >= g.NY() || c >= g.NX() || l >= g.NZ()) {
                std::printf("FAIL grid %d %d %d node %lld\n", dm[0], dm[1], dm[2], (long long)n);
                return 1;
            }
        }
    }
    std::printf("ok %ld\n", checks);
    return 0;
}
