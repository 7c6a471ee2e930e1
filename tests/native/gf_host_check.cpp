// Host build of the library's GF(2^32) helpers (gf.hpp) for the CPU tests:
// reads pairs "a b" (decimal) from stdin and prints, per pair, the
// carry-less product by integer multiplies (clmul32_holes) and the reduced
// product; then compares clmul32_holes with the bit-serial product on
// random pairs and prints the number of mismatches.
#include <cstdint>
#include <cstdio>
#include <random>

#include "../../ligerito-rust_amd/csrc/gf.hpp"

int main() {
    unsigned long long a, b;
    while (std::scanf("%llu %llu", &a, &b) == 2) {
        const uint64_t p = lig::clmul32_holes((uint32_t)a, (uint32_t)b);
        std::printf("%llu %u\n", (unsigned long long)p, lig::gf32_reduce64(p));
    }
    std::mt19937_64 r(5);
    unsigned long long bad = 0;
    for (int i = 0; i < 2000000; i++) {
        const uint64_t v = r();
        const uint32_t x = (uint32_t)v, y = (uint32_t)(v >> 32);
        bad += lig::clmul32_holes(x, y) != lig::clmul32_portable(x, y);
    }
    std::printf("mismatches %llu\n", bad);
    return 0;
}
