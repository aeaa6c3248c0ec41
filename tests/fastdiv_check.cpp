// Brute-force check of lbm::FastDiv against the hardware division: every
// divisor 1 .. 70000 and a spread of larger ones, dividends 0 .. 4095, both
// sides of multiples of the divisor, and random 32-bit values.  Prints the
// first mismatch and exits 1, or prints the count checked and exits 0.
#include <cstdio>
#include <cstdint>
#include <random>

#include "fastdiv.h"

static int check(uint32_t d, uint32_t n, unsigned long long& count)
{
    const lbm::FastDiv f = lbm::FastDiv::of(d);
    ++count;
    if (f.div(n) != n / d) {
        std::printf("mismatch: n=%u d=%u got %u want %u\n", n, d, f.div(n), n / d);
        return 1;
    }
    return 0;
}

int main()
{
    std::mt19937_64 rng(12345);
    unsigned long long count = 0;
    for (uint32_t d = 1; d <= 70000; ++d) {
        for (uint32_t n = 0; n < 4096; n += (d < 64 ? 1 : 61)) if (check(d, n, count)) return 1;
        for (int r = 0; r < 24; ++r) {
            const uint64_t k = rng() % (0xffffffffull / d + 1);
            const uint64_t base = k * d;
            for (int o = -1; o <= 1; ++o) {
                const uint64_t n = base + o;
                if (n <= 0xffffffffull && check(d, (uint32_t)n, count)) return 1;
            }
            if (check(d, (uint32_t)rng(), count)) return 1;
        }
        if (check(d, 0xffffffffu, count)) return 1;
    }
    for (int r = 0; r < 200000; ++r) {
        const uint32_t d = (uint32_t)(rng() % 0xffffffffull) + 1;
        if (check(d, (uint32_t)rng(), count) || check(d, d - 1, count) || check(d, d, count)) return 1;
    }
    std::printf("ok %llu\n", count);
    return 0;
}
