// Checks ratio_exceeds (strobealign_amd/csrc/sa_seed.hpp), the division-free
// form of the chainer's diagonal-ratio test (float)mx / (float)mn > R, against
// the float division itself: every (mx, mn) up to 2100 and random pairs up to
// 2^24, for R = 10 (the default), thresholds whose significand is odd (ties
// round up), integers, fractions and values just beside them.
// Prints "ok <cases>" or the first failing case, and exits non-zero on failure.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "sa_seed.hpp"

using namespace sa;

int main() {
    std::vector<float> Rs = {10.0f, 1.0f, 2.5f, 3.0f, 0.75f, 7.0f, 1.5f, 100.0f, 1e-3f, 1e6f};
    Rs.push_back(std::nextafter(10.0f, 20.0f));
    Rs.push_back(std::nextafter(10.0f, 0.0f));
    Rs.push_back(std::nextafter(3.0f, 20.0f));
    Rs.push_back(1.0f + std::ldexp(1.0f, -23));  // odd significand
    uint64_t cases = 0;
    std::mt19937_64 rng(7);
    for (float R : Rs) {
        const RatioTest t = make_ratio_test(R);
        auto check = [&](uint32_t mx, uint32_t mn) {
            const bool want = (float)mx / (float)mn > R;
            if (ratio_exceeds(mx, mn, t) != want) {
                std::printf("fail R=%.9g mx=%u mn=%u want %d\n", (double)R, mx, mn, (int)want);
                return false;
            }
            ++cases;
            return true;
        };
        for (uint32_t mn = 0; mn <= 2100; ++mn)
            for (uint32_t mx = 0; mx <= 2100; ++mx)
                if (!check(mx, mn)) return 1;
        for (int i = 0; i < 2000000; ++i) {
            const uint32_t mn = (uint32_t)(rng() % (1u << 24));
            // quotients near R, where rounding decides
            const double q = (double)R * (1.0 + ((double)(rng() % 2001) - 1000.0) * 1e-9);
            const double mxd = std::floor(q * mn) + (double)(rng() % 3);
            if (mxd < 0 || mxd >= (double)(1u << 24)) continue;
            if (!check((uint32_t)mxd, mn)) return 1;
            if (!check((uint32_t)(rng() % (1u << 24)), mn)) return 1;
        }
    }
    std::printf("ok %llu\n", (unsigned long long)cases);
    return 0;
}
