// Checks the division used by the wave kernel for its uniform divisors:
// for a float divisor 0 < b < 2 and rb = 1/b rounded once in double,
// (float)((double)a * rb) equals the IEEE float quotient a / b for every
// float a.  Modes:
//   exhaustive <b>   every one of the 2^32 float bit patterns a
//   random <n> <m>   n random divisors in (0, 2), m random a each
//   tie              the counter-example for b >= 2 (an exact subnormal tie)
//   pow2 <b>         b a power of two: a * (1/b) in float equals a / b for
//                    every float a (the flow kernels' dx instantiations)
// Prints "mismatches N" and exits non-zero on a mismatch (or, for "tie",
// when the expected mismatch is absent).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

static float fb(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static uint32_t bf(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static bool same(float x, float y) { return (x != x && y != y) || bf(x) == bf(y); }
static float recip_div(float a, double rb) { return (float)((double)a * rb); }

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    if (!std::strcmp(argv[1], "exhaustive")) {
        const float b = std::strtof(argv[2], nullptr);
        const double rb = 1.0 / (double)b;
        const int T = 8;
        std::vector<long> bad(T, 0);
        std::vector<std::thread> th;
        for (int t = 0; t < T; t++)
            th.emplace_back([&, t] {
                for (uint64_t u = t; u < (1ull << 32); u += T) {
                    float a = fb((uint32_t)u);
                    if (!same(recip_div(a, rb), a / b)) bad[t]++;
                }
            });
        for (auto& x : th) x.join();
        long n = 0;
        for (long v : bad) n += v;
        std::printf("b=%a mismatches %ld\n", b, n);
        return n != 0;
    }
    if (!std::strcmp(argv[1], "pow2")) {
        const float b = std::strtof(argv[2], nullptr);
        int ex = 0;
        if (std::frexp(b, &ex) != 0.5f) return 2;
        const float rb = std::ldexp(1.0f, 1 - ex);
        const int T = 8;
        std::vector<long> bad(T, 0);
        std::vector<std::thread> th;
        for (int t = 0; t < T; t++)
            th.emplace_back([&, t] {
                for (uint64_t u = t; u < (1ull << 32); u += T) {
                    float a = fb((uint32_t)u);
                    if (!same(a * rb, a / b)) bad[t]++;
                }
            });
        for (auto& x : th) x.join();
        long n = 0;
        for (long v : bad) n += v;
        std::printf("pow2 b=%a mismatches %ld\n", b, n);
        return n != 0;
    }
    if (!std::strcmp(argv[1], "random")) {
        const long nd = std::atol(argv[2]), na = std::atol(argv[3]);
        std::mt19937_64 rng(12345);
        long n = 0;
        for (long d = 0; d < nd; d++) {
            // any float in (0, 2): random mantissa, exponent 2^-40 .. 2^0
            float b = fb((uint32_t)((rng() & 0x7fffff) | ((uint64_t)(87 + rng() % 41) << 23)));
            const double rb = 1.0 / (double)b;
            for (long k = 0; k < na; k++) {
                float a = fb((uint32_t)rng());
                if (!same(recip_div(a, rb), a / b)) n++;
            }
        }
        std::printf("random mismatches %ld\n", n);
        return n != 0;
    }
    if (!std::strcmp(argv[1], "tie")) {
        // 650000 * 2^-149 / 100000 = 6.5 * 2^-149, halfway between two
        // subnormals: IEEE rounds the tie to even (6 * 2^-149); the product
        // with RN(1/100000) lies just above it and rounds to 7 * 2^-149.
        const float a = fb(650000u), b = 100000.0f;
        const float q = a / b, r = recip_div(a, 1.0 / (double)b);
        std::printf("tie b=1e5: ieee %a recip %a mismatches %d\n", q, r, same(q, r) ? 0 : 1);
        return same(q, r);
    }
    return 2;
}
