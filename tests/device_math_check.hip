// Host-side checks of the integer and floating-point helpers in
// chllmodel_amd/csrc/chll_device.h (compiled by tests/test_device_math.py).
//   fastdiv   : fdiv(n, make_fastdiv(d)) == n / d
//   philox    : the device Philox4x32-10 against the published known answers
//   recip div : div_by_recip(x, b, RN(1/b)) == x / b (IEEE) on the trial-move
//               domain and on random binary64 operands
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "chll_device.h"

using namespace chll;

int main(int argc, char **argv)
{
    const long ndiv = argc > 1 ? atol(argv[1]) : 20000000;
    long bad = 0;
    std::mt19937_64 rng(12345);
    // fastdiv: every divisor up to 4096 on random and edge numerators, and
    // random divisors up to 2^31
    for (uint32_t d = 1; d <= 4096; ++d) {
        FastDiv f = make_fastdiv(d);
        const uint32_t edges[] = {0u, 1u, d - 1, d, d + 1, 0xFFFFFFFFu, 0xFFFFFFFEu, 0x80000000u, 0x7FFFFFFFu};
        for (uint32_t n : edges) if (fdiv(n, f) != n / d) { if (bad++ < 5) printf("fastdiv %u/%u\n", n, d); }
        for (int t = 0; t < 2000; ++t) {
            uint32_t n = (uint32_t)rng();
            if (fdiv(n, f) != n / d) { if (bad++ < 5) printf("fastdiv %u/%u\n", n, d); }
        }
    }
    for (int t = 0; t < 200000; ++t) {
        uint32_t d = 1 + (uint32_t)(rng() % 0x7FFFFFFFu), n = (uint32_t)rng();
        FastDiv f = make_fastdiv(d);
        if (fdiv(n, f) != n / d) { if (bad++ < 5) printf("fastdiv %u/%u\n", n, d); }
    }
    printf("fastdiv bad=%ld\n", bad);
    // Philox known answers
    struct { U4 c; uint32_t k0, k1; U4 w; } kat[3] = {
        {{0, 0, 0, 0}, 0, 0, {0x6627E8D5u, 0xE169C58Du, 0xBC57AC4Cu, 0x9B00DBD8u}},
        {{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, 0xFFFFFFFFu, 0xFFFFFFFFu,
         {0x408F276Du, 0x41C83B0Eu, 0xA20BC7C6u, 0x6D5451FDu}},
        {{0x243F6A88u, 0x85A308D3u, 0x13198A2Eu, 0x03707344u}, 0xA4093822u, 0x299F31D0u,
         {0xD16CFE09u, 0x94FDCCEBu, 0x5001E420u, 0x24126EA1u}}};
    long pbad = 0;
    for (auto &k : kat) {
        U4 w = philox4x32_10(k.c, k.k0, k.k1);
        if (w.x != k.w.x || w.y != k.w.y || w.z != k.w.z || w.w != k.w.w) ++pbad;
    }
    printf("philox bad=%ld\n", pbad);
    bad += pbad;
    // reciprocal-corrected division
    long dbad = 0;
    std::uniform_real_distribution<double> ux(-1.6, 1.6), ub(1e-3, 1.8);
    for (long t = 0; t < ndiv; ++t) {
        double x, b;
        if (t % 4 == 3) {  // random normal binary64 operands, exponents in [-60, 60]
            x = std::ldexp(1.0 + (double)(rng() >> 12) * 0x1p-52, (int)(rng() % 121) - 60) * ((rng() & 1) ? -1 : 1);
            b = std::ldexp(1.0 + (double)(rng() >> 12) * 0x1p-52, (int)(rng() % 121) - 60);
        } else {
            x = ux(rng);
            b = ub(rng);
        }
        const double y = 1.0 / b;
        const double q = div_by_recip(x, b, y);
        if (q != x / b) { if (dbad++ < 5) printf("div %a / %a: %a vs %a\n", x, b, q, x / b); }
    }
    printf("recipdiv bad=%ld of %ld\n", dbad, ndiv);
    bad += dbad;
    printf(bad ? "FAIL\n" : "OK\n");
    return bad ? 1 : 0;
}
