// Host check of the portable logarithm's table index (ptx_common.h,
// plog_settle / plog_index): for every m the index equals the one a
// bisection of the table finds, for guesses one below, at and one above it
// (the device's guess comes from the single-precision hardware logarithm),
// around every table boundary and at random points; plog itself equals the
// bisection form bit for bit.  Prints "ok" and exits 0, else the first
// failure and exits 1.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "ptx_common.h"

using namespace ptx;

static const double kTab[32] = PTX_POW2_NEG_J32;

static int bisect(double m) {
    int lo = 0, hi = 31;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (m * kTab[mid] >= 1.0) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// plog with the bisection index: the form the guess-and-settle index replaced
static double plog_bisect(double x) {
    uint64_t b;
    std::memcpy(&b, &x, 8);
    int k = 0;
    if ((b >> 52) == 0) {
        x = x * 0x1p54;
        k = -54;
        std::memcpy(&b, &x, 8);
    }
    k += (int)((b >> 52) & 0x7FF) - 1023;
    const uint64_t mb = (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    std::memcpy(&m, &mb, 8);
    const int lo = bisect(m);
    const double t = m * kTab[lo] - 1.0;
    double p = 1.0 / 13.0;
    p = p * t + -1.0 / 12.0;
    p = p * t + 1.0 / 11.0;
    p = p * t + -1.0 / 10.0;
    p = p * t + 1.0 / 9.0;
    p = p * t + -1.0 / 8.0;
    p = p * t + 1.0 / 7.0;
    p = p * t + -1.0 / 6.0;
    p = p * t + 1.0 / 5.0;
    p = p * t + -1.0 / 4.0;
    p = p * t + 1.0 / 3.0;
    p = p * t + -1.0 / 2.0;
    p = p * t + 1.0;
    p = p * t;
    const double ln2 = 6.93147180559945286227e-01;
    const double kj = (double)k + (double)lo * 0.03125;
    return kj * ln2 + p;
}

static int check(double m) {
    const int j = bisect(m);
    for (int d = -1; d <= 1; ++d) {
        const int got = plog_settle(m, j + d, kTab);
        if (got != j) {
            std::printf("plog_settle(%.17g, %d) = %d, bisection %d\n", m, j + d, got, j);
            return 1;
        }
    }
    if (plog_index(m, kTab) != j) {
        std::printf("plog_index(%.17g) = %d, bisection %d\n", m, plog_index(m, kTab), j);
        return 1;
    }
    const int hf = (int)(32.0f * std::log2((float)m));  // a single-precision guess
    if (hf < j - 1 || hf > j + 1) {
        std::printf("single-precision guess %d for %.17g, index %d\n", hf, m, j);
        return 1;
    }
    return 0;
}

int main() {
    for (int j = 1; j < 32; ++j) {  // m around 1 / tab[j], where the index changes
        double m = 1.0 / kTab[j];
        for (int s = 0; s < 64; ++s) m = std::nextafter(m, 0.0);
        for (int s = 0; s < 128; ++s, m = std::nextafter(m, 4.0))
            if (check(m)) return 1;
    }
    for (double m : {1.0, std::nextafter(2.0, 0.0)})
        if (check(m)) return 1;
    std::mt19937_64 g(5);
    std::uniform_real_distribution<double> U(1.0, 2.0);
    for (int i = 0; i < 2000000; ++i)
        if (check(U(g))) return 1;
    for (uint32_t r : {1u, 2u, 3u, 0x80000000u, 0xFFFFFFFFu, 0x12345678u}) {
        const double x = (double)r * 0x1p-32;
        if (plog(x) != plog_bisect(x)) {
            std::printf("plog(%.17g) differs from the bisection form\n", x);
            return 1;
        }
    }
    for (int i = 0; i < 1000000; ++i) {
        const double x = (double)(uint32_t)g() * 0x1p-32 + 0x1p-40;
        if (plog(x) != plog_bisect(x)) {
            std::printf("plog(%.17g) differs from the bisection form\n", x);
            return 1;
        }
    }
    std::printf("ok\n");
    return 0;
}
