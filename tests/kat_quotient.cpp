// Host check of mega::quotient_below_one against the division it replaces:
// for a >= 0, b > 0 it must equal (a / b < 1.0) bit for bit.  Cases: quotients
// at and around the rounding midpoint 1 - 2^-54 (a = b(1 - 2^-54) +- k ulps,
// built exactly), a = b, a = b/2 and their neighbors, and random pairs over
// many binades.  Prints the number of cases and of mismatches.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include "mega_math.hpp"

static long checked = 0, bad = 0;

static void check(double a, double b)
{
    if (!(a >= 0.0) || !(b > 0.0) || !std::isnormal(b)) return;
    const bool want = (a / b) < 1.0;
    const bool got = mega::quotient_below_one(a, b);
    ++checked;
    if (want != got) {
        if (bad < 10) std::printf("mismatch a=%a b=%a want=%d got=%d\n", a, b, (int)want, (int)got);
        ++bad;
    }
}

int main()
{
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> mant(1.0, 2.0);
    std::uniform_int_distribution<int> ex(-40, 40);
    for (int n = 0; n < 200000; ++n) {
        const double b = std::ldexp(mant(rng), ex(rng));
        // a close to b (1 - 2^-54): step through neighbors of the rounded product
        double a = b * (1.0 - 0x1p-54);
        for (int k = 0; k < 6; ++k) a = std::nextafter(a, 0.0);
        for (int k = 0; k < 13; ++k) { check(a, b); a = std::nextafter(a, INFINITY); }
        double c = b;
        for (int k = 0; k < 4; ++k) c = std::nextafter(c, 0.0);
        for (int k = 0; k < 9; ++k) { check(c, b); c = std::nextafter(c, INFINITY); }
        double h = 0.5 * b;
        for (int k = 0; k < 4; ++k) h = std::nextafter(h, 0.0);
        for (int k = 0; k < 9; ++k) { check(h, b); h = std::nextafter(h, INFINITY); }
        check(b * mant(rng) * 0.5, b);
        check(b * mant(rng) * 0.25, b);
        check(std::ldexp(mant(rng), ex(rng)), b);
        check(0.0, b);
    }
    // b with few significant bits, a next to b - b 2^-54
    for (int e = -30; e <= 30; ++e)
        for (int m = 1; m < 64; ++m) {
            const double b = std::ldexp((double)(64 + m), e);
            const double a = b - std::ldexp(b, -54);
            check(a, b);
            check(std::nextafter(a, 0.0), b);
            check(std::nextafter(a, INFINITY), b);
        }
    std::printf("checked %ld mismatches %ld\n", checked, bad);
    return bad == 0 ? 0 : 1;
}
