// Host check of dd::pow_dd (taylor_amd/csrc/ddmath.hpp).
//   y = 0.5, 2, -1: against sqrt(x), x*x and 1/x, which IEEE arithmetic rounds
//   correctly, so any misrounding of pow_dd shows;
//   y = 0.333333333333333 (THRD of PRE_MODEL_SMOOTHING_LENGTHS.F90): against
//   the C library's pow, over many binades and over the range of nodal volumes.
//   That pow is accurate to 0.52 ulp but not correctly rounded, so the two may
//   differ by one ulp where the power lies near a rounding midpoint; more than
//   one ulp, or more than 0.5 % of the cases, is a failure.
// Prints the counts; exit status 1 on a failure.
#include <cmath>
#include <cstdio>
#include <random>
#include "ddmath.hpp"

int main()
{
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> mant(1.0, 2.0);
    std::uniform_int_distribution<int> ex(-100, 100);
    long n = 0, bad_exact = 0, bad_libm = 0, far_libm = 0;
    auto vs_libm = [&](double x) {
        const double a = dd::pow_dd(x, 0.333333333333333), b = std::pow(x, 0.333333333333333);
        if (a != b) {
            ++bad_libm;
            if (a != std::nextafter(b, a)) ++far_libm;
        }
        ++n;
    };
    for (int i = 0; i < 300000; ++i) {
        const double x = std::ldexp(mant(rng), ex(rng));
        if (dd::pow_dd(x, 0.5) != std::sqrt(x)) ++bad_exact;
        if (dd::pow_dd(x, 2.0) != x * x) ++bad_exact;
        if (dd::pow_dd(x, -1.0) != 1.0 / x) ++bad_exact;
        vs_libm(x);
        n += 3;
    }
    std::uniform_real_distribution<double> vol(0.5, 2.0);
    for (int e = -40; e <= 10; ++e)          // nodal volumes, window volumes
        for (int i = 0; i < 4000; ++i) {
            vs_libm(std::ldexp(vol(rng), e));
        }
    if (dd::pow_dd(0.0, 0.333333333333333) != 0.0) ++bad_exact;
    std::printf("checked %ld exact_mismatches %ld libm_one_ulp %ld libm_beyond_one_ulp %ld\n", n, bad_exact,
                bad_libm, far_libm);
    return (bad_exact || far_libm || bad_libm * 200 > n) ? 1 : 0;
}
