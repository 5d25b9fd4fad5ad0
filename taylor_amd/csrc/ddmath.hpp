// ddmath.hpp -- double-double arithmetic (Dekker's and Knuth's error-free
// transformations: two_sum, two_prod by fused multiply-add) and a power
// function x^y for x > 0 built on it, for host and device.
//
// pow_dd(x, y) evaluates exp(y log x) with about 1e-29 relative error before
// its single final rounding, so it returns the correctly rounded x^y except
// when x^y lies within that distance of a rounding midpoint.  The
// preprocessing of the model (VOL**THRD in PRE_MODEL_SMOOTHING_LENGTHS.F90:61)
// uses it: the device's own pow and the host C library's pow are each
// accurate to about half an ulp to one ulp and round differently near
// midpoints; this one has a precise answer, the correctly rounded power, and
// differs from the host's pow by at most one ulp where that one misrounds.
//   log: x = m 2^k with m in [1/sqrt 2, sqrt 2), log m = 2 atanh(s),
//        s = (m - 1) / (m + 1), the series summed in double-double;
//   exp: y = k ln 2 + r, exp(r) = (Taylor series of r / 2^10)^(2^10).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#define DD_HD __host__ __device__ __forceinline__

namespace dd {

struct Num {
    double hi, lo;
};

DD_HD Num two_sum(double a, double b)
{
#pragma clang fp contract(off)
    const double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}

DD_HD Num quick_two_sum(double a, double b)   // |a| >= |b|
{
#pragma clang fp contract(off)
    const double s = a + b;
    return {s, b - (s - a)};
}

DD_HD Num two_prod(double a, double b)
{
#pragma clang fp contract(off)
    const double p = a * b;
    return {p, fma(a, b, -p)};
}

DD_HD Num add(Num a, Num b)
{
#pragma clang fp contract(off)
    Num s = two_sum(a.hi, b.hi), t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}

DD_HD Num mul(Num a, Num b)
{
#pragma clang fp contract(off)
    Num p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return quick_two_sum(p.hi, p.lo);
}

DD_HD Num neg(Num a) { return {-a.hi, -a.lo}; }

DD_HD Num div(Num a, Num b)
{
#pragma clang fp contract(off)
    const double q1 = a.hi / b.hi;
    Num r = add(a, neg(mul(b, {q1, 0.0})));
    const double q2 = r.hi / b.hi;
    r = add(r, neg(mul(b, {q2, 0.0})));
    const double q3 = r.hi / b.hi;
    return add(quick_two_sum(q1, q2), {q3, 0.0});
}

// ln 2 to 106 bits
constexpr double kLn2Hi = 0x1.62e42fefa39efp-1, kLn2Lo = 0x1.abc9e3b39803fp-56;

// log x for finite x > 0
DD_HD Num log(double x)
{
    int k = 0;
    double m = frexp(x, &k);                     // m in [1/2, 1)
    if (m < 0.70710678118654752440) { m *= 2.0; --k; }
    const double f = m - 1.0;                    // exact (Sterbenz)
    const Num s = div({f, 0.0}, two_sum(2.0, f));
    const Num s2 = mul(s, s);                    // s^2 <= 0.0295
    constexpr int kTerms = 24;                   // 0.0295^24 / 49 < 1e-38
    Num sum = div({1.0, 0.0}, {2.0 * kTerms + 1.0, 0.0});
    for (int n = kTerms - 1; n >= 0; --n) sum = add(mul(sum, s2), div({1.0, 0.0}, {2.0 * n + 1.0, 0.0}));
    const Num lm = mul(mul({2.0, 0.0}, s), sum);
    const Num kl = add(two_prod((double)k, kLn2Hi), two_prod((double)k, kLn2Lo));
    return add(kl, lm);
}

// exp of a double-double argument; the caller keeps it in range
DD_HD Num exp(Num y, int *scale)
{
    const double k = nearbyint(y.hi / kLn2Hi);
    Num r = add(y, neg(add(two_prod(k, kLn2Hi), two_prod(k, kLn2Lo))));   // |r| <= 0.35
    r = {ldexp(r.hi, -10), ldexp(r.lo, -10)};                              // exact scaling
    constexpr int kTerms = 12;                                             // |r| < 3.5e-4
    Num e = {1.0, 0.0};
    for (int n = kTerms; n >= 1; --n) e = add({1.0, 0.0}, div(mul(e, r), {(double)n, 0.0}));
    for (int i = 0; i < 10; ++i) e = mul(e, e);
    *scale = (int)k;
    return e;
}

// x^y for x >= 0 (x = 0 gives 0 for y > 0), rounded once at the end
DD_HD double pow_dd(double x, double y)
{
    if (!(x > 0.0) || !(x <= 0x1.fffffffffffffp+1023) || !(y == y)) return pow(x, y);   // 0, inf, NaN
    const Num l = mul({y, 0.0}, log(x));
    if (!(fabs(l.hi) < 700.0)) return pow(x, y);   // outside the range used here
    int k = 0;
    const Num e = exp(l, &k);
    return ldexp(e.hi + e.lo, k);
}

}  // namespace dd
