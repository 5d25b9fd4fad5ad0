// Host check of the FP32 prefilter of the neighbor search's box test
// (mega::box_rel, box_limits, box_axis) against the exact FP64 test
// (mega::quotient_below_one): wherever the prefilter decides a pair, its
// answer must equal the exact one.  Cases: bins of several sizes and corners
// far from the origin, windows from 1/60 bin to a full bin, the own point
// within its bin (and drifted up to one bin out), the candidate within the
// 27 bins, and candidates placed at |dx| = w (1 + k 2^-40), at w (1 -+ 1e-4)
// and their neighbors, on every axis.  Prints cases, decided cases (also
// among the random candidates alone) and mismatches.
#include <cmath>
#include <cstdio>
#include <random>
#include "mega_math.hpp"

static long checked = 0, decided = 0, bad = 0, random_checked = 0, random_decided = 0;

static bool check(const double *xi, const double *xj, const double *w, const double *corner, double bin)
{
    const float reach = (float)(4.0 * bin), wmin = (float)(0.02 * bin);
    bool add = true, rej = false;
    bool exact = true;
    for (int k = 0; k < 3; ++k) {
        float lo, hi;
        mega::box_limits(w[k], wmin, lo, hi);
        const int a = mega::box_axis(mega::box_rel(xj[k], corner[k], reach), mega::box_rel(xi[k], corner[k], reach),
                                     lo, hi);
        add = add && a > 0;
        rej = rej || a < 0;
        exact = exact && mega::quotient_below_one(std::fabs(xj[k] - xi[k]), w[k]);
    }
    ++checked;
    if (add || rej) {
        ++decided;
        if (add != exact) {
            if (bad < 10)
                std::printf("mismatch xi=(%a %a %a) xj=(%a %a %a) w=(%a %a %a) add=%d exact=%d\n", xi[0], xi[1], xi[2],
                            xj[0], xj[1], xj[2], w[0], w[1], w[2], (int)add, (int)exact);
            ++bad;
        }
    }
    return add || rej;
}

int main()
{
    std::mt19937_64 rng(2024);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (int n = 0; n < 400000; ++n) {
        const double bin = std::ldexp(1.0 + u(rng), -12 + (int)(u(rng) * 10));   // ~0.25 mm .. 0.5 m scale in m
        double corner[3], xi[3], xj[3], w[3];
        for (int k = 0; k < 3; ++k) {
            corner[k] = (u(rng) - 0.5) * 200.0 * bin * (1 + n % 7);
            xi[k] = corner[k] + (u(rng) * 1.5 - 0.25) * bin * (n % 5 == 0 ? 3.0 : 1.0);
            w[k] = bin * (1.0 / 60.0 + u(rng) * (1.0 - 1.0 / 60.0));
        }
        // random candidates in the 27 bins
        for (int k = 0; k < 3; ++k) xj[k] = corner[k] + (u(rng) * 3.0 - 1.0) * bin;
        ++random_checked;
        random_decided += check(xi, xj, w, corner, bin) ? 1 : 0;
        // candidates at the window's edge on one axis, inside on the others
        const int ax = n % 3;
        for (int k = 0; k < 3; ++k) xj[k] = xi[k] + (u(rng) - 0.5) * w[k];
        const double sgn = (n & 1) ? 1.0 : -1.0;
        const double factors[] = {1.0, 1.0 + 0x1p-40, 1.0 - 0x1p-40, 1.0 + 1e-4, 1.0 - 1e-4, 1.0 + 1.1e-4,
                                  1.0 - 1.1e-4, 1.0 + 5e-5, 1.0 - 5e-5, 1.0 + 2e-4, 1.0 - 2e-4};
        for (double f : factors) {
            double e = xi[ax] + sgn * w[ax] * f;
            for (int s = -2; s <= 2; ++s) {
                double v = e;
                for (int t = 0; t < (s < 0 ? -s : s); ++t) v = std::nextafter(v, s < 0 ? -INFINITY : INFINITY);
                xj[ax] = v;
                check(xi, xj, w, corner, bin);
            }
        }
    }
    std::printf("checked %ld decided %ld random %ld random_decided %ld mismatches %ld\n", checked, decided,
                random_checked, random_decided, bad);
    return bad == 0 ? 0 : 1;
}
