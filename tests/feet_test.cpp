// Host check of the swirl foot evaluation with line constants
// (face_foot_swirl) against the general face_foot: bitwise equal feet for
// random lines, faces, step sizes and times, in both sweep directions.  The
// table form (face_foot_swirl_tab: face sines from a table, displaced sines
// by angle addition when pi tau <= 1/4) agrees with it to 4e-16 absolute
// (feet lie in [0, 1]); the angle addition itself (sin_pi_shift) is within
// 5e-16 of sin(pi (s - delta)) in long double over |pi delta| <= 1/4, no
// further than the direct double evaluation.
// Built with hipcc for the host by tests/test_feet.py.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "ccfv_feet.h"

using namespace ccfv;

int main()
{
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long bad = 0, total = 0, smalls = 0;
    double worst = 0.0, worst_shift = 0.0, worst_direct = 0.0;
    // the angle addition alone, over the range it is used on, against
    // sin(pi (s - delta)) in long double: it is no less accurate than the
    // direct double evaluation it replaces
    const long double PI = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < 200000; i++) {
        const double s = U(rng), delta = (2.0 * U(rng) - 1.0) * 0.25 / M_PI;
        const long double ex = sinl(PI * ((long double)s - (long double)delta));
        const double a = sin(M_PI * (s - delta));
        const double b = sin_pi_shift(sin(M_PI * s), cos(M_PI * s), delta);
        worst_direct = std::fmax(worst_direct, (double)fabsl(a - ex));
        worst_shift = std::fmax(worst_shift, (double)fabsl(b - ex));
    }
    if (worst_shift > 5e-16 || worst_shift > worst_direct) {
        std::printf("sin_pi_shift off by %.3g (direct %.3g)\n", worst_shift, worst_direct);
        bad++;
    }
    for (int trial = 0; trial < 2000; trial++) {
        VelParams v{CCFV_VEL_SWIRL, 0.5 + 9.5 * U(rng), 0.0, 0.0};
        const int n = 8 << (trial % 8);
        const double d = 1.0 / n;
        const double h = (0.01 + 2.0 * U(rng)) * d * (1 + trial % 40);
        const double t_lo = 5.0 * U(rng);
        const double t_hi = t_lo + h * U(rng);
        const int L = (int)(U(rng) * n);
        const double o0 = corner(0.0, d, L), o1 = corner(0.0, d, L + 1);
        const SwirlLine sl = swirl_line(v, o0, o1, t_hi, t_lo);
        for (int q = 0; q <= n; q += 1 + n / 64) {
            for (int dir = 0; dir < 2; dir++) {
                const double tau = dir == 0 ? 0.5 * h : h;
                double a = dir == 0 ? face_foot<0>(v, 0.0, d, tau, t_hi, t_lo, q, o0, o1)
                                    : face_foot<1>(v, 0.0, d, tau, t_hi, t_lo, q, o0, o1);
                double b = dir == 0 ? face_foot_swirl<0>(sl, 0.0, d, tau, q)
                                    : face_foot_swirl<1>(sl, 0.0, d, tau, q);
                total++;
                if (std::memcmp(&a, &b, sizeof(double)) != 0) {
                    if (bad < 5) std::printf("differ: n=%d q=%d dir=%d %.17g %.17g\n", n, q,
                                             dir, a, b);
                    bad++;
                }
                const double s = 0.0 + d * q;
                const double sp = sin(M_PI * s), cp = cos(M_PI * s);
                const bool small = M_PI * tau <= 0.25;
                double c = dir == 0 ? face_foot_swirl_tab<0>(sl, 0.0, d, tau, q, sp, cp, small)
                                    : face_foot_swirl_tab<1>(sl, 0.0, d, tau, q, sp, cp, small);
                if (small) smalls++;
                const double e = std::fabs(c - a);
                if (e > worst) worst = e;
                if (e > 4e-16) {
                    if (bad < 5) std::printf("table form: n=%d q=%d dir=%d tau=%g %.17g %.17g\n",
                                             n, q, dir, tau, a, c);
                    bad++;
                }
            }
        }
    }
    if (bad) {
        std::printf("FAILED %ld of %ld\n", bad, total);
        return 1;
    }
    if (smalls < total / 4) {
        std::printf("too few small-step cases: %ld of %ld\n", smalls, total);
        return 1;
    }
    std::printf("ok %ld (table form within %.3g, %ld by angle addition; shift within %.3g)\n",
                total, worst, smalls, worst_shift);
    return 0;
}
