// Host check of the swirl foot evaluation with line constants
// (face_foot_swirl) against the general face_foot: bitwise equal feet for
// random lines, faces, step sizes and times, in both sweep directions.
// Built with hipcc for the host by tests/test_feet.py.
#include <cstdio>
#include <cstring>
#include <random>

#include "ccfv_feet.h"

using namespace ccfv;

int main()
{
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long bad = 0, total = 0;
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
            }
        }
    }
    if (bad) {
        std::printf("FAILED %ld of %ld\n", bad, total);
        return 1;
    }
    std::printf("ok %ld\n", total);
    return 0;
}
