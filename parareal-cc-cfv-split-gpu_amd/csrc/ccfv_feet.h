// ccfv_feet.h — characteristic feet, shared by the host and the gfx950 kernels.
//
// Host and device evaluate the same expressions without fused multiply-add,
// so a foot (and the cell it falls in, a discrete decision) is bit-identical
// on both sides and in the CPU restatement.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/ccfv.h"

namespace ccfv {

struct VelParams {
    int kind;
    double v0, v1, v2;
};

// velocity (T2S4_Parareal.cu:44-45, generalised; see CCFV_VEL_*)
__host__ __device__ inline double vel_x(const VelParams &v, double x, double y,
                                        double t)
{
#pragma clang fp contract(off)
    if (v.kind == CCFV_VEL_SOLID_BODY) return -v.v0 * (y - v.v2);
    if (v.kind == CCFV_VEL_CONSTANT) return v.v0;
    double s = sin(M_PI * x);
    return s * s * sin(2.0 * M_PI * y) * cos(M_PI * t / v.v0);
}

__host__ __device__ inline double vel_y(const VelParams &v, double x, double y,
                                        double t)
{
#pragma clang fp contract(off)
    if (v.kind == CCFV_VEL_SOLID_BODY) return v.v0 * (x - v.v1);
    if (v.kind == CCFV_VEL_CONSTANT) return v.v1;
    double s = sin(M_PI * y);
    return -(s * s) * sin(2.0 * M_PI * x) * cos(M_PI * t / v.v0);
}

// Foot of face q of a line: mean of the Heun feet of its two corners
// (T2S4_Parareal.cu:215-225, :240, :255).  lo, d: along-line box start and
// cell width; o0, o1: the line's corner coordinates in the other direction;
// tau: the sub-step's trace interval (dt/2 for x, dt for y).
template <int DIR>
__host__ __device__ inline double face_foot(const VelParams &v, double lo,
                                            double d, double tau, double t_hi,
                                            double t_lo, int q, double o0,
                                            double o1)
{
#pragma clang fp contract(off)
    double s = lo + d * q;
    double ht = 0.5 * tau;
    if (DIR == 0) {
        double v0 = vel_x(v, s, o0, t_hi);
        double v1 = vel_x(v, s - tau * v0, o0, t_lo);
        double f0 = s - ht * (v0 + v1);
        double w0 = vel_x(v, s, o1, t_hi);
        double w1 = vel_x(v, s - tau * w0, o1, t_lo);
        double f1 = s - ht * (w0 + w1);
        return 0.5 * (f0 + f1);
    } else {
        double v0 = vel_y(v, o0, s, t_hi);
        double v1 = vel_y(v, o0, s - tau * v0, t_lo);
        double f0 = s - ht * (v0 + v1);
        double w0 = vel_y(v, o1, s, t_hi);
        double w1 = vel_y(v, o1, s - tau * w0, t_lo);
        double f1 = s - ht * (w0 + w1);
        return 0.5 * (f0 + f1);
    }
}

// The swirl velocity is separable: along a line only sin^2(pi s) changes;
// sin(2 pi o) at the line's two corner coordinates and cos(pi t / T) at the
// sub-step's two sampling times are line constants.  face_foot_swirl
// evaluates face_foot for CCFV_VEL_SWIRL with them, multiplying in the
// order of vel_x / vel_y, so the foot is bitwise the same.
struct SwirlLine {
    double a0, a1;    // sin(2 pi o0), sin(2 pi o1)
    double chi, clo;  // cos(pi t_hi / T), cos(pi t_lo / T)
};

__host__ __device__ inline SwirlLine swirl_line(const VelParams &v, double o0, double o1,
                                                double t_hi, double t_lo)
{
#pragma clang fp contract(off)
    SwirlLine l;
    l.a0 = sin(2.0 * M_PI * o0);
    l.a1 = sin(2.0 * M_PI * o1);
    l.chi = cos(M_PI * t_hi / v.v0);
    l.clo = cos(M_PI * t_lo / v.v0);
    return l;
}

template <int DIR>
__host__ __device__ inline double face_foot_swirl(const SwirlLine &l, double lo, double d,
                                                  double tau, int q)
{
#pragma clang fp contract(off)
    double s = lo + d * q;
    double ht = 0.5 * tau;
    // sign: u = +sin^2(pi x)..., v = -sin^2(pi y)...
    auto sq = [](double p) { return DIR == 0 ? p * p : -(p * p); };
    const double ps = sq(sin(M_PI * s));
    double v0 = ps * l.a0 * l.chi;
    double v1 = sq(sin(M_PI * (s - tau * v0))) * l.a0 * l.clo;
    double f0 = s - ht * (v0 + v1);
    double w0 = ps * l.a1 * l.chi;
    double w1 = sq(sin(M_PI * (s - tau * w0))) * l.a1 * l.clo;
    double f1 = s - ht * (w0 + w1);
    return 0.5 * (f0 + f1);
}

// sin(pi (s - delta)) from sp = sin(pi s), cp = cos(pi s) and a small shift:
// sin(a - e) = sin a cos e - cos a sin e with e = pi delta, and cos e, sin e
// by their Taylor series, for |e| <= 1/4.  The sums stop at e^12 / 12! and
// e^11 / 11!; the first terms left out, e^14 / 14! and e^13 / 13!, are below
// 5e-20 and 3e-18 at |e| = 1/4.  Explicit fma() keeps host and device
// arithmetic identical.
__host__ __device__ inline double sin_pi_shift(double sp, double cp, double delta)
{
#pragma clang fp contract(off)
    const double e = M_PI * delta, e2 = e * e;
    double c = 1.0 / 479001600.0;           // 1/12!
    c = fma(c, e2, -1.0 / 3628800.0);       // 1/10!
    c = fma(c, e2, 1.0 / 40320.0);          // 1/8!
    c = fma(c, e2, -1.0 / 720.0);           // 1/6!
    c = fma(c, e2, 1.0 / 24.0);             // 1/4!
    c = fma(c, e2, -0.5);                   // 1/2!
    c = fma(c, e2, 1.0);
    double s = -1.0 / 39916800.0;           // 1/11!
    s = fma(s, e2, 1.0 / 362880.0);         // 1/9!
    s = fma(s, e2, -1.0 / 5040.0);          // 1/7!
    s = fma(s, e2, 1.0 / 120.0);            // 1/5!
    s = fma(s, e2, -1.0 / 6.0);             // 1/3!
    s = fma(s * e2, e, e);
    return fma(sp, c, -(cp * s));
}

// face_foot_swirl with sin(pi s), cos(pi s) of the face from a table (sp,
// cp: the values sin(M_PI * s) and cos(M_PI * s) give for s = lo + d q) and,
// when small (pi tau <= 1/4; the swirl speed is at most 1, so the shift
// tau v of a trace is at most tau), the displaced sines by sin_pi_shift:
// one table read instead of a sine per face, and the two traced sines at the
// cost of two short polynomials.  Not bitwise face_foot_swirl: the displaced
// sines differ by the rounding of the two formulas (a few ulp).
template <int DIR>
__host__ __device__ inline double face_foot_swirl_tab(const SwirlLine &l, double lo, double d,
                                                      double tau, int q, double sp, double cp,
                                                      bool small)
{
#pragma clang fp contract(off)
    double s = lo + d * q;
    double ht = 0.5 * tau;
    auto sq = [](double p) { return DIR == 0 ? p * p : -(p * p); };
    auto shifted = [&](double delta) {
        return small ? sin_pi_shift(sp, cp, delta) : sin(M_PI * (s - delta));
    };
    const double ps = sq(sp);
    double v0 = ps * l.a0 * l.chi;
    double v1 = sq(shifted(tau * v0)) * l.a0 * l.clo;
    double f0 = s - ht * (v0 + v1);
    double w0 = ps * l.a1 * l.chi;
    double w1 = sq(shifted(tau * w0)) * l.a1 * l.clo;
    double f1 = s - ht * (w0 + w1);
    return 0.5 * (f0 + f1);
}

// cell of position x and the offset inside it: floor((x - lo)/d)
// (T2S4_Parareal.cu:325-328); n bounds the int conversion
__host__ __device__ inline int cell_of(double lo, double d, int n, double x,
                                       double &xi)
{
#pragma clang fp contract(off)
    double v = floor((x - lo) / d);
    v = fmin(fmax(v, -4.0 * n), 4.0 * n);
    int m = (int)v;
    xi = (x - (lo + d * m)) / d;
    return m;
}

// corner coordinate of line L in the other direction (T2S4_Parareal.cu:197-198)
__host__ __device__ inline double corner(double olo, double od, int L)
{
#pragma clang fp contract(off)
    return olo + od * L;
}

// Per-line filter weights of the uniform-displacement fast path: with every
// foot of a line at offset xi in cell q+S,
//   K_m = I_m(xi) + coef * D_m(xi) = sum_s kappa_s C_{m-2+s}
// and b_i = C_{i+S} + K_{i+S+1} - K_{i+S}.  The PPM antiderivative
// I = CL xi + (dl+c6)/2 xi^2 - c6/3 xi^3 (T2S4_Parareal.cu:63-65) has
// CL = (-1,7,7,-1,0)/12, dl+c6 = (4,-26,30,-10,2)/12, c6 = (3,-18,30,-18,3)/12
// in the five cell averages (:296-300); D is the quartic derivative (:481-486).
__host__ __device__ inline void fir_weights(double xi, double coef, double inv12d,
                                            double kap[5])
{
#pragma clang fp contract(off)
    const double x2 = xi * xi, x3 = x2 * xi;
    const double cl[5] = {-1.0, 7.0, 7.0, -1.0, 0.0};
    const double dc[5] = {4.0, -26.0, 30.0, -10.0, 2.0};
    const double c6[5] = {3.0, -18.0, 30.0, -18.0, 3.0};
    const double g[5] = {2.0 * x3 - 6.0 * x2 + 3.0 * xi + 1.0,
                         -8.0 * x3 + 18.0 * x2 + 6.0 * xi - 15.0,
                         12.0 * x3 - 18.0 * x2 - 24.0 * xi + 15.0,
                         -8.0 * x3 + 6.0 * x2 + 18.0 * xi - 1.0,
                         2.0 * x3 - 3.0 * xi};
    const double ci = coef * inv12d;
    for (int s = 0; s < 5; s++)
        kap[s] = (cl[s] * xi + 0.5 * dc[s] * x2 - (1.0 / 3.0) * c6[s] * x3) *
                     (1.0 / 12.0) +
                 ci * g[s];
}

}  // namespace ccfv
