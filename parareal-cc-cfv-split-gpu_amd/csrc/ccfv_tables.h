// ccfv_tables.h — the per-(direction, step size) solve tables, as functions
// shared by the host (get_table) and the device (the allocation-free entry
// point builds them in its static scratch).  Evaluated without fused
// multiply-add on both sides, so host- and device-built tables are bitwise
// the same.
#pragma once

#include "ccfv_feet.h"
#include "ccfv_internal.h"

namespace ccfv {

struct TabScalars {
    double ca, cb;      // off-diagonal, diagonal of the implicit matrix
    double coef;        // explicit diffusion factor (theta*nu)*(h/d)
    double vl, diag0;   // Sherman-Morrison: v's last entry, first pivot
    double smden;       // 1 + v.z (1 when not periodic)
};

// Thomas factors of the constant tridiagonal matrix of one sweep direction
// and step size (T2S4_Parareal.cu:550-561, :583-594, :666-673); theta is 1/4
// for the x half-steps and 1/2 for the y step.  tab (tab_doubles(n)):
//   [mneg | den / 12 | pneg | z | mf (64) | mb (64)]
// mneg = -a den, den the inverse pivots (with the compact operator's 1/12),
// pneg = -p, z = T^{-1} u of the Sherman-Morrison split when periodic (else
// 0), mf / mb the products of mneg / pneg over each lane's chunk of ch cells
// (1 on idle lanes); for split lines (fast_wpl(n) > 1) then the products
// over each of the 64 wpl lanes of 8 cells (tab_doubles(n) in all).  work:
// 2 n doubles.
__host__ __device__ inline TabScalars thomas_fill(int n, double d, double theta, double nu,
                                                  double h, bool per, int ch, double *tab,
                                                  double *work)
{
#pragma clang fp contract(off)
    TabScalars T;
    const double r2 = h / (d * d);
    T.ca = 1.0 / 12 - theta * nu * r2;
    T.cb = 10.0 / 12 + 2.0 * theta * nu * r2;
    T.coef = theta * nu * (h / d);
    T.vl = 0.0;
    T.smden = 1.0;
    double first = T.cb, last = T.cb;
    if (per) {
        // A = T + u v^T, gamma = -b, u = (gamma,0..0,a), v = (1,0..0,a/gamma)
        const double gamma = -T.cb;
        T.vl = T.ca / gamma;
        first = T.cb - gamma;
        last = T.cb - T.ca * T.vl;
    }
    T.diag0 = first;
    double *p = work, *den = work + n;
    p[0] = T.ca / first;
    den[0] = 1.0 / first;
    for (int i = 1; i < n; i++) {
        const double di = i == n - 1 ? last : T.cb;
        den[i] = 1.0 / (di - p[i - 1] * T.ca);
        p[i] = T.ca * den[i];
    }
    double *mneg = tab, *dv = tab + n, *pneg = tab + 2 * n, *z = tab + 3 * n;
    mneg[0] = 0.0;
    for (int i = 1; i < n; i++) mneg[i] = -T.ca * den[i];
    for (int i = 0; i < n; i++) dv[i] = den[i] / 12.0;
    for (int i = 0; i < n; i++) pneg[i] = -p[i];
    pneg[n - 1] = 0.0;
    if (per) {
        // z = T^{-1} u by the same recurrences (q in z's place, then z)
        z[0] = -T.cb / first;
        for (int i = 1; i < n; i++) {
            const double ui = i == n - 1 ? T.ca : 0.0;
            z[i] = (ui - z[i - 1] * T.ca) * den[i];
        }
        for (int i = n - 2; i >= 0; i--) z[i] = z[i] - p[i] * z[i + 1];
        T.smden = 1.0 + (z[0] + T.vl * z[n - 1]);
    } else {
        for (int i = 0; i < n; i++) z[i] = 0.0;
    }
    // chunk products in the order the lanes fold them; 1 on idle lanes
    const int nl = n / ch;
    for (int l = 0; l < 64; l++) {
        double f = 1.0, b = 1.0;
        if (l < nl) {
            for (int k = 0; k < ch; k++) f = mneg[l * ch + k] * f;
            for (int k = ch - 1; k >= 0; k--) b = pneg[l * ch + k] * b;
        }
        tab[4 * n + l] = f;
        tab[4 * n + 64 + l] = b;
    }
    // split lines: the same over the 64 wpl lanes of 8 cells
    const int ls = fast_wpl(n) > 1 ? 64 * fast_wpl(n) : 0;
    for (int l = 0; l < ls; l++) {
        double f = 1.0, b = 1.0;
        for (int k = 0; k < 8; k++) f = mneg[l * 8 + k] * f;
        for (int k = 7; k >= 0; k--) b = pneg[l * 8 + k] * b;
        tab[4 * n + 128 + l] = f;
        tab[4 * n + 128 + ls + l] = b;
    }
    return T;
}

// Uniform-displacement path, line L of direction dir (0 = x): locate face 0
// of the line exactly as the general path does; the line is accepted when
// its offset xi is >= 1e-9 from a cell edge (then every face's floor()
// agrees with face 0's).  Then b_i = C_m + K_{m+1} - K_m with
// K_m = sum_s kap_s C_{m-2+s} (m = i + S) is one 6-tap filter over
// C_{m-2} .. C_{m+3}: *S the shift, w the weights.
__host__ __device__ inline bool line_fill(const ccfv_problem &P, int dir, double h, double coef,
                                          int L, double *S, double w[6])
{
#pragma clang fp contract(off)
    const int n = dir == 0 ? P.nx : P.ny;
    const double lo = dir == 0 ? P.x_lo : P.y_lo;
    const double olo = dir == 0 ? P.y_lo : P.x_lo;
    const double d = dir == 0 ? (P.x_hi - P.x_lo) / P.nx : (P.y_hi - P.y_lo) / P.ny;
    const double od = dir == 0 ? (P.y_hi - P.y_lo) / P.ny : (P.x_hi - P.x_lo) / P.nx;
    const double tau = dir == 0 ? 0.5 * h : h;
    const double inv12d = 1.0 / (12.0 * d);
    const VelParams v{P.vel_kind, P.vel[0], P.vel[1], P.vel[2]};
    const double o0 = corner(olo, od, L), o1 = corner(olo, od, L + 1);
    const double xb = dir == 0 ? face_foot<0>(v, lo, d, tau, 0.0, 0.0, 0, o0, o1)
                               : face_foot<1>(v, lo, d, tau, 0.0, 0.0, 0, o0, o1);
    double xi;
    const int m = cell_of(lo, d, n, xb, xi);
    if (!(xi > 1e-9 && xi < 1.0 - 1e-9) || m > 2 * n || m < -2 * n) return false;
    *S = (double)m;
    double kap[5];
    fir_weights(xi, coef, inv12d, kap);
    w[0] = -kap[0];
    w[1] = kap[0] - kap[1];
    w[2] = kap[1] - kap[2] + 1.0;
    w[3] = kap[2] - kap[3];
    w[4] = kap[3] - kap[4];
    w[5] = kap[4];
    return true;
}

// whether a problem's sweeps may take the uniform-displacement path
__host__ __device__ inline bool uniform_problem(const ccfv_problem &P)
{
    return P.bc == CCFV_BC_PERIODIC && P.src_kind == CCFV_SRC_NONE &&
           (P.vel_kind == CCFV_VEL_SOLID_BODY || P.vel_kind == CCFV_VEL_CONSTANT);
}

}  // namespace ccfv
