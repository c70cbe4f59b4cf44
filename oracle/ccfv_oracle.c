/*
 * ccfv_oracle.c — CPU oracle: a plain-C restatement of the T2S4 CC-CFV step
 * and Parareal driver of T2S4_Parareal.cu.
 *
 * TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py's
 * cpu_baseline leg).  The product library never links or calls this file.
 *
 * Structure: the reference computes every stage as a separate whole-field
 * kernel (PPM coefficients, remap, foot derivative, RHS, compact operator,
 * Thomas).  Here each sweep is restated line by line — gather one x- or
 * y-line, run the same stages on it, scatter — which is the same arithmetic
 * (every stage of the reference is line-local) and lets OpenMP split the
 * lines across host cores for the CPU baseline.
 *
 * Expression order follows the reference so that results agree to rounding;
 * build with -ffp-contract=off so that no multiply-add is fused.
 *
 * Deliberate deviations from the reference, each documented where it occurs:
 *   - the y-remap clip epsilon 1e-16 (T2S4_Parareal.cu:414-421) can give
 *     jR == Ny, an out-of-bounds read; the cell index is clamped to Ny-1
 *     (effect <= ~1e-14 * |C| at the top boundary);
 *   - x_b_RHS receives Nx for Ny (:859-862); the true line length is used;
 *   - the source term f == 0 (:59-61) is elided, so b = Cbar + diffusion.
 */
#include "ccfv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static int g_threads = 1;

void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------ */
/* velocity (T2S4_Parareal.cu:44-45, generalised)                           */
/* ------------------------------------------------------------------------ */
static double vel_x(const ccfv_problem *p, double x, double y, double t)
{
    switch (p->vel_kind) {
    case CCFV_VEL_SOLID_BODY: /* -omega*(y-cy); reference: -4*y */
        return -p->vel[0] * (y - p->vel[2]);
    case CCFV_VEL_CONSTANT:
        return p->vel[0];
    default: { /* swirl */
        double s = sin(M_PI * x);
        return s * s * sin(2.0 * M_PI * y) * cos(M_PI * t / p->vel[0]);
    }
    }
}

static double vel_y(const ccfv_problem *p, double x, double y, double t)
{
    switch (p->vel_kind) {
    case CCFV_VEL_SOLID_BODY: /* omega*(x-cx); reference: 4*x */
        return p->vel[0] * (x - p->vel[1]);
    case CCFV_VEL_CONSTANT:
        return p->vel[1];
    default: {
        double s = sin(M_PI * y);
        return -(s * s) * sin(2.0 * M_PI * x) * cos(M_PI * t / p->vel[0]);
    }
    }
}

/* Heun backward trace of a corner along x over tau (T2S4_Parareal.cu:215-217):
 *   x* = x - tau V(x,y);  xbar = x - (tau/2) (V(x,y) + V(x*,y)).
 * The reference passes tau = dt/2 for the x sub-steps; its (0.5*dt) and
 * (0.25*dt) factors are tau and tau/2 exactly. */
static double trace_x(const ccfv_problem *p, double x, double y, double tau,
                      double t_hi, double t_lo)
{
    double v0 = vel_x(p, x, y, t_hi);
    double xs = x - tau * v0;
    double v1 = vel_x(p, xs, y, t_lo);
    return x - (0.5 * tau) * (v0 + v1);
}

/* Heun backward trace along y over tau (T2S4_Parareal.cu:219-221). */
static double trace_y(const ccfv_problem *p, double x, double y, double tau,
                      double t_hi, double t_lo)
{
    double v0 = vel_y(p, x, y, t_hi);
    double ys = y - tau * v0;
    double v1 = vel_y(p, x, ys, t_lo);
    return y - (0.5 * tau) * (v0 + v1);
}

/* Face feet of x-line j: mean of the feet of the corners (x_p, y_j) and
 * (x_p, y_{j+1}) (T2S4_Parareal.cu:231-243; corners :197-198). */
void orc_feet_x(const ccfv_problem *p, int j, double tau, double t_hi,
                double t_lo, double *feet)
{
    double dx = (p->x_hi - p->x_lo) / p->nx;
    double dy = (p->y_hi - p->y_lo) / p->ny;
    double y0 = p->y_lo + dy * j;
    double y1 = p->y_lo + dy * (j + 1);
    for (int q = 0; q <= p->nx; q++) {
        double x = p->x_lo + dx * q;
        feet[q] = 0.5 * (trace_x(p, x, y0, tau, t_hi, t_lo) +
                         trace_x(p, x, y1, tau, t_hi, t_lo));
    }
}

/* Face feet of y-line i: mean of the feet of the corners (x_i, y_q) and
 * (x_{i+1}, y_q) (T2S4_Parareal.cu:245-258). */
void orc_feet_y(const ccfv_problem *p, int i, double tau, double t_hi,
                double t_lo, double *feet)
{
    double dx = (p->x_hi - p->x_lo) / p->nx;
    double dy = (p->y_hi - p->y_lo) / p->ny;
    double x0 = p->x_lo + dx * i;
    double x1 = p->x_lo + dx * (i + 1);
    for (int q = 0; q <= p->ny; q++) {
        double y = p->y_lo + dy * q;
        feet[q] = 0.5 * (trace_y(p, x0, y, tau, t_hi, t_lo) +
                         trace_y(p, x1, y, tau, t_hi, t_lo));
    }
}

/* ------------------------------------------------------------------------ */
/* reconstruction                                                            */
/* ------------------------------------------------------------------------ */
double orc_intc(double c6, double dltc, double cl, double xi)
{
    /* T2S4_Parareal.cu:63-65 */
    return -1.0 / 3 * c6 * (xi * xi * xi) + 0.5 * (dltc + c6) * (xi * xi) +
           cl * xi;
}

void orc_ghosts(const double *c, int n, double g[4])
{
    /* T2S4_Parareal.cu:275-286: g = {Cm2, Cm1, CN, CNp1} */
    const double e = 1.0 / 11.0;
    g[2] = e * (-c[n - 3] + 3 * c[n - 2] + 9 * c[n - 1]);
    g[3] = e * (-15 * c[n - 3] + 56 * c[n - 2] - 30 * c[n - 1]);
    g[1] = e * (-c[2] + 3 * c[1] + 9 * c[0]);
    g[0] = e * (-15 * c[2] + 56 * c[1] - 30 * c[0]);
}

/* Extended line value E[k]: ghosts outside [0,n) in the closed box, wrap in
 * periodic mode. */
static inline double ext(const double *c, int n, int periodic,
                         const double g[4], int k)
{
    if (periodic) {
        int m = k % n;
        if (m < 0) m += n;
        return c[m];
    }
    if (k < 0) return k == -1 ? g[1] : g[0];
    if (k >= n) return k == n ? g[2] : g[3];
    return c[k];
}

static void ppm_coef_g(const double *c, int n, int periodic, const double g[4],
                       int i, double out[4])
{
    /* T2S4_Parareal.cu:288-300 */
    const double s12 = 7.0 / 12.0, o12 = 1.0 / 12.0;
    double ci = ext(c, n, periodic, g, i);
    double cip1 = ext(c, n, periodic, g, i + 1);
    double cip2 = ext(c, n, periodic, g, i + 2);
    double cim1 = ext(c, n, periodic, g, i - 1);
    double cim2 = ext(c, n, periodic, g, i - 2);
    double cl = s12 * (ci + cim1) - o12 * (cip1 + cim2);
    double cr = s12 * (cip1 + ci) - o12 * (cip2 + cim1);
    out[0] = cl;
    out[1] = cr;
    out[2] = cr - cl;
    out[3] = 6.0 * (ci - 0.5 * (cl + cr));
}

void orc_ppm_coef(const double *c, int n, int periodic, int i, double out[4])
{
    double g[4] = {0, 0, 0, 0};
    if (!periodic) orc_ghosts(c, n, g);
    ppm_coef_g(c, n, periodic, g, i, out);
}

static double deriv_g(const double *c, int n, int periodic, int quirk,
                      const double g[4], int iL, double xi, double d)
{
    /* T2S4_Parareal.cu:463-486 */
    double ci = ext(c, n, periodic, g, iL);
    double cip1 = ext(c, n, periodic, g, iL + 1);
    double cip2 = ext(c, n, periodic, g, iL + 2);
    double cim1 = ext(c, n, periodic, g, iL - 1);
    double cim2 = ext(c, n, periodic, g, iL - 2);
    if (!periodic && quirk) {
        /* :476-479 — iL==Nx-2 takes CNp1 and iL==1 takes Cm2 */
        if (iL > n - 3) cip2 = g[3];
        if (iL < 2) cim2 = g[0];
    }
    double x2 = xi * xi, x3 = x2 * xi;
    return 1.0 / (12.0 * d) *
           ((2.0 * x3 - 6.0 * x2 + 3.0 * xi + 1.0) * cim2 +
            (-8.0 * x3 + 18.0 * x2 + 6.0 * xi - 15.0) * cim1 +
            (12.0 * x3 - 18.0 * x2 - 24.0 * xi + 15.0) * ci +
            (-8.0 * x3 + 6.0 * x2 + 18.0 * xi - 1.0) * cip1 +
            (2.0 * x3 - 3.0 * xi) * cip2);
}

double orc_deriv(const double *c, int n, int periodic, int quirk, int iL,
                 double xi, double d)
{
    double g[4] = {0, 0, 0, 0};
    if (!periodic) orc_ghosts(c, n, g);
    return deriv_g(c, n, periodic, quirk, g, iL, xi, d);
}

/* ------------------------------------------------------------------------ */
/* tridiagonal solves                                                        */
/* ------------------------------------------------------------------------ */
/* Thomas with a per-row diagonal and constant off-diagonals a (sub) and
 * c (super); the recurrences of T2S4_Parareal.cu:666-683. */
static void thomas_diag(double a, const double *diag, double c, const double *r,
                        int n, double *x, double *p, double *q)
{
    p[0] = c / diag[0];
    q[0] = r[0] / diag[0];
    for (int i = 1; i < n; i++) {
        double den = 1.0 / (diag[i] - p[i - 1] * a);
        p[i] = c * den;
        q[i] = (r[i] - q[i - 1] * a) * den;
    }
    x[n - 1] = q[n - 1];
    for (int i = n - 2; i >= 0; i--) x[i] = q[i] - p[i] * x[i + 1];
}

void orc_thomas(double a, double b, double c, const double *r, int n, double *x)
{
    double *w = (double *)malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < n; i++) w[i] = b;
    thomas_diag(a, w, c, r, n, x, w + n, w + 2 * n);
    free(w);
}

/* Cyclic tridiag(a, b, a): A = T + u v^T with gamma = -b,
 * u = (gamma, 0..0, a), v = (1, 0..0, a/gamma); T has diagonal
 * (b - gamma, b, ..., b, b - a*a/gamma).  x = y - (v.y / (1 + v.z)) z with
 * T y = r, T z = u. */
static void cyclic_ws(double a, double b, const double *r, int n, double *x,
                      double *ws)
{
    double *diag = ws, *p = ws + n, *q = ws + 2 * n, *u = ws + 3 * n,
           *z = ws + 4 * n;
    double gamma = -b;
    double vl = a / gamma;
    for (int i = 0; i < n; i++) { diag[i] = b; u[i] = 0.0; }
    diag[0] = b - gamma;
    diag[n - 1] = b - a * vl;
    u[0] = gamma;
    u[n - 1] = a;
    thomas_diag(a, diag, a, r, n, x, p, q);
    thomas_diag(a, diag, a, u, n, z, p, q);
    double fac = (x[0] + vl * x[n - 1]) / (1.0 + (z[0] + vl * z[n - 1]));
    for (int i = 0; i < n; i++) x[i] = x[i] - fac * z[i];
}

void orc_cyclic(double a, double b, const double *r, int n, double *x)
{
    double *ws = (double *)malloc(sizeof(double) * 5 * n);
    cyclic_ws(a, b, r, n, x, ws);
    free(ws);
}

/* ------------------------------------------------------------------------ */
/* one line                                                                  */
/* ------------------------------------------------------------------------ */
/* workspace doubles for a line of n cells */
#define LINE_WS(n) (15 * ((size_t)(n) + 4))

/* PPM coefficients and the remap of one line over the faces' feet
 * (T2S4_Parareal.cu:260-347, :349-438); cl/dl/c6 n doubles each, fm/fxi
 * the periodic feet cells (n+1).  Returns 0 or CCFV_ECFL. */
static int remap_line(const double *c, const double *feet, int n, double lo,
                      double hi, int periodic, double eps, double *cl,
                      double *dl, double *c6, int *fm, double *fxi,
                      double *cbar)
{
    const double d = (hi - lo) / n;
    double g[4] = {0, 0, 0, 0};
    if (!periodic) orc_ghosts(c, n, g);

    /* PPM coefficients (T2S4_Parareal.cu:260-304) */
    for (int i = 0; i < n; i++) {
        double o[4];
        ppm_coef_g(c, n, periodic, g, i, o);
        cl[i] = o[0];
        dl[i] = o[2];
        c6[i] = o[3];
    }

    /* remap (T2S4_Parareal.cu:306-347, :396-438) */
    if (!periodic) {
        for (int i = 0; i < n; i++) {
            double xbL = feet[i], xbR = feet[i + 1];
            double v = 0.0;
            if (xbR < xbL) return CCFV_ECFL; /* crossed characteristics */
            if (!(xbL > hi || xbR < lo || xbL == xbR)) {
                double ixL = fmax(xbL, lo + eps);
                double ixR = fmin(xbR, hi - eps);
                int iL = (int)floor((ixL - lo) / d);
                int iR = (int)floor((ixR - lo) / d);
                if (iL > n - 1) iL = n - 1; /* clamp: see file header */
                if (iR > n - 1) iR = n - 1;
                double xiL = (ixL - (lo + d * iL)) / d;
                double xiR = (ixR - (lo + d * iR)) / d;
                if (iR == iL) {
                    v = orc_intc(c6[iR], dl[iR], cl[iR], xiR) -
                        orc_intc(c6[iL], dl[iL], cl[iL], xiL);
                } else {
                    double vl = orc_intc(c6[iL], dl[iL], cl[iL], 1.0) -
                                orc_intc(c6[iL], dl[iL], cl[iL], xiL);
                    double vr = orc_intc(c6[iR], dl[iR], cl[iR], xiR);
                    double s = 0.0;
                    for (int m = iL + 1; m <= iR - 1; m++) s += c[m];
                    v = vl + s + vr;
                }
            }
            cbar[i] = v;
        }
    } else {
        /* periodic: unwrapped foot cells; face n is face 0 shifted by n */
        for (int q = 0; q < n; q++) {
            int m = (int)floor((feet[q] - lo) / d);
            fm[q] = m;
            fxi[q] = (feet[q] - (lo + d * m)) / d;
        }
        fm[n] = fm[0] + n;
        fxi[n] = fxi[0];
        for (int i = 0; i < n; i++) {
            int mL = fm[i], mR = fm[i + 1];
            if (mR < mL || (mR == mL && fxi[i + 1] < fxi[i]) || mR - mL > n)
                return CCFV_ECFL;
            int wL = ((mL % n) + n) % n, wR = ((mR % n) + n) % n;
            double v;
            if (mR == mL) {
                v = orc_intc(c6[wR], dl[wR], cl[wR], fxi[i + 1]) -
                    orc_intc(c6[wL], dl[wL], cl[wL], fxi[i]);
            } else {
                double vl = orc_intc(c6[wL], dl[wL], cl[wL], 1.0) -
                            orc_intc(c6[wL], dl[wL], cl[wL], fxi[i]);
                double vr = orc_intc(c6[wR], dl[wR], cl[wR], fxi[i + 1]);
                double s = 0.0;
                for (int m = mL + 1; m <= mR - 1; m++)
                    s += c[((m % n) + n) % n];
                v = vl + s + vr;
            }
            cbar[i] = v;
        }
    }
    return 0;
}

static int line_sweep_ws(const double *c, const double *feet, int n, double lo,
                         double hi, int periodic, int quirk, double eps,
                         double nu, double h, double theta, int deriv_postremap,
                         const double *f0, const double *f1, double *out,
                         double *ws)
{
    const double d = (hi - lo) / n;
    double *cl = ws, *dl = ws + (n + 4), *c6 = ws + 2 * (n + 4);
    double *cbar = ws + 3 * (n + 4), *D = ws + 4 * (n + 4);
    double *b = ws + 5 * (n + 4), *r = ws + 6 * (n + 4);
    int *fm = (int *)(ws + 7 * (n + 4)); /* n+1 ints fit in (n+4) doubles */
    double *fxi = ws + 8 * (n + 4);
    double *tws = ws + 9 * (n + 4);     /* 5 blocks of solver workspace */
    double *fbar = ws + 14 * (n + 4);
    int rc = remap_line(c, feet, n, lo, hi, periodic, eps, cl, dl, c6, fm, fxi, cbar);
    if (rc) return rc;
    /* the source's earlier cell averages remapped over the same feet
     * (T2S4_Parareal.cu:846-853, :893-900, :943-950) */
    if (f0) {
        rc = remap_line(f0, feet, n, lo, hi, periodic, eps, tws, tws + (n + 4),
                        tws + 2 * (n + 4), fm, fxi, fbar);
        if (rc) return rc;
    }

    /* foot derivatives (T2S4_Parareal.cu:440-547) on the pre-remap line, or
     * on the remapped line for the reference's third sub-step (:929-932) */
    const double *src = deriv_postremap ? cbar : c;
    double gs[4] = {0, 0, 0, 0};
    if (!periodic) orc_ghosts(src, n, gs);
    if (!periodic) {
        for (int q = 0; q <= n; q++) {
            double xb = feet[q];
            if (xb >= lo && xb <= hi) {
                double ib = fmin(fmax(xb, lo + 1e-15), hi - 1e-15);
                int iL = (int)floor((ib - lo) / d);
                if (iL > n - 1) iL = n - 1;
                double xi = (ib - (lo + d * iL)) / d;
                D[q] = deriv_g(src, n, 0, quirk, gs, iL, xi, d);
            } else {
                D[q] = 0.0;
            }
        }
    } else {
        for (int q = 0; q < n; q++)
            D[q] = deriv_g(src, n, 1, 0, gs, fm[q], fxi[q], d);
        D[n] = D[0];
    }

    /* RHS (T2S4_Parareal.cu:564-580, :596-613) */
    const double coef = theta * nu * (h / d);
    if (f0) {
        const double w = theta * h;
        for (int i = 0; i < n; i++)
            b[i] = cbar[i] + w * (fbar[i] + f1[i]) + coef * (D[i + 1] - D[i]);
    } else {
        for (int i = 0; i < n; i++) b[i] = cbar[i] + coef * (D[i + 1] - D[i]);
    }

    /* compact operator (T2S4_Parareal.cu:616-654) */
    if (!periodic) {
        r[0] = (10.0 * b[0] + b[1]) / 12.0;
        r[n - 1] = (10.0 * b[n - 1] + b[n - 2]) / 12.0;
        for (int i = 1; i < n - 1; i++)
            r[i] = (b[i + 1] + 10.0 * b[i] + b[i - 1]) / 12.0;
    } else {
        for (int i = 0; i < n; i++) {
            int ip = i + 1 == n ? 0 : i + 1, im = i == 0 ? n - 1 : i - 1;
            r[i] = (b[ip] + 10.0 * b[i] + b[im]) / 12.0;
        }
    }

    /* implicit half (T2S4_Parareal.cu:550-561, :583-594, :657-719) */
    const double r2 = h / (d * d);
    const double ca = 1.0 / 12 - theta * nu * r2;
    const double cb = 10.0 / 12 + 2.0 * theta * nu * r2;
    if (!periodic) {
        for (int i = 0; i < n; i++) tws[i] = cb;
        thomas_diag(ca, tws, ca, r, n, out, tws + n, tws + 2 * n);
    } else {
        cyclic_ws(ca, cb, r, n, out, tws);
    }
    return 0;
}

int orc_line_sweep(const double *c, const double *feet, int n, double lo,
                   double hi, int periodic, int quirk, double eps, double nu,
                   double h, double theta, int deriv_postremap, double *out)
{
    if (n < 4) return CCFV_EINVAL;
    double *ws = (double *)malloc(sizeof(double) * LINE_WS(n));
    int rc = line_sweep_ws(c, feet, n, lo, hi, periodic, quirk, eps, nu, h,
                           theta, deriv_postremap, NULL, NULL, out, ws);
    free(ws);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* source term (T2S4_Parareal.cu:59-61, :143-165)                           */
/* ------------------------------------------------------------------------ */
static double src_point(const ccfv_problem *p, double x, double y, double t)
{
    if (p->src_kind == CCFV_SRC_CONSTANT) return p->src[0];
    const double tp = 2.0 * M_PI, w = p->src[0];
    const double sx = sin(tp * x), sy = sin(tp * y);
    const double cx = cos(tp * x), cy = cos(tp * y);
    const double ct = cos(w * t), st = sin(w * t);
    return (p->src[1] * ct + p->src[2] * st) * sx * sy +
           st * (p->src[3] * cx * sy + p->src[4] * sx * cy);
}

/* cell average of f by the 3x3 Simpson rule, summed as :155-160 */
static double src_avg(const ccfv_problem *p, int i, int j, double t)
{
    const double dx = (p->x_hi - p->x_lo) / p->nx, dy = (p->y_hi - p->y_lo) / p->ny;
    const double x = p->x_lo + (i + 0.5) * dx, y = p->y_lo + (j + 0.5) * dy;
    const double xm = x - 0.5 * dx, xp = x + 0.5 * dx;
    const double ym = y - 0.5 * dy, yp = y + 0.5 * dy;
    return 1.0 / 36.0 *
           (src_point(p, xm, ym, t) + src_point(p, xp, ym, t) + src_point(p, xm, yp, t) +
            src_point(p, xp, yp, t) +
            4.0 * (src_point(p, xm, y, t) + src_point(p, xp, y, t) + src_point(p, x, ym, t) +
                   src_point(p, x, yp, t)) +
            16.0 * src_point(p, x, y, t));
}

/* ------------------------------------------------------------------------ */
/* sweeps over the whole field                                              */
/* ------------------------------------------------------------------------ */
/* The reference's coefficients (:556-558, :589-591) and RHS factors (:576,
 * :608) are theta*K*... with theta = 1/4 for x (two half steps of the CN
 * weight 1/2 over dt/2) and 1/2 for y.  (0.25*K)*r equals theta*nu*r here
 * because theta*nu is formed first in both. */
int orc_sweep_x(const ccfv_problem *p, const double *cin, double *cout,
                double tn, double h, int substep)
{
    const int nx = p->nx, ny = p->ny;
    const int periodic = p->bc == CCFV_BC_PERIODIC;
    const int quirk = (p->flags & CCFV_FLAG_DERIV_GHOST_QUIRK) != 0;
    const int post = substep == 3 && (p->flags & CCFV_FLAG_STEP3_POSTREMAP);
    const double tau = 0.5 * h;
    const double t_lo = substep == 3 ? tn + 0.5 * h : tn;
    const double t_hi = substep == 3 ? tn + h : tn + 0.5 * h;
    /* source times: sub-step 1 f(tn) remapped + f(tn+h/2) (:839-870);
     * sub-step 3 f(tn+h) for both (:935-958) */
    const int src = p->src_kind != CCFV_SRC_NONE;
    const double tf0 = substep == 3 ? tn + h : tn;
    const double tf1 = substep == 3 ? tn + h : tn + 0.5 * h;
    int err = 0;
#pragma omp parallel num_threads(g_threads) reduction(| : err)
    {
        double *line = (double *)malloc(sizeof(double) * (nx + 1));
        double *res = (double *)malloc(sizeof(double) * (nx + 1));
        double *feet = (double *)malloc(sizeof(double) * (nx + 1));
        double *f0 = (double *)malloc(sizeof(double) * (nx + 1));
        double *f1 = (double *)malloc(sizeof(double) * (nx + 1));
        double *ws = (double *)malloc(sizeof(double) * LINE_WS(nx));
#pragma omp for schedule(static)
        for (int j = 0; j < ny; j++) {
            for (int i = 0; i < nx; i++) line[i] = cin[(size_t)i * ny + j];
            if (src)
                for (int i = 0; i < nx; i++) {
                    f0[i] = src_avg(p, i, j, tf0);
                    f1[i] = src_avg(p, i, j, tf1);
                }
            orc_feet_x(p, j, tau, t_hi, t_lo, feet);
            int rc = line_sweep_ws(line, feet, nx, p->x_lo, p->x_hi, periodic,
                                   quirk, 1e-15, p->nu, h, 0.25, post,
                                   src ? f0 : NULL, src ? f1 : NULL, res, ws);
            if (rc) err |= 1;
            for (int i = 0; i < nx; i++) cout[(size_t)i * ny + j] = res[i];
        }
        free(line); free(res); free(feet); free(f0); free(f1); free(ws);
    }
    return err ? CCFV_ECFL : 0;
}

int orc_sweep_y(const ccfv_problem *p, const double *cin, double *cout,
                double tn, double h)
{
    const int nx = p->nx, ny = p->ny;
    const int periodic = p->bc == CCFV_BC_PERIODIC;
    const int quirk = (p->flags & CCFV_FLAG_DERIV_GHOST_QUIRK) != 0;
    /* source: f(tn) remapped + f(tn+h) (:873-917) */
    const int src = p->src_kind != CCFV_SRC_NONE;
    int err = 0;
#pragma omp parallel num_threads(g_threads) reduction(| : err)
    {
        double *feet = (double *)malloc(sizeof(double) * (ny + 1));
        double *f0 = (double *)malloc(sizeof(double) * (ny + 1));
        double *f1 = (double *)malloc(sizeof(double) * (ny + 1));
        double *ws = (double *)malloc(sizeof(double) * LINE_WS(ny));
#pragma omp for schedule(static)
        for (int i = 0; i < nx; i++) {
            if (src)
                for (int j = 0; j < ny; j++) {
                    f0[j] = src_avg(p, i, j, tn);
                    f1[j] = src_avg(p, i, j, tn + h);
                }
            orc_feet_y(p, i, h, tn + h, tn, feet);
            int rc = line_sweep_ws(cin + (size_t)i * ny, feet, ny, p->y_lo,
                                   p->y_hi, periodic, quirk, 1e-16, p->nu, h,
                                   0.5, 0, src ? f0 : NULL, src ? f1 : NULL,
                                   cout + (size_t)i * ny, ws);
            if (rc) err |= 1;
        }
        free(feet); free(f0); free(f1); free(ws);
    }
    return err ? CCFV_ECFL : 0;
}

int orc_step(const ccfv_problem *p, double *c, double tn, double h)
{
    /* GpuOneStep (T2S4_Parareal.cu:785-968): x(h/2) -> y(h) -> x(h/2) */
    size_t n = (size_t)p->nx * p->ny;
    double *a = (double *)malloc(sizeof(double) * n);
    int rc = orc_sweep_x(p, c, a, tn, h, 1);
    if (!rc) rc = orc_sweep_y(p, a, c, tn, h);
    if (!rc) rc = orc_sweep_x(p, c, a, tn, h, 3);
    if (!rc) memcpy(c, a, sizeof(double) * n);
    free(a);
    return rc;
}

int orc_plain(const ccfv_problem *p, double *c, double t0, double t1, int nt)
{
    /* T2S4_Parareal.cu:1060, :1234-1236 */
    double dt = (t1 - t0) / nt;
    for (int n = 0; n < nt; n++) {
        int rc = orc_step(p, c, t0 + n * dt, dt);
        if (rc) return rc;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Parareal, scheduled as the reference schedules it                        */
/* ------------------------------------------------------------------------ */
int orc_parareal(const ccfv_problem *p, double *u, double t0, double t1,
                 int nt_fine, int n_slices, int kmax, double tol, double *incr,
                 int *iters, double *slices_out)
{
    if (n_slices < 1 || nt_fine % n_slices) return CCFV_EINVAL;
    const size_t n = (size_t)p->nx * p->ny;
    const double DT = (t1 - t0) / n_slices;      /* :1251 */
    const int J = nt_fine / n_slices;            /* :1252 */
    const double dt = DT / (double)J;            /* :979 */
    double *U = (double *)malloc(sizeof(double) * n * (n_slices + 1));
    double *U1 = (double *)malloc(sizeof(double) * n * (n_slices + 1));
    double *g1 = (double *)malloc(sizeof(double) * n);
    double *g0 = (double *)malloc(sizeof(double) * n);
    double *f = (double *)malloc(sizeof(double) * n);
    int rc = 0, k = 0;
    memcpy(U, u, sizeof(double) * n);
    /* k = 0 coarse sweep (:1271-1280; the repeat at :1282-1291 recomputes
     * the same values) */
    for (int s = 0; s < n_slices && !rc; s++) {
        memcpy(U + (s + 1) * n, U + s * n, sizeof(double) * n);
        rc = orc_step(p, U + (s + 1) * n, t0 + s * DT, DT);
    }
    for (k = 0; k < kmax && !rc; k++) {
        memcpy(U1, U, sizeof(double) * n); /* :1299 */
        for (int s = 0; s < n_slices && !rc; s++) {
            double tn = t0 + s * DT;
            memcpy(g1, U1 + s * n, sizeof(double) * n);
            rc = orc_step(p, g1, tn, DT);                 /* :1306-1310 */
            memcpy(g0, U + s * n, sizeof(double) * n);
            if (!rc) rc = orc_step(p, g0, tn, DT);        /* :1313-1317 */
            memcpy(f, U + s * n, sizeof(double) * n);
            for (int jj = 0; jj < J && !rc; jj++)         /* :977-984 */
                rc = orc_step(p, f, tn + jj * dt, dt);
            double *dst = U1 + (s + 1) * n;
            for (size_t c = 0; c < n; c++) dst[c] = g1[c] + (f[c] - g0[c]);
        }
        double num = 0.0, den = 0.0;
        for (size_t c = n; c < n * (n_slices + 1); c++) {
            double dd = U1[c] - U[c];
            num += dd * dd;
            den += U1[c] * U1[c];
        }
        double inc = den > 0 ? sqrt(num / den) : sqrt(num);
        if (incr) incr[k] = inc;
        double *tmp = U; U = U1; U1 = tmp; /* :1352-1354 */
        if (tol > 0 && inc <= tol) { k++; break; }
    }
    if (iters) *iters = k;
    if (!rc) {
        memcpy(u, U + (size_t)n_slices * n, sizeof(double) * n);
        if (slices_out)
            memcpy(slices_out, U, sizeof(double) * n * (n_slices + 1));
    }
    free(U); free(U1); free(g1); free(g0); free(f);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* initial / exact cell averages                                            */
/* ------------------------------------------------------------------------ */
static double rot_gauss(const double *q, double x, double y, double t)
{
    /* C_Exact (T2S4_Parareal.cu:51-56) about centre (cx, cy) */
    double x0 = q[0], y0 = q[1], s = q[2], om = q[3], cx = q[4], cy = q[5],
           nu = q[6];
    double co = cos(om * t), si = sin(om * t);
    double ax = x - cx, ay = y - cy;
    double xs = cx + (ax * co + ay * si);
    double ys = cy + (-ax * si + ay * co);
    double den = s + 4 * nu * t;
    double ex = xs - x0, ey = ys - y0;
    return s / den * exp(-(ex * ex + ey * ey) / den);
}

static double gaussians(const ccfv_ic *ic, double x, double y, double t)
{
    double u = ic->p[0], v = ic->p[1], nu = ic->p[2], acc = 0.0;
    for (int k = 0; k < ic->n && k < 3; k++) {
        const double *g = ic->p + 3 + 4 * k;
        double w2 = g[2] * g[2], s2 = w2 + 2.0 * nu * t;
        double ex = x - g[0] - u * t, ey = y - g[1] - v * t;
        acc += g[3] * (w2 / s2) * exp(-(ex * ex + ey * ey) / (2.0 * s2));
    }
    return acc;
}

static double cosine_bell(const ccfv_ic *ic, double x, double y)
{
    double ex = x - ic->p[0], ey = y - ic->p[1];
    double r = sqrt(ex * ex + ey * ey), R = ic->p[2];
    return r < R ? 0.5 * (1.0 + cos(M_PI * r / R)) : 0.0;
}

/* point value; periodic problems sum the 3x3 periodic images */
static double point_value(const ccfv_problem *p, const ccfv_ic *ic,
                          const double *q, double x, double y, double t)
{
    int img = p->bc == CCFV_BC_PERIODIC ? 1 : 0;
    double Lx = p->x_hi - p->x_lo, Ly = p->y_hi - p->y_lo, acc = 0.0;
    for (int a = -img; a <= img; a++)
        for (int b = -img; b <= img; b++) {
            double xx = x + a * Lx, yy = y + b * Ly;
            double v;
            if (ic->kind == CCFV_IC_ROT_GAUSSIAN) v = rot_gauss(q, xx, yy, t);
            else if (ic->kind == CCFV_IC_GAUSSIANS) v = gaussians(ic, xx, yy, t);
            else v = cosine_bell(ic, xx, yy);
            acc += v;
        }
    return acc;
}

int orc_exact(const ccfv_problem *p, const ccfv_ic *ic, double t, double *out)
{
    const int nx = p->nx, ny = p->ny;
    const double dx = (p->x_hi - p->x_lo) / nx, dy = (p->y_hi - p->y_lo) / ny;
    double q[7] = {0};
    if (ic->kind == CCFV_IC_ROT_GAUSSIAN) {
        for (int k = 0; k < 6; k++) q[k] = ic->p[k];
        q[6] = p->nu;
    }
    if (ic->kind < 0 || ic->kind > CCFV_IC_COSINE_BELL) return CCFV_EINVAL;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i = 0; i < nx; i++) {
        for (int j = 0; j < ny; j++) {
            /* cell centre (T2S4_Parareal.cu:179-180) and half points
             * (:126-129) */
            double x = p->x_lo + (i + 0.5) * dx, y = p->y_lo + (j + 0.5) * dy;
            double xm = x - 0.5 * dx, xp = x + 0.5 * dx;
            double ym = y - 0.5 * dy, yp = y + 0.5 * dy;
            double v;
            if (ic->kind == CCFV_IC_SINSIN) {
                const double tp = 2.0 * M_PI;
                double u = ic->p[0], w = ic->p[1], nu = ic->p[2];
                double ax = (cos(tp * (xm - u * t)) - cos(tp * (xp - u * t))) /
                            (tp * dx);
                double ay = (cos(tp * (ym - w * t)) - cos(tp * (yp - w * t))) /
                            (tp * dy);
                v = ax * ay * exp(-2.0 * tp * tp * nu * t);
            } else {
#define PV(a, b) point_value(p, ic, q, a, b, t)
                /* 3x3 tensor Simpson (T2S4_Parareal.cu:131-136) */
                v = 1.0 / 36.0 *
                    (PV(xm, ym) + PV(xp, ym) + PV(xm, yp) + PV(xp, yp) +
                     4.0 * (PV(xm, y) + PV(xp, y) + PV(x, ym) + PV(x, yp)) +
                     16.0 * PV(x, y));
#undef PV
            }
            out[(size_t)i * ny + j] = v;
        }
    }
    return 0;
}
