/*
 * ccfv_oracle.h — CPU oracle for the T2S4 CC-CFV step and its Parareal driver.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm in T2S4_Parareal.cu, used by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg as the checker / CPU baseline.  It is
 * never linked into or called by the product library.
 *
 * Problem and initial-condition descriptors are the product's own
 * (include/ccfv.h), so both sides read the same parameters.
 */
#ifndef CCFV_ORACLE_H
#define CCFV_ORACLE_H

#include "../include/ccfv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Threads used by the OpenMP loops over lines (1 = scalar). */
void orc_set_threads(int n);
int  orc_get_threads(void);

/* ---- building blocks (exported so tests can pin each one) -------------- */
/* Parabola antiderivative, T2S4_Parareal.cu:63-65. */
double orc_intc(double c6, double dltc, double cl, double xi);
/* PPM coefficients of cell i of a line (T2S4_Parareal.cu:260-304):
 * out = {CL, CR, DltC, C6}.  periodic selects wrapped stencils. */
void orc_ppm_coef(const double *c, int n, int periodic, int i, double out[4]);
/* Zero-slope cubic ghosts {Cm2, Cm1, CN, CNp1} (T2S4_Parareal.cu:275-286). */
void orc_ghosts(const double *c, int n, double out[4]);
/* Derivative at position xi in cell iL of the quartic matching five cell
 * averages (T2S4_Parareal.cu:440-494); quirk = reference ghost selection. */
double orc_deriv(const double *c, int n, int periodic, int quirk, int iL,
                 double xi, double d);
/* Thomas solve of tridiag(a, b, c) x = r with constant coefficients
 * (T2S4_Parareal.cu:657-686). */
void orc_thomas(double a, double b, double c, const double *r, int n, double *x);
/* Cyclic tridiag(a, b, a) x = r by Sherman-Morrison (periodic mode). */
void orc_cyclic(double a, double b, const double *r, int n, double *x);

/* One line sweep: remap (feet[0..n] of the faces), foot derivatives,
 * RHS, compact operator and the tridiagonal solve.
 *   theta    : 0.25 for x sweeps, 0.5 for y sweeps (the CN weights of
 *              T2S4_Parareal.cu:556-558, :576, :589-591, :608)
 *   h        : the step's dt (dt_step)
 *   eps      : remap clip epsilon (1e-15 for x, 1e-16 for y; reference mode)
 *   deriv_postremap : evaluate the derivative on the remapped line
 * Returns 0, or CCFV_ECFL if feet cross. */
int orc_line_sweep(const double *c, const double *feet, int n, double lo,
                   double hi, int periodic, int quirk, double eps, double nu,
                   double h, double theta, int deriv_postremap, double *out);

/* Feet of the x faces of line j (T2S4_Parareal.cu:203-243): n_x+1 values. */
void orc_feet_x(const ccfv_problem *p, int j, double tau, double t_hi,
                double t_lo, double *feet);
/* Feet of the y faces of line i (T2S4_Parareal.cu:203-258): n_y+1 values. */
void orc_feet_y(const ccfv_problem *p, int i, double tau, double t_hi,
                double t_lo, double *feet);

/* ---- sub-steps and steps ----------------------------------------------- */
/* x sweep of the whole field over h/2 (sub-step 1 or 3), out-of-place. */
int orc_sweep_x(const ccfv_problem *p, const double *cin, double *cout,
                double tn, double h, int substep);
/* y sweep of the whole field over h (sub-step 2), out-of-place. */
int orc_sweep_y(const ccfv_problem *p, const double *cin, double *cout,
                double tn, double h);
/* One Strang step in place (GpuOneStep, T2S4_Parareal.cu:785-968). */
int orc_step(const ccfv_problem *p, double *c, double tn, double h);
/* nt steps of (t1-t0)/nt (T2S4_Parareal.cu:1224-1243). */
int orc_plain(const ccfv_problem *p, double *c, double t0, double t1, int nt);

/* Parareal exactly as the reference schedules it (T2S4_Parareal.cu:1250-1355):
 * k=0 coarse sweep, then per iteration and slice G(U^{k+1}_n), G(U^k_n),
 * F(U^k_n) and the correction, serially.  Stops after kmax iterations or when
 * the relative L2 increment over slice end-points is <= tol (tol <= 0: never).
 * u (in/out): initial state in, U^k(t1) out.  incr[kmax] receives the
 * increments, *iters the iterations run.  slices_out, if not NULL, receives
 * all (n_slices+1) slice states of the last iterate. */
int orc_parareal(const ccfv_problem *p, double *u, double t0, double t1,
                 int nt_fine, int n_slices, int kmax, double tol,
                 double *incr, int *iters, double *slices_out);

/* Cell averages of the initial/exact solution (T2S4_Parareal.cu:118-140). */
int orc_exact(const ccfv_problem *p, const ccfv_ic *ic, double t, double *out);

#ifdef __cplusplus
}
#endif
#endif
