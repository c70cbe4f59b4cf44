/* Drives the CPU oracle (oracle/ccfv_oracle.c) over every problem kind it
 * restates, for a build with AddressSanitizer and UndefinedBehaviorSanitizer:
 * the reference problem (closed box, both quirks: the ghost and clip index
 * paths), periodic rotation, constant velocity on a non-square grid, the
 * time-dependent swirl, both source kinds, and a Parareal solve.  Prints
 * "ok" when every call returns 0 and the states stay finite. */
#define _DEFAULT_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../oracle/ccfv_oracle.h"

static ccfv_problem make(int nx, int ny, int bc, int vel, double v0, double v1, double v2,
                         double nu, double lo, double hi, unsigned flags)
{
    ccfv_problem p;
    memset(&p, 0, sizeof p);
    p.nx = nx;
    p.ny = ny;
    p.bc = bc;
    p.vel_kind = vel;
    p.x_lo = p.y_lo = lo;
    p.x_hi = p.y_hi = hi;
    p.nu = nu;
    p.vel[0] = v0;
    p.vel[1] = v1;
    p.vel[2] = v2;
    p.flags = flags;
    return p;
}

static int run(const char *name, const ccfv_problem *p, const ccfv_ic *ic, double t1, int nt)
{
    const size_t n = (size_t)p->nx * p->ny;
    double *u = malloc(sizeof(double) * n);
    if (!u) return 1;
    int rc = orc_exact(p, ic, 0.0, u);
    if (!rc) rc = orc_plain(p, u, 0.0, t1, nt);
    for (size_t i = 0; !rc && i < n; i++)
        if (!isfinite(u[i])) rc = 99;
    if (!rc) {
        double incr[4];
        int iters = 0;
        rc = orc_exact(p, ic, 0.0, u);
        if (!rc) rc = orc_parareal(p, u, 0.0, t1, nt, 4, 2, 0.0, incr, &iters, NULL);
        if (!rc && iters != 2) rc = 98;
    }
    free(u);
    if (rc) printf("%s: status %d\n", name, rc);
    return rc;
}

int main(void)
{
    orc_set_threads(1);
    int bad = 0;
    ccfv_ic gauss;
    memset(&gauss, 0, sizeof gauss);
    gauss.kind = CCFV_IC_ROT_GAUSSIAN;
    double gp[6] = {-0.35, 0.0, 0.005, 4.0, 0.0, 0.0};
    memcpy(gauss.p, gp, sizeof gp);

    ccfv_problem ref = make(16, 16, CCFV_BC_REFERENCE, CCFV_VEL_SOLID_BODY, 4.0, 0.0, 0.0,
                            0.01, -1.0, 1.0,
                            CCFV_FLAG_STEP3_POSTREMAP | CCFV_FLAG_DERIV_GHOST_QUIRK);
    bad |= run("reference", &ref, &gauss, 0.25, 8);

    ccfv_ic rot = gauss;
    double rp[6] = {0.5, 0.75, 0.005, 2.0 * M_PI, 0.5, 0.5};
    memcpy(rot.p, rp, sizeof rp);
    ccfv_problem per = make(16, 16, CCFV_BC_PERIODIC, CCFV_VEL_SOLID_BODY, 2.0 * M_PI, 0.5,
                            0.5, 1e-3, 0.0, 1.0, 0);
    bad |= run("rotation", &per, &rot, 0.125, 8);

    ccfv_ic sinsin;
    memset(&sinsin, 0, sizeof sinsin);
    sinsin.kind = CCFV_IC_SINSIN;
    sinsin.p[0] = 1.0;
    sinsin.p[1] = 0.5;
    sinsin.p[2] = 0.01;
    ccfv_problem cst = make(24, 8, CCFV_BC_PERIODIC, CCFV_VEL_CONSTANT, 1.0, 0.5, 0.0, 0.01,
                            0.0, 1.0, 0);
    bad |= run("constant", &cst, &sinsin, 0.25, 8);

    ccfv_ic bell;
    memset(&bell, 0, sizeof bell);
    bell.kind = CCFV_IC_COSINE_BELL;
    bell.p[0] = 0.5;
    bell.p[1] = 0.75;
    bell.p[2] = 0.15;
    ccfv_problem swl = make(16, 16, CCFV_BC_PERIODIC, CCFV_VEL_SWIRL, 1.0, 0.0, 0.0, 1e-4,
                            0.0, 1.0, 0);
    bad |= run("swirl", &swl, &bell, 0.25, 8);

    ccfv_problem sc = cst;
    sc.src_kind = CCFV_SRC_CONSTANT;
    sc.src[0] = 0.5;
    bad |= run("source constant", &sc, &sinsin, 0.25, 8);
    ccfv_problem ss = ref;
    ss.src_kind = CCFV_SRC_SINSIN;
    ss.src[0] = 2.0 * M_PI;
    ss.src[1] = 0.3;
    ss.src[2] = 0.2;
    ss.src[3] = 0.1;
    ss.src[4] = -0.1;
    bad |= run("source sinsin (reference box)", &ss, &gauss, 0.25, 8);

    if (!bad) printf("ok\n");
    return bad ? 1 : 0;
}
