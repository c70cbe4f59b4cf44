/*
 * ccfv.h — C ABI of the MI355X (gfx950) T2S4 CC-CFV solver and its Parareal driver.
 *
 * Everything here is plain C: doubles, fixed-width integers, sizes and opaque
 * handles.  No HIP or torch types cross the boundary.  All state is FP64.
 *
 * The reference program (T2S4_Parareal.cu) has no library surface: its solver
 * lives behind four internal C++ entry points, listed next to their
 * replacements below:
 *
 *   GpuOneStep(GPUContext&, FLOAT tn, FLOAT dt)          T2S4_Parareal.cu:785   -> ccfv_step
 *   CoarsePropagator(GPUContext&, FLOAT tn, FLOAT DT)    T2S4_Parareal.cu:971   -> ccfv_coarse
 *   FinePropagator(GPUContext&, FLOAT tn, FLOAT DT, J)   T2S4_Parareal.cu:977   -> ccfv_fine
 *   parareal_correction_kernel(Gk1, Gk, F, U, N)         T2S4_Parareal.cu:771   -> ccfv_parareal_correction
 *
 * plus the two hot loops of main():
 *
 *   serial fine loop                                     T2S4_Parareal.cu:1234  -> ccfv_plain_solve
 *   Parareal init + iterations                           T2S4_Parareal.cu:1250  -> ccfv_parareal_solve
 *
 * and the context set-up of main() (config keys N, K, T_min, T_max, Nt at
 * T2S4_Parareal.cu:1031-1049, buffers at :1088-1135, GPUContext at :724).
 *
 * State layout in memory is the reference's: C[i*ny + j] with i the x index
 * (slow) and j the y index (contiguous), T2S4_Parareal.cu:176-180.
 *
 * Every function returns an int status (CCFV_OK == 0, negative on error) and
 * never aborts; ccfv_last_error() returns a thread-local message describing
 * the most recent failure on the calling thread.
 */
#ifndef CCFV_H
#define CCFV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CCFV_OK          0
#define CCFV_EINVAL     (-1)   /* bad argument or shape                         */
#define CCFV_ENOMEM     (-2)   /* device or host allocation failed              */
#define CCFV_EDEVICE    (-3)   /* HIP runtime error / no device                 */
#define CCFV_ENOTCONV   (-4)   /* Parareal stopped at kmax above tolerance      */
#define CCFV_ECFL       (-5)   /* characteristic feet crossed (time step too large) */
#define CCFV_EBUSY      (-6)   /* serving: request refused by the memory budget */
#define CCFV_ENONFINITE (-7)   /* NaN/Inf detected in a state                   */

/* ---- boundary mode ----------------------------------------------------- */
/* CCFV_BC_REFERENCE: the reference's closed box.  Feet are clipped to the box
 * and cells whose feet leave it get zero inflow (T2S4_Parareal.cu:322-324),
 * ghost cells use the zero-slope cubic closure (:275-286), the compact RHS
 * uses zero ghosts (:624-631) and the tridiagonal system is truncated
 * (:666-683).
 * CCFV_BC_PERIODIC: feet wrap, every stencil wraps, the compact RHS and the
 * tridiagonal system are cyclic (solved by Sherman-Morrison).  The reference
 * has no periodic mode; this is the same scheme on a torus. */
#define CCFV_BC_REFERENCE 0
#define CCFV_BC_PERIODIC  1

/* ---- velocity fields ---------------------------------------------------- */
/* SOLID_BODY: u = -omega*(y-cy), v = omega*(x-cx); vel = {omega, cx, cy}.
 *             omega=4, c=(0,0) is the reference's Vx=-4y, Vy=4x (:44-45).
 * CONSTANT:   u = vel[0], v = vel[1].
 * SWIRL:      u =  sin^2(pi x) sin(2 pi y) cos(pi t / T),
 *             v = -sin^2(pi y) sin(2 pi x) cos(pi t / T);   vel = {T}.      */
#define CCFV_VEL_SOLID_BODY 0
#define CCFV_VEL_CONSTANT   1
#define CCFV_VEL_SWIRL      2

/* ---- scheme flags -------------------------------------------------------- */
/* Reproduce the third sub-step's explicit diffusion derivative evaluated on
 * the remapped field (T2S4_Parareal.cu:929-932) instead of the pre-remap
 * field used by the first two sub-steps (:834-837, :882-885). */
#define CCFV_FLAG_STEP3_POSTREMAP  0x1u
/* Reproduce the reference's ghost selection in the foot-derivative stencil
 * (iL==1 takes Cm2, iL==Nx-2 takes CNp1; :476-479, :530-533). */
#define CCFV_FLAG_DERIV_GHOST_QUIRK 0x2u
/* Both reference-mode quirks. */
#define CCFV_FLAGS_REFERENCE (CCFV_FLAG_STEP3_POSTREMAP | CCFV_FLAG_DERIV_GHOST_QUIRK)

typedef struct ccfv_problem {
    int32_t  nx, ny;        /* cells in x and y (reference: N, with nx == ny)  */
    int32_t  bc;            /* CCFV_BC_*                                        */
    int32_t  vel_kind;      /* CCFV_VEL_*                                       */
    double   x_lo, x_hi;    /* box (reference: X_min, X_max = -1, 1; :30-31)   */
    double   y_lo, y_hi;    /* box (reference: Y_min, Y_max = -1, 1; :32-33)   */
    double   nu;            /* diffusion coefficient (reference config key K)  */
    double   vel[4];        /* velocity parameters, see CCFV_VEL_*             */
    uint32_t flags;         /* CCFV_FLAG_*                                      */
    uint32_t reserved;
    int32_t  src_kind;      /* CCFV_SRC_* (reference: f == 0, :59-61)          */
    uint32_t src_reserved;
    double   src[5];        /* source parameters, see CCFV_SRC_*               */
} ccfv_problem;

/* ---- source term f(x, y, t) ---------------------------------------------
 * Every sub-step adds theta*h*(fbar0 + f1) to its right-hand side, as the
 * reference does (:564-613, :839-958): f1 the cell averages of f (3x3
 * Simpson, :143-165) at the sub-step's later time, fbar0 those at its
 * earlier time remapped over the sub-step's own feet like the solution;
 * x sub-step 1: times tn, tn+h/2; y: tn, tn+h; x sub-step 3: tn+h for both;
 * theta = 1/4 for x and 1/2 for y.  Over a step this adds 2*h*f for a
 * constant f, twice the source of the equation: a solver for
 * u_t + V.grad u = nu lap u + g takes f = g/2.
 * NONE:     f = 0 (the reference).
 * CONSTANT: f = src[0].
 * SINSIN:   f = (src[1] cos(w t) + src[2] sin(w t)) S(x) S(y)
 *             + sin(w t) (src[3] C(x) S(y) + src[4] S(x) C(y)),
 *           w = src[0], S(z) = sin(2 pi z), C(z) = cos(2 pi z).           */
#define CCFV_SRC_NONE     0
#define CCFV_SRC_CONSTANT 1
#define CCFV_SRC_SINSIN   2

/* ---- initial / exact data ----------------------------------------------- */
/* ROT_GAUSSIAN: the reference's C_Exact (:51-56) generalised to a centre of
 *   rotation: s/(s+4 nu t) exp(-|x* - x0|^2/(s + 4 nu t)) with x* the point
 *   rotated back by omega*t about (cx, cy).  p = {x0, y0, s, omega, cx, cy}.
 *   Reference IC: {-0.35, 0, 0.005, 4, 0, 0}.  Cell averages by the 3x3
 *   tensor Simpson rule of :131-136.
 * GAUSSIANS: sum of n <= 3 Gaussians a exp(-r^2/(2 w^2)) translated by a
 *   constant velocity (u, v) and diffused (exact on the plane):
 *   p = {u, v, nu, (x0, y0, w, a) * n}.  Simpson cell averages.
 * SINSIN: sin(2 pi (x - u t)) sin(2 pi (y - v t)) exp(-8 pi^2 nu t), exact
 *   cell averages (closed form).  p = {u, v, nu}.
 * COSINE_BELL: 0.5 (1 + cos(pi r / R)) for r < R, else 0; p = {x0, y0, R}.
 *   Simpson cell averages (t is ignored).                                    */
#define CCFV_IC_ROT_GAUSSIAN 0
#define CCFV_IC_GAUSSIANS    1
#define CCFV_IC_SINSIN       2
#define CCFV_IC_COSINE_BELL  3

typedef struct ccfv_ic {
    int32_t kind;
    int32_t n;              /* number of Gaussians for CCFV_IC_GAUSSIANS */
    double  p[16];
} ccfv_ic;

/* ---- Parareal report ---------------------------------------------------- */
#define CCFV_MAX_ITERS 64
typedef struct ccfv_parareal_report {
    int32_t iterations;          /* iterations performed (k)                   */
    int32_t converged;           /* 1 if the increment fell to <= tol, or all  */
                                 /* n_slices iterations ran (every slice exact) */
    int32_t n_slices, j_fine;    /* slices and fine steps per slice            */
    double  increment[CCFV_MAX_ITERS]; /* rel. L2 increment per iteration over
                                        * all slice end-points                */
    double  fine_steps_run;      /* single-field fine steps executed           */
    double  coarse_steps_run;    /* single-field coarse steps executed         */
    double  wall_ms;             /* hipEvent time of the whole solve           */
    int32_t serial_finish;       /* 1: kmax was reached above tol and the result
                                  * is the serial fine solve from the initial
                                  * state (CCFV_OPT_SERIAL_FINISH); converged
                                  * is then 1                                  */
    int32_t reserved;
} ccfv_parareal_report;

typedef struct ccfv_ctx ccfv_ctx;

/* Thread-local text of the last error on this thread ("" if none). */
const char *ccfv_last_error(void);

/* Library and device information. */
int ccfv_version(int32_t *major, int32_t *minor);
int ccfv_device_count(int32_t *count);

/* Fill *p with the reference problem: box [-1,1]^2, Vx=-4y, Vy=4x, closed-box
 * boundaries, both quirks on, diffusion nu = K (T2S4_Parareal.cu:27-61,
 * :1037-1048).  n = config key N. */
int ccfv_problem_reference(ccfv_problem *p, int32_t n, double K);
/* Fill *ic with the reference initial condition (C_Exact at t=0, :1147). */
int ccfv_ic_reference(ccfv_ic *ic, double K);

/* Context: owns device memory for one state and the scratch the step needs
 * (replaces the buffers of :1088-1135 and GPUContext of :724-768).  The
 * device is the HIP device current on the calling thread. */
int  ccfv_create(const ccfv_problem *p, ccfv_ctx **out);
void ccfv_destroy(ccfv_ctx *ctx);

/* Copy the state in / out (host arrays of nx*ny doubles, reference layout). */
int ccfv_set_state(ccfv_ctx *ctx, const double *host);
int ccfv_get_state(ccfv_ctx *ctx, double *host);
/* Device-side copies for callers that keep fields in HBM (pointers to
 * nx*ny doubles on the context's device). */
int ccfv_set_state_device(ccfv_ctx *ctx, const double *dev);
int ccfv_get_state_device(ccfv_ctx *ctx, double *dev);
/* Device address of the context's state (nx*ny doubles, reference layout),
 * for callers that exchange fields in HBM (e.g. wrap it as a tensor).  Valid
 * until ccfv_destroy. */
int ccfv_state_device_ptr(ccfv_ctx *ctx, void **dev);
/* Cell averages of an initial/exact solution at time t into the state
 * (replaces get_exact_solution_on_gpu, :118-140).  Computed on the device. */
int ccfv_fill_exact(ccfv_ctx *ctx, const ccfv_ic *ic, double t);
/* Cell averages into a host array, computed on the device (state untouched). */
int ccfv_exact_to_host(ccfv_ctx *ctx, const ccfv_ic *ic, double t, double *host);

/* One Strang-split step x(h/2) -> y(h) -> x(h/2) in place (GpuOneStep, :785). */
int ccfv_step(ccfv_ctx *ctx, double tn, double h);
/* Coarse propagator: one step of DT (CoarsePropagator, :971). */
int ccfv_coarse(ccfv_ctx *ctx, double tn, double DT);
/* Fine propagator: J steps of DT/J (FinePropagator, :977). */
int ccfv_fine(ccfv_ctx *ctx, double tn, double DT, int32_t J);
/* Plain serial-in-time solve: nt steps of (t1-t0)/nt from t0 (:1224-1243). */
int ccfv_plain_solve(ccfv_ctx *ctx, double t0, double t1, int32_t nt);
/* The plain serial-in-time solve of ccfv_plain_solve (nt steps of
 * (t1-t0)/nt from t0, in place) on a caller's device field u of nx*ny
 * doubles, queued on the caller's stream (NULL: the default stream), that
 * allocates no device memory.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under stream capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Everything the sweeps need
 * beyond u (the crossed-feet flag, both directions' solve tables, the
 * uniform-offset sweep's per-line tables, the swirl's face table) lives in
 * one static 512 KiB device array per GPU, built there by kernels: grids up
 * to 2048 x 2048 fit, larger ones are refused with CCFV_EINVAL.  Calls that
 * overlap on one GPU race on that array.  Results are bitwise those of
 * ccfv_plain_solve.  Blocks until the solve is done (it reads the flag);
 * returns CCFV_ECFL when feet crossed, as ccfv_plain_solve. */
int ccfv_plain_solve_noalloc(const ccfv_problem *p, double *u, double t0, double t1,
                             int32_t nt, void *stream);

/* Steps n = n_begin .. n_end-1 of the serial fine loop, step n from
 * t_origin + n*dt with step size dt, as the reference's loop computes them
 * (:1234-1236); lets a caller stop every T_span steps for output. */
int ccfv_advance(ccfv_ctx *ctx, double t_origin, double dt, int32_t n_begin,
                 int32_t n_end);

/* Parareal over [t0, t1]: n_slices coarse steps of DT=(t1-t0)/n_slices, fine
 * propagator J = nt_fine/n_slices steps of DT/J; iterations
 *   U_{n+1}^{k+1} = G(U_n^{k+1}) + (F(U_n^k) - G(U_n^k))
 * (:1250-1355).  All fine propagators of one iteration run as one batched
 * launch per sweep; the correction is fused into the coarse step.  Stops after
 * kmax iterations or once the relative L2 increment over all slice end-points
 * is <= tol (tol <= 0: run all kmax iterations, as the reference does).
 * After n_slices iterations every slice end-point is the serial fine solution
 * and further iterations change nothing, so min(kmax, n_slices) run.
 * The initial state is the context state; on return the state holds U^k(t1).
 * Returns CCFV_OK, or CCFV_ENOTCONV when tol > 0 was not reached (the state
 * still holds the last iterate).  rep may be NULL. */
int ccfv_parareal_solve(ccfv_ctx *ctx, double t0, double t1, int32_t nt_fine,
                        int32_t n_slices, int32_t kmax, double tol,
                        ccfv_parareal_report *rep);

/* Parareal schedule options (no counterpart in the reference, which runs
 * every slice serially, :1302-1330):
 *   CCFV_OPT_PIPELINE   1 (default): the fine propagators of iteration k+1
 *                       run on a second stream and overlap the coarse chain
 *                       of iteration k group by group; 2: in addition the
 *                       coarse chains of consecutive iterations run on two
 *                       streams and overlap (faster when the chain dominates:
 *                       many slices of few fine steps); 0: one stream, fine
 *                       batch then chain.  All give identical results.
 *   CCFV_OPT_PIPE_GROUP slices per group of the pipelined schedule (32).
 *   CCFV_OPT_GRAPHS     1: each pipelined phase (coarse chain of iteration k
 *                       with the fine propagators of k+1) is captured once
 *                       as a hipGraph and replayed by later solves with the
 *                       same times, slices and shapes (0, default: issued
 *                       directly, which keeps the overlap across phases).
 *   CCFV_OPT_PLAIN_GRAPHS 1: ccfv_plain_solve captures the whole solve as a
 *                       hipGraph and replays it for identical solves.
 *   CCFV_OPT_FUSE       1 (default): in a run of steps the last x sweep of a
 *                       step and the first x sweep of the next run as one
 *                       launch on the same lines (one HBM round trip less);
 *                       0: one launch per sweep.  Both give identical
 *                       results.
 *   CCFV_OPT_PANEL      1 (default): the batched fine propagators keep their
 *                       states between the first and the last x sweep of a
 *                       run in a work buffer laid out as panels of 8 columns
 *                       (contiguous x tiles, whole cache lines for y tiles),
 *                       where the uniform-offset sweeps apply and the grid
 *                       is a multiple of 8 both ways; 0: row-major.  Both
 *                       give identical results. */
#define CCFV_OPT_PIPELINE     1
#define CCFV_OPT_PIPE_GROUP   2
#define CCFV_OPT_GRAPHS       3
#define CCFV_OPT_PLAIN_GRAPHS 4
#define CCFV_OPT_FUSE         5
#define CCFV_OPT_PANEL        6
/*   CCFV_OPT_SERIAL_FINISH 1: a solve that reaches kmax with tol > 0 and an
 *                       increment still above tol finishes with the serial
 *                       fine solve of the same nt_fine steps from its initial
 *                       state (the steps of ccfv_plain_solve), so the result
 *                       always meets the accuracy contract: it is then the
 *                       serial fine solution itself; the report says
 *                       serial_finish = 1, converged = 1 and the call returns
 *                       CCFV_OK.  0 (default): the reference's behaviour, the
 *                       last iterate and CCFV_ENOTCONV. */
#define CCFV_OPT_SERIAL_FINISH 7
/*   CCFV_OPT_KEEP_ITERATES 1: ccfv_parareal_solve keeps the end state
 *                       U^k(t1) of every iteration k in device memory (kmax
 *                       fields, allocated by the solve) for
 *                       ccfv_parareal_iterate; 0 (default): not kept. */
#define CCFV_OPT_KEEP_ITERATES 8
int ccfv_set_option(ccfv_ctx *ctx, int32_t key, int64_t value);
int ccfv_get_option(ccfv_ctx *ctx, int32_t key, int64_t *value);

/* End state U^k(t1) of iteration k (1 <= k <= iterations of the last
 * ccfv_parareal_solve on this context, run with CCFV_OPT_KEEP_ITERATES) into
 * a host array of nx*ny doubles: the iterate whose distance to the serial
 * fine solution the reference prints after each iteration
 * (T2S4_Parareal.cu:1333-1346), from one solve. */
int ccfv_parareal_iterate(ccfv_ctx *ctx, int32_t k, double *host);

/* U = G_new + (F - G_old) over n doubles, device pointers, on the context's
 * stream (parareal_correction_kernel, :771-782). */
int ccfv_parareal_correction(ccfv_ctx *ctx, const double *g_new,
                             const double *g_old, const double *f,
                             double *u, size_t n);

/* Block until all work queued on the context's stream is done. */
int ccfv_sync(ccfv_ctx *ctx);

/* ---- launch shapes ------------------------------------------------------
 * Each sweep launch stages W lines per workgroup and sweeps them with NW
 * wavefronts.  Shapes are kept per sweep direction for one field (plain
 * sweep, Parareal coarse chain) and for batches (fine propagators).
 * tiles[8] = {x1.W, x1.NW, xb.W, xb.NW, y1.W, y1.NW, yb.W, yb.NW}.
 * ccfv_autotune times every valid shape with hipEvents on the context's
 * stream (one field and batch_fields fields, step size h; the state is
 * copied into scratch fields, the context state is untouched) and keeps the
 * fastest.  It has no counterpart in the reference, whose launch shape is a
 * fixed 128-thread 1-D grid (T2S4_Parareal.cu:1073-1080). */
int ccfv_get_tiles(ccfv_ctx *ctx, int32_t tiles[8]);
int ccfv_set_tiles(ccfv_ctx *ctx, const int32_t tiles[8]);
int ccfv_autotune(ccfv_ctx *ctx, int32_t batch_fields, double h,
                  int32_t tiles_out[8]);

/* ---- kernel timing (instrumentation; no counterpart in the reference,
 * which only reads clock() around the whole program, :1137, :1357) -------
 * While enabled, every sweep launch on the context's stream is bracketed by
 * hipEvents and accumulated per kernel class together with its algorithmic
 * HBM bytes (8 B read + 8 B written per cell and field; the fused Parareal
 * epilogue adds 32 B per cell).  Enabling clears the totals.  Meant for a
 * dedicated measurement run: the events add launch gaps. */
#define CCFV_KSTAT_X_BATCH   0   /* x sweeps over a batch of fields           */
#define CCFV_KSTAT_Y_BATCH   1   /* y sweeps over a batch of fields           */
#define CCFV_KSTAT_X_SINGLE  2   /* x sweeps of one field                     */
#define CCFV_KSTAT_Y_SINGLE  3   /* y sweeps of one field                     */
#define CCFV_KSTAT_X_EPILOG  4   /* last x sweep of a coarse step + correction */
#define CCFV_KSTAT_COUNT     5
typedef struct ccfv_kernel_stat {
    int32_t kind;
    int32_t launches;
    double  total_ms;   /* summed hipEvent durations                      */
    double  bytes;      /* summed algorithmic HBM bytes                   */
} ccfv_kernel_stat;
int ccfv_profile_enable(ccfv_ctx *ctx, int32_t enable);
/* Fill out[CCFV_KSTAT_COUNT]; synchronises the stream. */
int ccfv_profile_read(ccfv_ctx *ctx, ccfv_kernel_stat *out);

/* Mean launch time of one sweep class in isolation (instrumentation; no
 * counterpart in the reference): `launches` back-to-back sweeps of direction
 * dir (0 = x, 1 = y) over nfields fields of scratch memory holding copies of
 * the state (the state is not changed), step size h, between two hipEvents
 * on the context's stream.  fused = 1 (x only): the launch a run of steps
 * issues, the last x sub-step of one step and the first x sub-step of the
 * next.  layout 0: row-major fields; 1: the panel layout (8-column panels,
 * uniform-displacement sweeps only; x shapes 8 lines wide).  tile_w, tile_nw:
 * the launch shape (0, 0: the context's shape for that many fields).
 * graph = 1: the launches are captured once as a hipGraph and the replay is
 * timed (no host submission cost per launch).  Each launch reads and writes
 * every cell of every field once: 16 B per cell and field.  *us_per_launch
 * includes the gaps between back-to-back launches. */
int ccfv_time_sweep(ccfv_ctx *ctx, int32_t dir, int32_t fused, int32_t nfields, double h,
                    int32_t launches, int32_t layout, int32_t tile_w, int32_t tile_nw,
                    int32_t graph, double *us_per_launch);

/* Launch-rate floor of the one-field sweeps (instrumentation; no counterpart
 * in the reference).  `launches` back-to-back launches on the context's
 * stream of, for direction dir (0 = x, 1 = y) and the context's one-field
 * launch shape (one wave per line):
 *   kind 0: an empty kernel with that grid;
 *   kind 1: a copy of the state field into a scratch field with the sweep's
 *           memory pattern and block-to-line mapping (y: each lane moves its
 *           contiguous chunk of a row; x: each lane its strided chunk of a
 *           column, column tiles dealt to the XCDs in contiguous runs) and
 *           no arithmetic;
 *   kind 2: a plain solve's pattern without its arithmetic: launches
 *           alternate the y and the x copy of kind 1, each in place on one
 *           scratch copy of the state, so every launch reads what the
 *           previous one wrote (dir is ignored).
 * graph = 1 captures the launches as one hipGraph and times its replay.
 * *us_per_launch: hipEvent time / launches.  The state is not changed. */
int ccfv_dispatch_floor(ccfv_ctx *ctx, int32_t dir, int32_t kind, int32_t launches,
                        int32_t graph, double *us_per_launch);

/* ---- batches of independent problems -----------------------------------
 * (workloads of many independent problems; no counterpart in the reference, which solves
 * one hard-coded problem).  Problems may differ in grid (multiples of 8),
 * velocity, diffusion, boundary mode, time interval, steps and slices.  Every
 * sub-step of every problem is issued through ragged launches (one per sweep
 * stage and line class) in lock-step rounds; each problem's result is
 * bitwise the one ccfv_parareal_solve / ccfv_plain_solve gives it alone.
 * Problem i's state is reached through ccfv_batch_context (set/get state,
 * fill_exact); the batch owns those contexts. */
typedef struct ccfv_batch ccfv_batch;
int  ccfv_batch_create(const ccfv_problem *problems, int32_t n, ccfv_batch **out);
void ccfv_batch_destroy(ccfv_batch *b);
int  ccfv_batch_size(ccfv_batch *b, int32_t *n);
int  ccfv_batch_context(ccfv_batch *b, int32_t i, ccfv_ctx **ctx);
/* plain sweeps: problem i takes nt[i] steps of (t1[i]-t0[i])/nt[i] */
int  ccfv_batch_plain(ccfv_batch *b, const double *t0, const double *t1,
                      const int32_t *nt);
/* Parareal for every problem; each stops on its own increment <= tol (or
 * after kmax iterations, or when all its slices are exact).  reports[i] (may
 * be NULL) receives problem i's report; wall_ms is the batch's. */
int  ccfv_batch_parareal(ccfv_batch *b, const double *t0, const double *t1,
                         const int32_t *nt_fine, const int32_t *n_slices,
                         int32_t kmax, double tol, ccfv_parareal_report *reports);
/* ccfv_set_option on every member context (CCFV_OPT_SERIAL_FINISH: problems
 * of ccfv_batch_parareal that reach kmax above tol finish with their serial
 * fine solve, in the same lock-step rounds as the others). */
int  ccfv_batch_set_option(ccfv_batch *b, int32_t key, int64_t value);
/* For the batch, CCFV_OPT_GRAPHS = 1 also makes its phases of ragged
 * launches hipGraphs: a phase whose launch list has been seen before (same
 * kernels, shapes, item counts and descriptor offsets) is captured once and
 * replayed with one graph launch; the descriptors are uploaded before each
 * replay.  Results are the same either way.
 *
 * Launch shapes of the batch's sweeps, tuned per line class (direction,
 * cells per line, sweep path, boundary mode, source): for each class the
 * member with the most lines is timed (ccfv_autotune's timing, step size
 * (t1-t0)/nt of that member, batch_fields fields for the batch shape) and
 * every member of the class takes the fastest shapes, so members of one
 * class keep sharing one launch per stage.  The coarse chain's shape is not
 * changed (the increments' block sums depend on it): each problem's results
 * stay bitwise those it gets alone.  (ccfv_batch_create gives every member
 * fixed default shapes.) */
int  ccfv_batch_autotune(ccfv_batch *b, const double *t0, const double *t1,
                         const int32_t *nt, int32_t batch_fields);

/* ---- serving engine ------------------------------------------------------
 * (no counterpart in the reference, which solves one problem per process.)
 * Parareal solve requests of mixed sizes, admitted by a device-memory budget
 * and advanced together: continuous batching.  The engine allocates its
 * arena once (mem_budget bytes).  A request waits in arrival order until an
 * arena block of ccfv_request_bytes() and one of max_active slots are free;
 * a request that does not fit is passed over by later ones at most
 * max_bypass times.  Its whole Parareal state then stays resident until it
 * finishes.  Each ccfv_server_step admits what fits and advances every
 * resident request by one Parareal iteration (a new one also runs its
 * k = 0 coarse sweep) through shared ragged launches.  A request's result
 * and report are bitwise those of ccfv_parareal_solve on the same problem
 * alone.  One engine is driven by one host thread. */
typedef struct ccfv_server ccfv_server;
typedef struct ccfv_server_config {
    uint64_t mem_budget;   /* arena bytes (0: half of the free device memory) */
    int32_t  max_active;   /* resident requests, 1..4096 (0: 256)             */
    int32_t  max_slices;   /* largest n_slices accepted, 1..1024 (0: 64)      */
    int32_t  max_bypass;   /* see above (0: 8)                                */
    int32_t  adaptive_above; /* > 0: a Parareal request submitted while this
                              * many requests are queued or resident runs as
                              * the plain serial solve (0: as requested)     */
    int32_t  serial_finish;  /* 1: a Parareal request that reaches kmax above
                              * its tol continues as its serial fine solve
                              * from its initial state, still resident, so
                              * every result meets its tol (as
                              * CCFV_OPT_SERIAL_FINISH); 0: it is returned
                              * with CCFV_ENOTCONV                             */
    int32_t  reserved;
} ccfv_server_config;

typedef struct ccfv_request {
    ccfv_problem problem;  /* grids: multiples of 8 cells                      */
    ccfv_ic  ic;           /* initial data at t0 when submitted without u0    */
    double   t0, t1;
    int32_t  nt_fine, n_slices, kmax, reserved;  /* n_slices 0: the plain
                              * serial solve of nt_fine steps (bitwise
                              * ccfv_plain_solve), PLAIN_CHUNK = 64 steps per
                              * engine step; its report has n_slices 0     */
    double   tol;          /* as ccfv_parareal_solve                          */
} ccfv_request;

typedef struct ccfv_server_status {
    int32_t  active, queued, finished, reserved;  /* finished: not yet taken */
    uint64_t mem_used, mem_budget;
    uint64_t steps;
    double   last_step_ms;  /* device time of the last step                  */
} ccfv_server_status;

/* device bytes a request holds while resident */
int  ccfv_request_bytes(const ccfv_request *r, uint64_t *bytes);
int  ccfv_server_create(const ccfv_server_config *cfg, ccfv_server **out);
void ccfv_server_destroy(ccfv_server *s);
/* queues a request (u0: nx*ny initial cell averages, or NULL for r->ic);
 * CCFV_EBUSY if it can never fit the budget.  *id identifies it. */
int  ccfv_server_submit(ccfv_server *s, const ccfv_request *r, const double *u0,
                        int64_t *id);
/* admission + one iteration of every resident request; blocks until done */
int  ccfv_server_step(ccfv_server *s, ccfv_server_status *st);
/* oldest finished request not yet returned here, or *id = -1 */
int  ccfv_server_next_done(ccfv_server *s, int64_t *id);
/* result of a finished request (u: nx*ny, may be NULL; rep may be NULL;
 * *status: CCFV_OK, CCFV_ENOTCONV, CCFV_ECFL or CCFV_ENONFINITE); frees it */
int  ccfv_server_result(ccfv_server *s, int64_t id, double *u,
                        ccfv_parareal_report *rep, int32_t *status);

#ifdef __cplusplus
}
#endif
#endif /* CCFV_H */
