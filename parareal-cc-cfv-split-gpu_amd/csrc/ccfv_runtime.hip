// ccfv_runtime.hip — host side of the C ABI (include/ccfv.h).
//
// A context owns one state field in HBM, the per-step-size solve tables and,
// once Parareal runs, the slice arrays.  All work is queued on the context's
// own stream (default priority).  Host code never touches field data except
// through ccfv_set_state / ccfv_get_state.
#include "ccfv_ctx.h"
#include "ccfv_feet.h"
#include "ccfv_tables.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <utility>
#include <vector>

using namespace ccfv;
using namespace ccfv::rt;

namespace ccfv {
namespace rt {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(CCFV_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace rt
}  // namespace ccfv


namespace ccfv {
namespace rt {

Tile env_tile(const char *name, Tile dflt)
{
    const char *s = getenv(name);
    if (!s) return dflt;
    Tile t = dflt;
    if (sscanf(s, "%d,%d", &t.w, &t.nw) != 2) return dflt;
    return t;
}

bool tile_ok(Tile t, int n, int nl, Dir dir, bool post, bool src, int bc)
{
    // NW <= 4: the sweep kernels are compiled for at most 256 work-items
    if (t.w <= 0 || t.nw <= 0 || t.nw > 4 || t.w % t.nw || nl % t.w) return false;
    if (t.w & (t.w - 1)) return false;  // power of two: tiles split by shifts
    if (dir == DIR_X && (t.w & 1)) return false;
    // the general sweeps' layout: with the auxiliary line unless the RHS stays
    // in registers (the closed box, the default, always has it)
    const bool aux = (dir == DIR_X && post) || !rhs_in_regs(chunk_for(n), bc, src, false);
    return std::max(sweep_lds_bytes_generic(n, t, src, aux), sweep_lds_bytes_fast(n, t)) <=
           160 * 1024;
}

// 4 lines per workgroup, one per wave (measured at 512^2: pipelined
// Parareal 18.7 ms against 19.3 ms with the one-field sweeps' 2 x 2 shape,
// whose 256 small workgroups crowd the concurrent fine batches harder)
Tile chain_tile_for(const ccfv_problem &p, Tile fallback)
{
    const bool post = (p.flags & CCFV_FLAG_STEP3_POSTREMAP) != 0;
    const Tile t = env_tile("CCFV_TILE_XC", Tile{4, 4});
    return tile_ok(t, p.nx, p.ny, DIR_X, post, p.src_kind != CCFV_SRC_NONE) ? t : fallback;
}

// default tiles (measured on MI355X at 512^2); ccfv_autotune refines them
Tile pick_tile(Dir dir, int n, int nl, bool batch, bool post, bool src)
{
    static const Tile xb[] = {{8, 4}, {16, 4}, {4, 4}, {4, 2}, {2, 2}, {2, 1}};
    static const Tile x1[] = {{2, 2}, {4, 4}, {4, 2}, {2, 1}, {8, 4}};
    static const Tile yb[] = {{2, 1}, {4, 2}, {8, 4}, {1, 1}};
    static const Tile y1[] = {{2, 2}, {1, 1}, {2, 1}, {4, 4}};
    const Tile *c = dir == DIR_X ? (batch ? xb : x1) : (batch ? yb : y1);
    int nc = dir == DIR_X ? (batch ? 6 : 5) : 4;
    for (int i = 0; i < nc; i++)
        if (tile_ok(c[i], n, nl, dir, post, src)) return c[i];
    return Tile{0, 0};
}

double cell_width(const ccfv_problem &P, Dir dir)
{
    return dir == DIR_X ? (P.x_hi - P.x_lo) / P.nx : (P.y_hi - P.y_lo) / P.ny;
}

// lane-scan doubling steps of a table: after s steps a lane's carry
// neglects the product of the factors over 2^s whole chunks; stop once that
// is below 1e-18 (far under FP64 rounding of the carried value).  ch cells
// per lane, the scan over at most 64 lanes (one wave).
static void scan_steps_ch(const double *tab, int n, int ch, int &sf, int &sb)
{
    const int nl = n / ch, lanes = std::min(nl, 64);
    double pf = 0.0, pb = 0.0;
    for (int l = 0; l < nl; l++) {
        double f = 1.0, b = 1.0;
        for (int k = 0; k < ch; k++) {
            f *= std::fabs(tab[l * ch + k]);
            b *= std::fabs(tab[2 * n + l * ch + k]);
        }
        pf = std::max(pf, f);
        pb = std::max(pb, b);
    }
    auto steps = [lanes](double p) {
        int full = 0;
        while ((1 << full) < lanes) full++;
        if (full < 1) full = 1;
        if (p <= 0.0) return 1;
        if (p >= 1.0) return full;
        double need = std::log(1e-18) / std::log(p);  // chunks to span
        int s = 1;
        while ((double)(1 << s) < need && s < full) s++;
        return std::min(s, full);
    };
    sf = steps(pf);
    sb = steps(pb);
}

void scan_steps(const double *tab, int n, SolveTable &T)
{
    scan_steps_ch(tab, n, chunk_for(n), T.sf, T.sb);
    if (fast_wpl(n) > 1) scan_steps_ch(tab, n, 8, T.sfs, T.sbs);
    else T.sfs = T.sf, T.sbs = T.sb;
}

// the scalars of a table (thomas_fill) and, in tab / lines, its arrays on
// the host; lines: 7 nl doubles (shift per line, then 6 weights per line),
// filled and *fast set when every line takes the uniform-offset sweep
void table_host(const ccfv_problem &P, Dir dir, double h, SolveTable &T,
                std::vector<double> &tab, std::vector<double> &lines, bool *fast)
{
    const int n = dir == DIR_X ? P.nx : P.ny;
    const int nl = dir == DIR_X ? P.ny : P.nx;
    std::vector<double> work(2 * (size_t)n);
    tab.assign((size_t)tab_doubles(n), 1.0);
    const TabScalars S =
        thomas_fill(n, cell_width(P, dir), dir == DIR_X ? 0.25 : 0.5, P.nu, h,
                    P.bc == CCFV_BC_PERIODIC, chunk_for(n), tab.data(), work.data());
    T.ca = S.ca;
    T.cb = S.cb;
    T.coef = S.coef;
    T.vl = S.vl;
    T.diag0 = S.diag0;
    T.smden = S.smden;
    scan_steps(tab.data(), n, T);
    *fast = false;
    lines.clear();
    if (!uniform_problem(P) || getenv("CCFV_NO_FAST")) return;
    lines.assign(7 * (size_t)nl, 0.0);
    for (int L = 0; L < nl; L++)
        if (!line_fill(P, (int)dir, h, T.coef, L, &lines[L], &lines[nl + 6 * (size_t)L])) {
            lines.clear();
            return;
        }
    *fast = true;
}

// Thomas factors of the constant tridiagonal matrix of one sweep direction
// and step size, and for the uniform-offset sweep the per-line shift and
// filter weights (thomas_fill, line_fill), in device memory of the context
int get_table(ccfv_ctx *c, Dir dir, double h, const SolveTable **out)
{
    if (c->preset[dir] && h == c->preset_h) {
        *out = c->preset[dir];
        return CCFV_OK;
    }
    if (c->no_alloc) return fail(CCFV_EINVAL, "no table for this step size");
    auto key = std::make_pair((int)dir, h);
    auto it = c->tables.find(key);
    if (it != c->tables.end()) {
        *out = &it->second;
        return CCFV_OK;
    }
    SolveTable T;
    std::vector<double> tab, ld;
    bool fast = false;
    table_host(c->P, dir, h, T, tab, ld, &fast);
    HIPCHK(hipMalloc(&T.dev, sizeof(double) * tab.size()));
    HIPCHK(hipMemcpy(T.dev, tab.data(), sizeof(double) * tab.size(),
                     hipMemcpyHostToDevice));
    if (fast) {
        HIPCHK(hipMalloc(&T.line_dev, sizeof(double) * ld.size()));
        HIPCHK(hipMemcpy(T.line_dev, ld.data(), sizeof(double) * ld.size(),
                         hipMemcpyHostToDevice));
        T.fast_ok = true;
    }
    auto res = c->tables.emplace(key, T);
    *out = &res.first->second;
    return CCFV_OK;
}

SweepArgs make_args(ccfv_ctx *c, Dir dir, int substep, const SolveTable &T,
                    double h, double t0, double tDT, int field0, double jdt)
{
    const ccfv_problem &P = c->P;
    SweepArgs a;
    memset(&a, 0, sizeof(a));
    a.ld = P.ny;
    if (dir == DIR_X) {
        a.n = P.nx;
        a.nl = P.ny;
        a.lo = P.x_lo;
        a.hi = P.x_hi;
        a.olo = P.y_lo;
        a.eps = 1e-15;   // :323-324
        a.tau = 0.5 * h; // the x sub-steps trace over dt/2 (:215-217)
    } else {
        a.n = P.ny;
        a.nl = P.nx;
        a.lo = P.y_lo;
        a.hi = P.y_hi;
        a.olo = P.x_lo;
        a.eps = 1e-16;   // :414-415
        a.tau = h;       // the y sub-step traces over dt (:219-221)
    }
    a.d = cell_width(P, dir);
    a.od = cell_width(P, dir == DIR_X ? DIR_Y : DIR_X);
    a.quirk = (P.flags & CCFV_FLAG_DERIV_GHOST_QUIRK) ? 1 : 0;
    a.sf = T.sf;
    a.sb = T.sb;
    a.sfs = T.sfs;
    a.sbs = T.sbs;
    a.coef = T.coef;
    a.ca = T.ca;
    a.diag0 = T.diag0;
    a.vl = T.vl;
    a.smrcp = 1.0 / T.smden;
    a.tab = T.dev;
    a.lineS = T.line_dev;
    a.lineK = T.line_dev ? T.line_dev + a.nl : nullptr;
    a.id = 1.0 / a.d;
    if (c->trig) {
        a.trig = dir == DIR_X ? c->trig : c->trig + 2 * (size_t)(P.nx + 1);
        a.trig_small = M_PI * std::fabs(a.tau) <= 0.25 ? 1 : 0;
    }
    a.vel_kind = P.vel_kind;
    a.v0 = P.vel[0];
    a.v1 = P.vel[1];
    a.v2 = P.vel[2];
    a.t0 = t0;
    a.tDT = tDT;
    a.jdt = jdt;
    a.h = h;
    a.field0 = field0;
    a.substep = substep;
    a.flag = c->d_flag;
    a.src_kind = P.src_kind;
    for (int i = 0; i < 5; i++) a.src[i] = P.src[i];
    // theta*h: 0.25*dt for x, 0.5*dt for y (T2S4_Parareal.cu:575, :607)
    a.src_w = (dir == DIR_X ? 0.25 : 0.5) * h;
    return a;
}

hipEvent_t pool_event(ccfv_ctx *c)
{
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// fold finished records into the totals (caller has synchronised)
void drain_records(ccfv_ctx *c)
{
    for (auto &r : c->recs) {
        float ms = 0.f;
        if (r.a && r.b && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            ccfv_kernel_stat &s = c->stats[r.kind];
            s.launches += 1;
            s.total_ms += ms;
            s.bytes += r.bytes;
        }
        if (r.a) c->ev_pool.push_back(r.a);
        if (r.b) c->ev_pool.push_back(r.b);
    }
    c->recs.clear();
}

// launch one sweep, timed when profiling is on
hipError_t run_sweep(ccfv_ctx *c, Dir dir, bool post, Epi epi, const SweepArgs &a,
                     int nf, Tile t)
{
    if (nf <= 0) return hipSuccess;
    if (a.lineS) return launch_sweep_fast(dir, epi, a, nf, t, c->cur);
    // the general sweeps read and write row-major fields only
    if (a.in_panel || a.out_panel) return hipErrorInvalidValue;
    return launch_sweep_generic(dir, c->P.bc, post, epi, a, nf, t, c->cur);
}

hipError_t timed_sweep(ccfv_ctx *c, int kind, double bytes, Dir dir, bool post,
                       Epi epi, const SweepArgs &a, int nf, Tile t)
{
    if (!c->prof) return run_sweep(c, dir, post, epi, a, nf, t);
    if (c->recs.size() > 4096) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream2);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream3);
        if (e != hipSuccess) return e;
        drain_records(c);
    }
    ccfv_ctx::Rec r{kind, bytes, pool_event(c), pool_event(c)};
    if (r.a) (void)hipEventRecord(r.a, c->cur);
    hipError_t e = run_sweep(c, dir, post, epi, a, nf, t);
    if (r.b) (void)hipEventRecord(r.b, c->cur);
    c->recs.push_back(r);
    return e;
}

bool post_flag(const ccfv_ctx *c)
{
    return (c->P.flags & CCFV_FLAG_STEP3_POSTREMAP) != 0;
}

hipError_t run_sweep_fused(ccfv_ctx *c, bool post, Epi epi, const SweepArgs &a, int nf,
                           Tile t)
{
    if (nf <= 0) return hipSuccess;
    if (a.lineS) return launch_sweep_fast(DIR_X, epi, a, nf, t, c->cur, true);
    if (a.in_panel || a.out_panel) return hipErrorInvalidValue;
    return launch_sweep_generic(DIR_X, c->P.bc, post, epi, a, nf, t, c->cur, true);
}

namespace {

// profiling classes and algorithmic bytes of the x launches of a run
int x_kind(Epi epi, bool batch)
{
    return epi != EPI_NONE ? CCFV_KSTAT_X_EPILOG
                           : (batch ? CCFV_KSTAT_X_BATCH : CCFV_KSTAT_X_SINGLE);
}

double x_bytes(const ccfv_ctx *c, Epi epi, int nf, bool fused)
{
    const double f = 8.0 * (double)c->ncell;
    // epilogues: read G_old, F, U_old; write G, U[n+1] (EPI_INIT: write G,
    // U[n+1]); a fused launch also writes the next step's first x sweep
    double b = epi == EPI_CORRECT ? 6.0 * f : epi == EPI_INIT ? 3.0 * f : 2.0 * f * nf;
    if (fused && epi != EPI_NONE) b += f;
    return b;
}

}  // namespace

hipError_t timed_sweep_fused(ccfv_ctx *c, int kind, double bytes, bool post, Epi epi,
                             const SweepArgs &a, int nf, Tile t)
{
    if (!c->prof) return run_sweep_fused(c, post, epi, a, nf, t);
    if (c->recs.size() > 4096) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream2);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream3);
        if (e != hipSuccess) return e;
        drain_records(c);
    }
    ccfv_ctx::Rec r{kind, bytes, pool_event(c), pool_event(c)};
    if (r.a) (void)hipEventRecord(r.a, c->cur);
    hipError_t e = run_sweep_fused(c, post, epi, a, nf, t);
    if (r.b) (void)hipEventRecord(r.b, c->cur);
    c->recs.push_back(r);
    return e;
}

// Strang steps x(h/2) -> y(h) -> x(h/2) on nf fields (GpuOneStep,
// T2S4_Parareal.cu:785-968).  A step's first sweep reads `in` and writes
// `out`; the other two work in place on `out` (every sweep is line-local).
// In a run of steps where step i+1 reads step i's result, the last x sweep
// of step i and the first x sweep of step i+1 are one launch over the same
// lines: the same arithmetic on the same values, one HBM round trip and one
// launch fewer per step.
// the panel layout needs uniform-offset sweeps in both directions (the
// generic sweeps read row-major fields only), x batch tiles of PANEL_W
// lines, whole panels in both directions and a work buffer of nf fields
bool panel_ready(const ccfv_ctx *c, const SolveTable &tx, const SolveTable &ty, int nf)
{
    return c->panel_on && c->panel && nf <= c->panel_fields && tx.line_dev && ty.line_dev &&
           c->tile[DIR_X][1].w == PANEL_W && c->P.nx % PANEL_W == 0 &&
           c->P.ny % PANEL_W == 0;
}

// the panel work buffer for batches of up to nf fields stepped by h; left
// unallocated (batches then sweep row-major) where the layout cannot apply
// or the memory is not there
int ensure_panel(ccfv_ctx *c, int nf, double h)
{
    const SolveTable *tx, *ty;
    int rc;
    if ((rc = get_table(c, DIR_X, h, &tx)) || (rc = get_table(c, DIR_Y, h, &ty))) return rc;
    if (!c->panel_on || nf < 2 || !tx->line_dev || !ty->line_dev ||
        c->P.nx % PANEL_W || c->P.ny % PANEL_W || c->panel_fields >= nf)
        return CCFV_OK;
    (void)hipFree(c->panel);
    c->panel = nullptr;
    c->panel_fields = 0;
    if (hipMalloc(&c->panel, sizeof(double) * c->ncell * (size_t)nf) != hipSuccess) {
        (void)hipGetLastError();
        c->panel = nullptr;
        return CCFV_OK;
    }
    c->panel_fields = nf;
    return CCFV_OK;
}

int run_steps(ccfv_ctx *c, int count, const std::function<StepDesc(int)> &desc,
              const std::function<int(int)> &pre, const std::function<int(int)> &post)
{
    if (count <= 0) return CCFV_OK;
    const bool postflag = post_flag(c);
    const long long fs = (long long)c->ncell;
    int rc;
    StepDesc D = desc(0);
    const SolveTable *tx, *ty;
    if ((rc = get_table(c, DIR_X, D.h, &tx)) || (rc = get_table(c, DIR_Y, D.h, &ty)))
        return rc;
    // a batch of uniform-offset sweeps keeps its states between the run's
    // first and last x sweeps in the panel work buffer; every step of the run
    // has the same step size and field count
    bool panel = D.nf > 1 && panel_ready(c, *tx, *ty, D.nf) && D.epi == EPI_NONE;
    for (int i = 1; panel && i < count; i++) {
        const StepDesc S = desc(i);
        panel = S.h == D.h && S.nf == D.nf && S.epi == EPI_NONE;
    }
    // where the x1 sweep of a step writes, and the y and x3 sweeps work
    auto work = [&](const StepDesc &S) { return panel ? c->panel : S.out; };
    auto x1_args = [&](const StepDesc &S, const SolveTable &T) {
        SweepArgs a = make_args(c, DIR_X, 1, T, S.h, S.t0, S.tDT, S.field0, S.jdt);
        a.in = S.in;
        a.out = work(S);
        a.out_panel = panel;
        a.fstride = fs;
        return a;
    };
    if (pre && (rc = pre(0))) return rc;
    {
        const bool batch = D.nf > 1;
        HIPCHK(timed_sweep(c, batch ? CCFV_KSTAT_X_BATCH : CCFV_KSTAT_X_SINGLE,
                           x_bytes(c, EPI_NONE, D.nf, false), DIR_X, false, EPI_NONE,
                           x1_args(D, *tx), D.nf, c->tile[DIR_X][batch]));
    }
    for (int i = 0; i < count; i++) {
        const bool batch = D.nf > 1;
        SweepArgs b = make_args(c, DIR_Y, 2, *ty, D.h, D.t0, D.tDT, D.field0, D.jdt);
        b.in = work(D);
        b.out = work(D);
        b.in_panel = b.out_panel = panel;
        b.fstride = fs;
        HIPCHK(timed_sweep(c, batch ? CCFV_KSTAT_Y_BATCH : CCFV_KSTAT_Y_SINGLE,
                           16.0 * (double)c->ncell * D.nf, DIR_Y, false, EPI_NONE, b, D.nf,
                           c->tile[DIR_Y][batch]));
        SweepArgs x3 = make_args(c, DIR_X, 3, *tx, D.h, D.t0, D.tDT, D.field0, D.jdt);
        x3.in = work(D);
        x3.in_panel = panel;
        x3.out = D.out;
        x3.fstride = fs;
        x3.g_cache = D.g_cache;
        x3.f_in = D.f_in;
        x3.u_next = D.u_next;
        x3.partial = D.partial;
        const Tile xt = D.epi != EPI_NONE ? c->chain_tile : c->tile[DIR_X][batch];
        if (i + 1 == count) {
            HIPCHK(timed_sweep(c, x_kind(D.epi, batch), x_bytes(c, D.epi, D.nf, false),
                               DIR_X, postflag, D.epi, x3, D.nf, xt));
            if (post && (rc = post(i))) return rc;
            break;
        }
        StepDesc N = desc(i + 1);
        const double *res = D.epi == EPI_NONE ? D.out : D.u_next;
        const bool fuse = c->fuse && N.in == res && N.h == D.h && N.nf == D.nf &&
                          N.tDT == D.tDT;
        if (pre && (rc = pre(i + 1))) return rc;
        if (fuse) {
            // the next step's x1 result: its work field
            x3.out2 = work(N);
            x3.out_panel = panel;
            x3.field0_2 = N.field0;
            x3.t0_2 = N.t0;
            x3.jdt2 = N.jdt;
            HIPCHK(timed_sweep_fused(c, x_kind(D.epi, batch), x_bytes(c, D.epi, D.nf, true),
                                     postflag, D.epi, x3, D.nf, xt));
            if (post && (rc = post(i))) return rc;
        } else {
            HIPCHK(timed_sweep(c, x_kind(D.epi, batch), x_bytes(c, D.epi, D.nf, false),
                               DIR_X, postflag, D.epi, x3, D.nf, xt));
            if (post && (rc = post(i))) return rc;
            if (N.h != D.h &&
                ((rc = get_table(c, DIR_X, N.h, &tx)) || (rc = get_table(c, DIR_Y, N.h, &ty))))
                return rc;
            const bool nb = N.nf > 1;
            HIPCHK(timed_sweep(c, nb ? CCFV_KSTAT_X_BATCH : CCFV_KSTAT_X_SINGLE,
                               x_bytes(c, EPI_NONE, N.nf, false), DIR_X, false, EPI_NONE,
                               x1_args(N, *tx), N.nf, c->tile[DIR_X][nb]));
        }
        D = N;
    }
    return CCFV_OK;
}

int step_fields(ccfv_ctx *c, const double *in, double *out, int nf, double t0,
                double tDT, int field0, double jdt, double h, Epi epi,
                double *g_cache, const double *f_in, double *u_next, double *partial)
{
    StepDesc d{in, out, nf, t0, tDT, field0, jdt, h, epi, g_cache, f_in, u_next, partial};
    return run_steps(c, 1, [&](int) { return d; });
}

int check_flag(ccfv_ctx *c)
{
    HIPCHK(hipMemcpyAsync(c->h_flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (*c->h_flag) {
        *c->h_flag = 0;
        HIPCHK(hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return fail(CCFV_ECFL,
                    "characteristic feet crossed: the time step is too large "
                    "for this velocity field");
    }
    return CCFV_OK;
}

int validate_problem(const ccfv_problem *p)
{
    if (!p) return fail(CCFV_EINVAL, "problem is NULL");
    if (!supported_n(p->nx) || !supported_n(p->ny))
        return fail(CCFV_EINVAL,
                    "unsupported grid: nx and ny must be even, 8..2048, and a "
                    "multiple of ceil-pow2(n/64)");
    if (p->nx < 8 || p->ny < 8) return fail(CCFV_EINVAL, "grid too small");
    if (p->bc != CCFV_BC_REFERENCE && p->bc != CCFV_BC_PERIODIC)
        return fail(CCFV_EINVAL, "unknown boundary mode");
    if (p->vel_kind < 0 || p->vel_kind > CCFV_VEL_SWIRL)
        return fail(CCFV_EINVAL, "unknown velocity kind");
    if (!(p->x_hi > p->x_lo) || !(p->y_hi > p->y_lo))
        return fail(CCFV_EINVAL, "empty box");
    if (!(p->nu >= 0.0) || !std::isfinite(p->nu))
        return fail(CCFV_EINVAL, "diffusion coefficient must be >= 0");
    if (p->vel_kind == CCFV_VEL_SWIRL && !(p->vel[0] > 0.0))
        return fail(CCFV_EINVAL, "swirl period must be > 0");
    if (p->src_kind < CCFV_SRC_NONE || p->src_kind > CCFV_SRC_SINSIN)
        return fail(CCFV_EINVAL, "unknown source kind");
    for (int i = 0; i < 5; i++)
        if (!std::isfinite(p->src[i])) return fail(CCFV_EINVAL, "source parameters must be finite");
    if (p->bc == CCFV_BC_PERIODIC && (p->flags & CCFV_FLAG_STEP3_POSTREMAP))
        return fail(CCFV_EINVAL,
                    "the third-sub-step post-remap derivative is a reference "
                    "(closed box) behaviour only");
    return CCFV_OK;
}

void clear_graphs(ccfv_ctx *c)
{
    for (auto &kv : c->gcache) (void)hipGraphExecDestroy(kv.second);
    c->gcache.clear();
}

// A context keeps the tables of at most MAX_TABLES (direction, step size)
// pairs.  Past that, the next one-context call drops them all before it
// starts — with its streams idle and no capture in progress — together with
// the captured graphs whose launches point at them; the call then builds
// what it needs again.
constexpr size_t MAX_TABLES = 32;

int trim_tables(ccfv_ctx *c)
{
    if (c->tables.size() <= MAX_TABLES) return CCFV_OK;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(c->stream, &st));
    if (st != hipStreamCaptureStatusNone) return CCFV_OK;
    for (hipStream_t s : {c->stream, c->stream2, c->stream3})
        if (s) HIPCHK(hipStreamSynchronize(s));
    clear_graphs(c);
    for (auto &kv : c->tables) {
        HIPCHK(hipFree(kv.second.dev));
        HIPCHK(hipFree(kv.second.line_dev));
    }
    c->tables.clear();
    return CCFV_OK;
}

void free_para(ccfv_ctx *c)
{
    clear_graphs(c);
    (void)hipFree(c->U);
    (void)hipFree(c->G);
    (void)hipFree(c->F);
    (void)hipFree(c->Wk);
    (void)hipFree(c->Wk2);
    (void)hipFree(c->partial2);
    (void)hipFree(c->res2);
    c->Wk2 = c->partial2 = c->res2 = nullptr;
    (void)hipFree(c->partial);
    (void)hipFree(c->slot_sums);
    if (c->h_sums) (void)hipHostFree(c->h_sums);
    c->U = c->G = c->F = c->Wk = c->partial = c->slot_sums = c->h_sums = nullptr;
    (void)hipFree(c->panel);
    c->panel = nullptr;
    c->panel_fields = 0;
    c->para_ns = 0;
}

int ensure_trig(ccfv_ctx *c, hipStream_t s)
{
    if (c->trig || c->P.vel_kind != CCFV_VEL_SWIRL) return CCFV_OK;
    const ccfv_problem &P = c->P;
    const size_t nb = 2 * (size_t)(P.nx + 1) + 2 * (size_t)(P.ny + 1);
    if (hipMalloc(&c->trig, sizeof(double) * nb) != hipSuccess) {
        (void)hipGetLastError();
        c->trig = nullptr;
        return fail(CCFV_ENOMEM, "swirl face table");
    }
    HIPCHK(launch_fill_trig(P.x_lo, cell_width(P, DIR_X), P.nx, c->trig, s));
    HIPCHK(launch_fill_trig(P.y_lo, cell_width(P, DIR_Y), P.ny,
                            c->trig + 2 * (size_t)(P.nx + 1), s));
    return CCFV_OK;
}

int ensure_para(ccfv_ctx *c, int ns)
{
    if (c->para_ns == ns && c->chain_blocks == c->P.ny / c->chain_tile.w)
        return CCFV_OK;
    free_para(c);
    const size_t fb = sizeof(double) * c->ncell;
    c->chain_blocks = c->P.ny / c->chain_tile.w;
    hipError_t e;
    if ((e = hipMalloc(&c->U, fb * (ns + 1))) != hipSuccess ||
        (e = hipMalloc(&c->G, fb * ns)) != hipSuccess ||
        (e = hipMalloc(&c->F, fb * ns)) != hipSuccess ||
        (e = hipMalloc(&c->Wk, fb)) != hipSuccess ||
        (e = hipMalloc(&c->Wk2, fb)) != hipSuccess ||
        (e = hipMalloc(&c->res2, 2 * fb)) != hipSuccess ||
        (e = hipMalloc(&c->partial, sizeof(double) * 2 * c->chain_blocks *
                                        (size_t)(ns + 1))) != hipSuccess ||
        (e = hipMalloc(&c->partial2, sizeof(double) * 2 * c->chain_blocks *
                                         (size_t)(ns + 1))) != hipSuccess ||
        (e = hipMalloc(&c->slot_sums, sizeof(double) * 2 * (ns + 1) *
                                          CCFV_MAX_ITERS)) != hipSuccess ||
        (e = hipHostMalloc(&c->h_sums, sizeof(double) * 2 * (ns + 1) *
                                           CCFV_MAX_ITERS)) != hipSuccess) {
        free_para(c);
        return fail(CCFV_ENOMEM, std::string("Parareal buffers: ") +
                                     hipGetErrorString(e));
    }
    c->para_ns = ns;
    return CCFV_OK;
}

}  // namespace rt
}  // namespace ccfv

extern "C" {

const char *ccfv_last_error(void) { return g_err.c_str(); }

int ccfv_version(int32_t *major, int32_t *minor)
{
    if (major) *major = 0;
    if (minor) *minor = 1;
    return CCFV_OK;
}

int ccfv_device_count(int32_t *count)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        if (count) *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    if (count) *count = n;
    return CCFV_OK;
}

int ccfv_problem_reference(ccfv_problem *p, int32_t n, double K)
{
    if (!p) return fail(CCFV_EINVAL, "NULL problem");
    memset(p, 0, sizeof(*p));
    p->nx = p->ny = n;                    // :1037-1039
    p->bc = CCFV_BC_REFERENCE;
    p->vel_kind = CCFV_VEL_SOLID_BODY;    // Vx=-4y, Vy=4x (:44-45)
    p->vel[0] = 4.0;
    p->x_lo = -1.0;                       // :30-33
    p->x_hi = 1.0;
    p->y_lo = -1.0;
    p->y_hi = 1.0;
    p->nu = K;                            // :1047-1048
    p->flags = CCFV_FLAGS_REFERENCE;
    return CCFV_OK;
}

int ccfv_ic_reference(ccfv_ic *ic, double K)
{
    (void)K;  // the diffusion enters through the problem's nu
    if (!ic) return fail(CCFV_EINVAL, "NULL ic");
    memset(ic, 0, sizeof(*ic));
    ic->kind = CCFV_IC_ROT_GAUSSIAN;      // C_Exact (:51-56), x0,y0,sigma :27-29
    ic->p[0] = -0.35;
    ic->p[1] = 0.0;
    ic->p[2] = 0.005;
    ic->p[3] = 4.0;
    return CCFV_OK;
}

int ccfv_create(const ccfv_problem *p, ccfv_ctx **out)
{
    if (!out) return fail(CCFV_EINVAL, "NULL out");
    *out = nullptr;
    int rc = validate_problem(p);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(CCFV_EDEVICE, "no HIP device available");
    ccfv_ctx *c = new ccfv_ctx();
    c->P = *p;
    if (hipError_t e0 = hipGetDevice(&c->device); e0 != hipSuccess) {
        delete c;
        return hip_fail(e0, "hipGetDevice");
    }
    c->ncell = (size_t)p->nx * p->ny;
    const bool post = (p->flags & CCFV_FLAG_STEP3_POSTREMAP) != 0;
    const bool src = p->src_kind != CCFV_SRC_NONE;
    for (int b = 0; b < 2; b++) {
        c->tile[DIR_X][b] = pick_tile(DIR_X, p->nx, p->ny, b, post, src);
        c->tile[DIR_Y][b] = pick_tile(DIR_Y, p->ny, p->nx, b, false, src);
    }
    c->tile[DIR_X][0] = env_tile("CCFV_TILE_X1", c->tile[DIR_X][0]);
    c->tile[DIR_X][1] = env_tile("CCFV_TILE_XB", c->tile[DIR_X][1]);
    c->tile[DIR_Y][0] = env_tile("CCFV_TILE_Y1", c->tile[DIR_Y][0]);
    c->tile[DIR_Y][1] = env_tile("CCFV_TILE_YB", c->tile[DIR_Y][1]);
    c->chain_tile = chain_tile_for(*p, c->tile[DIR_X][0]);
    for (int d = 0; d < 2; d++)
        for (int b = 0; b < 2; b++)
            if (!tile_ok(c->tile[d][b], d == DIR_X ? p->nx : p->ny,
                         d == DIR_X ? p->ny : p->nx, (Dir)d, post && d == DIR_X, src, p->bc)) {
                delete c;
                return fail(CCFV_EINVAL, "no valid tile for this grid");
            }
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) !=
            hipSuccess ||
        (e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking)) !=
            hipSuccess ||
        (e = hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking)) !=
            hipSuccess ||
        (e = hipMalloc(&c->state, sizeof(double) * c->ncell)) != hipSuccess ||
        (e = hipMalloc(&c->d_flag, sizeof(int))) != hipSuccess ||
        (e = hipHostMalloc(&c->h_flag, sizeof(int))) != hipSuccess ||
        (e = hipEventCreate(&c->ev0)) != hipSuccess ||
        (e = hipEventCreate(&c->ev1)) != hipSuccess) {
        ccfv_destroy(c);
        return fail(CCFV_ENOMEM, std::string("context allocation: ") +
                                     hipGetErrorString(e));
    }
    *c->h_flag = 0;
    c->cur = c->stream;
    if (int rc2 = ensure_trig(c, c->stream)) {
        ccfv_destroy(c);
        return rc2;
    }
    // stream names in rocprofv3 traces
    roctxNameHipStream("ccfv chain", c->stream);
    roctxNameHipStream("ccfv fine", c->stream2);
    roctxNameHipStream("ccfv chain (odd iterations)", c->stream3);
    if (const char *s = getenv("CCFV_PIPE")) c->pipeline = std::min(2, std::max(0, atoi(s)));
    if (const char *s = getenv("CCFV_PIPE_GROUP")) c->pipe_group = std::max(1, atoi(s));
    if (const char *s = getenv("CCFV_GRAPHS")) c->graphs = atoi(s) != 0;
    if (const char *s = getenv("CCFV_PLAIN_GRAPHS")) c->plain_graphs = atoi(s) != 0;
    if (const char *s = getenv("CCFV_FUSE")) c->fuse = atoi(s) != 0;
    if (const char *s = getenv("CCFV_PANEL")) c->panel_on = atoi(s) != 0;
    (void)hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream);
    (void)hipMemsetAsync(c->state, 0, sizeof(double) * c->ncell, c->stream);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) {
        ccfv_destroy(c);
        return hip_fail(e, "context init");
    }
    *out = c;
    return CCFV_OK;
}

void ccfv_destroy(ccfv_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    if (c->stream3) (void)hipStreamSynchronize(c->stream3);
    free_para(c);
    (void)hipFree(c->iter_buf);
    drain_records(c);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->sync_ev) (void)hipEventDestroy(e);
    for (auto &kv : c->tables) {
        (void)hipFree(kv.second.dev);
        (void)hipFree(kv.second.line_dev);
    }
    (void)hipFree(c->state);
    (void)hipFree(c->trig);
    (void)hipFree(c->d_flag);
    if (c->h_flag) (void)hipHostFree(c->h_flag);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream3) (void)hipStreamDestroy(c->stream3);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int ccfv_set_state(ccfv_ctx *c, const double *host)
{
    if (!c || !host) return fail(CCFV_EINVAL, "NULL argument");
    HIPCHK(hipMemcpyAsync(c->state, host, sizeof(double) * c->ncell,
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CCFV_OK;
}

int ccfv_get_state(ccfv_ctx *c, double *host)
{
    if (!c || !host) return fail(CCFV_EINVAL, "NULL argument");
    HIPCHK(hipMemcpyAsync(host, c->state, sizeof(double) * c->ncell,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CCFV_OK;
}

int ccfv_set_state_device(ccfv_ctx *c, const double *dev)
{
    if (!c || !dev) return fail(CCFV_EINVAL, "NULL argument");
    HIPCHK(launch_copy(dev, c->state, c->ncell, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CCFV_OK;
}

int ccfv_get_state_device(ccfv_ctx *c, double *dev)
{
    if (!c || !dev) return fail(CCFV_EINVAL, "NULL argument");
    HIPCHK(launch_copy(c->state, dev, c->ncell, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CCFV_OK;
}

int ccfv_state_device_ptr(ccfv_ctx *c, void **dev)
{
    if (!c || !dev) return fail(CCFV_EINVAL, "NULL argument");
    *dev = c->state;
    return CCFV_OK;
}

int ccfv_fill_exact(ccfv_ctx *c, const ccfv_ic *ic, double t)
{
    if (!c || !ic) return fail(CCFV_EINVAL, "NULL argument");
    if (ic->kind < 0 || ic->kind > CCFV_IC_COSINE_BELL)
        return fail(CCFV_EINVAL, "unknown initial condition");
    HIPCHK(launch_fill_exact(c->P, *ic, t, c->state, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CCFV_OK;
}

int ccfv_exact_to_host(ccfv_ctx *c, const ccfv_ic *ic, double t, double *host)
{
    if (!c || !ic || !host) return fail(CCFV_EINVAL, "NULL argument");
    if (ic->kind < 0 || ic->kind > CCFV_IC_COSINE_BELL)
        return fail(CCFV_EINVAL, "unknown initial condition");
    double *tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, sizeof(double) * c->ncell));
    hipError_t e = launch_fill_exact(c->P, *ic, t, tmp, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(host, tmp, sizeof(double) * c->ncell,
                           hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "exact_to_host");
    return CCFV_OK;
}

int ccfv_step(ccfv_ctx *c, double tn, double h)
{
    if (!c) return fail(CCFV_EINVAL, "NULL context");
    if (!(h > 0.0) || !std::isfinite(h) || !std::isfinite(tn))
        return fail(CCFV_EINVAL, "step size must be finite and > 0");
    int rc = trim_tables(c);
    if (rc) return rc;
    rc = step_fields(c, c->state, c->state, 1, tn, 0.0, 0, 0.0, h);
    if (rc) return rc;
    return check_flag(c);
}

int ccfv_coarse(ccfv_ctx *c, double tn, double DT)
{
    // CoarsePropagator (T2S4_Parareal.cu:971-974): one step of DT
    return ccfv_step(c, tn, DT);
}

int ccfv_fine(ccfv_ctx *c, double tn, double DT, int32_t J)
{
    // FinePropagator (T2S4_Parareal.cu:977-984): J steps of DT/J
    if (!c) return fail(CCFV_EINVAL, "NULL context");
    if (J < 1) return fail(CCFV_EINVAL, "J must be >= 1");
    if (!(DT > 0.0) || !std::isfinite(DT))
        return fail(CCFV_EINVAL, "DT must be finite and > 0");
    double dt = DT / (double)J;
    int rc = trim_tables(c);
    if (rc) return rc;
    rc = run_steps(c, J, [&](int j) {
        return StepDesc{c->state, c->state, 1, tn, 0.0, 0, j * dt, dt};
    });
    if (rc) return rc;
    return check_flag(c);
}

int ccfv_advance(ccfv_ctx *c, double t_origin, double dt, int32_t n_begin, int32_t n_end)
{
    if (!c) return fail(CCFV_EINVAL, "NULL context");
    if (n_begin < 0 || n_end < n_begin) return fail(CCFV_EINVAL, "need 0 <= n_begin <= n_end");
    if (!(dt > 0.0) || !std::isfinite(dt) || !std::isfinite(t_origin))
        return fail(CCFV_EINVAL, "step size must be finite and > 0");
    ccfv::rt::TraceRange tr("ccfv_advance");
    int rc = trim_tables(c);
    if (rc) return rc;
    c->cur = c->stream;
    rc = run_steps(c, n_end - n_begin, [&](int i) {
        const int n = n_begin + i;
        return StepDesc{c->state, c->state, 1, t_origin + n * dt, 0.0, 0, 0.0, dt};
    });
    if (rc) return rc;
    return check_flag(c);
}

int ccfv_plain_solve(ccfv_ctx *c, double t0, double t1, int32_t nt)
{
    // serial fine loop (T2S4_Parareal.cu:1060, :1234-1236)
    if (!c) return fail(CCFV_EINVAL, "NULL context");
    if (nt < 1) return fail(CCFV_EINVAL, "nt must be >= 1");
    if (!(t1 > t0) || !std::isfinite(t0) || !std::isfinite(t1))
        return fail(CCFV_EINVAL, "need finite t1 > t0");
    double dt = (t1 - t0) / nt;
    ccfv::rt::TraceRange tr("ccfv_plain_solve");
    if (int r0 = trim_tables(c)) return r0;
    c->cur = c->stream;
    auto body = [&]() -> int {
        return run_steps(c, nt, [&](int n) {
            return StepDesc{c->state, c->state, 1, t0 + n * dt, 0.0, 0, 0.0, dt};
        });
    };
    if (c->plain_graphs && !c->prof) {
        // the whole solve as one hipGraph, replayed by later identical solves
        const SolveTable *tt;
        int rc;
        if ((rc = get_table(c, DIR_X, dt, &tt)) || (rc = get_table(c, DIR_Y, dt, &tt)))
            return rc;
        char key[160];
        snprintf(key, sizeof(key), "pl|%a|%a|%d|%d", t0, t1, nt, c->tiles_version);
        auto it = c->gcache.find(key);
        if (it == c->gcache.end()) {
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
            int r2 = body();
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamEndCapture(c->stream, &gr);
            if (r2 || e != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                return r2 ? r2 : hip_fail(e, "plain-solve graph capture");
            }
            hipGraphExec_t x = nullptr;
            e = hipGraphInstantiate(&x, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
            it = c->gcache.emplace(key, x).first;
        }
        HIPCHK(hipGraphLaunch(it->second, c->stream));
    } else {
        int rc = body();
        if (rc) return rc;
    }
    return check_flag(c);
}

int ccfv_parareal_iterate(ccfv_ctx *c, int32_t k, double *host)
{
    if (!c || !host) return fail(CCFV_EINVAL, "NULL argument");
    if (k < 1 || k > c->iter_valid)
        return fail(CCFV_EINVAL, "no kept iterate k: run ccfv_parareal_solve with "
                                 "CCFV_OPT_KEEP_ITERATES and 1 <= k <= its iterations");
    HIPCHK(hipMemcpyAsync(host, c->iter_buf + (size_t)(k - 1) * c->ncell,
                          sizeof(double) * c->ncell, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CCFV_OK;
}

int ccfv_parareal_correction(ccfv_ctx *c, const double *g_new,
                             const double *g_old, const double *f, double *u,
                             size_t n)
{
    if (!c || !g_new || !g_old || !f || !u) return fail(CCFV_EINVAL, "NULL argument");
    int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    if (blocks < 1) blocks = 1;
    HIPCHK(launch_correction(g_new, g_old, f, u, n, nullptr, blocks, c->stream));
    return CCFV_OK;
}

int ccfv_get_tiles(ccfv_ctx *c, int32_t t[8])
{
    if (!c || !t) return fail(CCFV_EINVAL, "NULL argument");
    const Tile v[4] = {c->tile[DIR_X][0], c->tile[DIR_X][1], c->tile[DIR_Y][0],
                       c->tile[DIR_Y][1]};
    for (int i = 0; i < 4; i++) {
        t[2 * i] = v[i].w;
        t[2 * i + 1] = v[i].nw;
    }
    return CCFV_OK;
}

int ccfv_set_tiles(ccfv_ctx *c, const int32_t t[8])
{
    if (!c || !t) return fail(CCFV_EINVAL, "NULL argument");
    const bool post = post_flag(c);
    Tile v[4];
    for (int i = 0; i < 4; i++) {
        v[i] = Tile{t[2 * i], t[2 * i + 1]};
        Dir d = i < 2 ? DIR_X : DIR_Y;
        int n = d == DIR_X ? c->P.nx : c->P.ny, nl = d == DIR_X ? c->P.ny : c->P.nx;
        if (!tile_ok(v[i], n, nl, d, post && d == DIR_X, c->P.src_kind != CCFV_SRC_NONE,
                     c->P.bc))
            return fail(CCFV_EINVAL, "invalid tile shape for this grid");
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamSynchronize(c->stream2));
    HIPCHK(hipStreamSynchronize(c->stream3));
    clear_graphs(c);
    c->tiles_version++;
    c->tile[DIR_X][0] = v[0];
    c->tile[DIR_X][1] = v[1];
    c->tile[DIR_Y][0] = v[2];
    c->tile[DIR_Y][1] = v[3];
    return CCFV_OK;
}

int ccfv_autotune(ccfv_ctx *c, int32_t batch_fields, double h, int32_t out[8])
{
    if (!c) return fail(CCFV_EINVAL, "NULL context");
    if (batch_fields < 1 || batch_fields > 4096)
        return fail(CCFV_EINVAL, "batch_fields must be in [1, 4096]");
    if (!(h > 0.0)) return fail(CCFV_EINVAL, "h must be > 0");
    int rc = trim_tables(c);
    if (rc) return rc;
    for (int d = 0; d < 2; d++)
        if ((rc = autotune_dir(c, (Dir)d, batch_fields, h))) return rc;
    if (out) ccfv_get_tiles(c, out);
    return check_flag(c);
}

int ccfv_profile_enable(ccfv_ctx *c, int32_t enable)
{
    if (!c) return fail(CCFV_EINVAL, "NULL context");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamSynchronize(c->stream2));
    HIPCHK(hipStreamSynchronize(c->stream3));
    drain_records(c);
    memset(c->stats, 0, sizeof(c->stats));
    for (int k = 0; k < CCFV_KSTAT_COUNT; k++) c->stats[k].kind = k;
    c->prof = enable != 0;
    return CCFV_OK;
}

int ccfv_profile_read(ccfv_ctx *c, ccfv_kernel_stat *out)
{
    if (!c || !out) return fail(CCFV_EINVAL, "NULL argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamSynchronize(c->stream2));
    HIPCHK(hipStreamSynchronize(c->stream3));
    drain_records(c);
    for (int k = 0; k < CCFV_KSTAT_COUNT; k++) {
        out[k] = c->stats[k];
        out[k].kind = k;
    }
    return CCFV_OK;
}

int ccfv_time_sweep(ccfv_ctx *c, int32_t dir, int32_t fused, int32_t nf, double h,
                    int32_t launches, int32_t layout, int32_t tile_w, int32_t tile_nw,
                    int32_t graph, double *us)
{
    if (!c || !us) return fail(CCFV_EINVAL, "NULL argument");
    if ((dir != DIR_X && dir != DIR_Y) || (fused && dir != DIR_X) || nf < 1 || nf > 4096 ||
        launches < 1 || launches > 1 << 20 || !(h > 0.0) || (layout != 0 && layout != 1))
        return fail(CCFV_EINVAL, "dir 0/1, fused only for x, 1 <= nfields <= 4096, "
                                 "1 <= launches <= 2^20, h > 0, layout 0/1");
    const Dir d = (Dir)dir;
    const SolveTable *tb;
    int rc = trim_tables(c);
    if (rc || (rc = get_table(c, d, h, &tb))) return rc;
    const int n = d == DIR_X ? c->P.nx : c->P.ny, nl = d == DIR_X ? c->P.ny : c->P.nx;
    const Tile t = (tile_w || tile_nw) ? Tile{tile_w, tile_nw} : c->tile[d][nf > 1 ? 1 : 0];
    if (!tile_ok(t, n, nl, d, post_flag(c) && d == DIR_X, c->P.src_kind != CCFV_SRC_NONE,
                 c->P.bc))
        return fail(CCFV_EINVAL, "invalid tile shape for this grid");
    if (layout && (!tb->line_dev || c->P.nx % PANEL_W || c->P.ny % PANEL_W ||
                   (d == DIR_X && t.w != PANEL_W)))
        return fail(CCFV_EINVAL, "the panel layout needs the uniform-displacement sweep, "
                                 "grids of multiples of 8 and x shapes 8 lines wide");
    const size_t N = c->ncell;
    double *scratch = nullptr;
    HIPCHK(hipMalloc(&scratch, sizeof(double) * N * (size_t)nf));
    hipError_t e = hipSuccess;
    for (int f = 0; f < nf && e == hipSuccess; f++)
        e = launch_copy(c->state, scratch + f * N, N, c->stream);
    SweepArgs a = make_args(c, d, d == DIR_X ? (fused ? 3 : 1) : 2, *tb, h, 0.0, 0.0, 0, 0.0);
    a.in = scratch;
    a.out = scratch;
    a.in_panel = a.out_panel = layout;
    a.fstride = (long long)N;
    if (fused) {
        a.out2 = scratch;
        a.field0_2 = 0;
        a.t0_2 = 0.0;
        a.jdt2 = h;
    }
    const bool post = post_flag(c) && d == DIR_X;
    hipStream_t keep = c->cur;
    c->cur = c->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraphExec_t gx = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto issue = [&]() {
        hipError_t r = hipSuccess;
        for (int i = 0; i < launches && r == hipSuccess; i++)
            r = fused ? run_sweep_fused(c, post, EPI_NONE, a, nf, t)
                      : run_sweep(c, d, post, EPI_NONE, a, nf, t);
        return r;
    };
    if (e == hipSuccess && graph) {
        // one eager launch first (a warm start for the replays)
        e = fused ? run_sweep_fused(c, post, EPI_NONE, a, nf, t)
                  : run_sweep(c, d, post, EPI_NONE, a, nf, t);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
        if (e == hipSuccess) {
            hipError_t r = issue();
            hipGraph_t gr = nullptr;
            e = hipStreamEndCapture(c->stream, &gr);
            if (e == hipSuccess) e = r;
            if (e == hipSuccess) e = hipGraphInstantiate(&gx, gr, nullptr, nullptr, 0);
            if (gr) (void)hipGraphDestroy(gr);
        }
    }
    // one untimed pass, then the timed one
    for (int pass = 0; pass < 2 && e == hipSuccess; pass++) {
        if (pass == 1) e = hipEventRecord(e0, c->stream);
        if (e == hipSuccess) e = gx ? hipGraphLaunch(gx, c->stream) : issue();
        if (e == hipSuccess && pass == 1) e = hipEventRecord(e1, c->stream);
    }
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    c->cur = keep;
    (void)hipStreamSynchronize(c->stream);
    if (gx) (void)hipGraphExecDestroy(gx);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(scratch);
    if (e != hipSuccess) return hip_fail(e, "time sweep");
    *us = 1e3 * (double)ms / launches;
    return check_flag(c);
}

int ccfv_dispatch_floor(ccfv_ctx *c, int32_t dir, int32_t kind, int32_t launches,
                        int32_t graph, double *us)
{
    if (!c || !us) return fail(CCFV_EINVAL, "NULL argument");
    if (kind == 2) dir = DIR_X;
    if ((dir != DIR_X && dir != DIR_Y) || kind < 0 || kind > 2 || launches < 1 ||
        launches > 1 << 20)
        return fail(CCFV_EINVAL, "dir 0/1, kind 0/1/2, 1 <= launches <= 2^20");
    const Dir d = (Dir)dir;
    const int n = d == DIR_X ? c->P.nx : c->P.ny, nl = d == DIR_X ? c->P.ny : c->P.nx;
    const Tile t = c->tile[d][0];
    double *scratch = nullptr;
    HIPCHK(hipMalloc(&scratch, sizeof(double) * c->ncell));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraphExec_t gx = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess && kind == 2)
        e = hipMemcpyAsync(scratch, c->state, sizeof(double) * c->ncell,
                           hipMemcpyDeviceToDevice, c->stream);
    auto issue = [&]() {
        hipError_t r = hipSuccess;
        for (int i = 0; i < launches && r == hipSuccess; i++) {
            if (kind == 2) {
                // y copy then x copy, in place on the scratch field
                const Dir di = (i & 1) ? DIR_X : DIR_Y;
                const int ni = di == DIR_X ? c->P.nx : c->P.ny;
                const int li = di == DIR_X ? c->P.ny : c->P.nx;
                r = launch_floor(1, di, scratch, scratch, ni, li, c->P.ny, c->tile[di][0],
                                 c->stream);
            } else {
                r = launch_floor(kind, d, c->state, scratch, n, nl, c->P.ny, t, c->stream);
            }
        }
        return r;
    };
    if (e == hipSuccess && graph) {
        e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
        if (e == hipSuccess) {
            hipError_t r = issue();
            hipGraph_t gr = nullptr;
            e = hipStreamEndCapture(c->stream, &gr);
            if (e == hipSuccess) e = r;
            if (e == hipSuccess) e = hipGraphInstantiate(&gx, gr, nullptr, nullptr, 0);
            if (gr) (void)hipGraphDestroy(gr);
        }
    }
    // one untimed pass, then the timed one
    for (int pass = 0; pass < 2 && e == hipSuccess; pass++) {
        if (pass == 1) e = hipEventRecord(e0, c->stream);
        if (e == hipSuccess) e = gx ? hipGraphLaunch(gx, c->stream) : issue();
        if (e == hipSuccess && pass == 1) e = hipEventRecord(e1, c->stream);
    }
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipStreamSynchronize(c->stream);
    if (gx) (void)hipGraphExecDestroy(gx);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(scratch);
    if (e != hipSuccess) return hip_fail(e, "dispatch floor");
    *us = 1e3 * (double)ms / launches;
    return CCFV_OK;
}

int ccfv_sync(ccfv_ctx *c)
{
    if (!c) return fail(CCFV_EINVAL, "NULL context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return CCFV_OK;
}

}  // extern "C"

namespace ccfv {
namespace rt {

// relative L2 increment of iteration k from the per-slot sums: slots up to k
// are unchanged (their |U|^2 comes from earlier iterations)
double increment_of(const ccfv_ctx *c, int ns, int k, std::vector<double> &slot_u2)
{
    const double *h = c->h_sums + (size_t)k * 2 * (ns + 1);
    double num = 0.0, den = 0.0;
    for (int s = 1; s <= ns; s++) {
        if (s >= k + 1) {
            num += h[2 * s];
            slot_u2[s] = h[2 * s + 1];
        }
        den += slot_u2[s];
    }
    return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

struct ParaPlan {
    double t0, DT, dt;
    int J, ns, kmax;
    double tol;
};

// U[ns] after iteration k -> kept iterate k+1 (CCFV_OPT_KEEP_ITERATES), on
// the stream that ran the iteration's coarse chain
hipError_t keep_iterate(ccfv_ctx *c, int ns, int k, hipStream_t s)
{
    if (!c->keep_iterates || !c->iter_buf || k >= c->iter_cap) return hipSuccess;
    const size_t N = c->ncell;
    return launch_copy(c->U + (size_t)ns * N, c->iter_buf + (size_t)k * N, N, s);
}

// the steps of the Parareal schedule
StepDesc init_step(ccfv_ctx *c, const ParaPlan &pl, int n)
{
    // k = 0 coarse sweep: U[n+1] = G(U[n]), G cached
    const size_t N = c->ncell;
    return StepDesc{c->U + n * N, c->Wk, 1, pl.t0, pl.DT, n, 0.0, pl.DT, EPI_INIT,
                    c->G + n * N, nullptr, c->U + (n + 1) * N, nullptr};
}

StepDesc fine_step(ccfv_ctx *c, const ParaPlan &pl, int lo, int nf, int j)
{
    // step j of the fine propagators of slices lo..lo+nf-1, one batch
    const size_t N = c->ncell;
    const double *src = j == 0 ? c->U + lo * N : c->F + lo * N;
    return StepDesc{src, c->F + lo * N, nf, pl.t0, pl.DT, lo, j * pl.dt, pl.dt};
}

StepDesc chain_step(ccfv_ctx *c, const ParaPlan &pl, int n)
{
    // coarse step of slice n with the correction of slot n+1 fused
    const size_t N = c->ncell;
    const int nb = c->chain_blocks;
    return StepDesc{c->U + n * N, c->Wk, 1, pl.t0, pl.DT, n, 0.0, pl.DT, EPI_CORRECT,
                    c->G + n * N, c->F + n * N, c->U + (n + 1) * N,
                    c->partial + 2 * (size_t)nb * (n + 1)};
}

// Parareal with everything on one stream: k=0 coarse sweep, then per
// iteration the batched fine propagators of the slices that can still change
// and the coarse chain with the correction fused into its last sweep.
int parareal_sequential(ccfv_ctx *c, const ParaPlan &pl, ccfv_parareal_report &R,
                        std::vector<double> &slot_u2, bool &converged, int &kdone)
{
    const size_t N = c->ncell;
    const int ns = pl.ns, J = pl.J, nb = c->chain_blocks;
    int rc;
    c->cur = c->stream;
    if ((rc = run_steps(c, ns, [&](int n) { return init_step(c, pl, n); }))) return rc;
    R.coarse_steps_run = ns;
    int k;
    for (k = 0; k < pl.kmax; k++) {
        char label[48];
        snprintf(label, sizeof(label), "parareal iteration %d", k);
        TraceRange tr(label);
        const int nf = ns - k;
        if ((rc = run_steps(c, J, [&](int j) { return fine_step(c, pl, k, nf, j); })))
            return rc;
        R.fine_steps_run += (double)nf * J;
        HIPCHK(launch_correction(c->G + k * N, c->G + k * N, c->F + k * N,
                                 c->U + (k + 1) * N, N,
                                 c->partial + 2 * (size_t)nb * (k + 1), nb, c->stream));
        rc = run_steps(c, ns - k - 1, [&](int i) { return chain_step(c, pl, k + 1 + i); });
        if (rc) return rc;
        R.coarse_steps_run += ns - k - 1;
        HIPCHK(keep_iterate(c, ns, k, c->stream));
        double *sums = c->slot_sums + (size_t)k * 2 * (ns + 1);
        HIPCHK(launch_reduce_segments(c->partial, nb, k + 1, ns - k, sums, c->stream));
        HIPCHK(hipMemcpyAsync(c->h_sums + (size_t)k * 2 * (ns + 1), sums,
                              sizeof(double) * 2 * (ns + 1), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double inc = increment_of(c, ns, k, slot_u2);
        R.increment[k] = inc;
        if (!std::isfinite(inc)) {
            kdone = k + 1;
            return fail(CCFV_ENONFINITE, "Parareal iterate is not finite");
        }
        if (pl.tol > 0 && inc <= pl.tol) {
            converged = true;
            k++;
            break;
        }
    }
    kdone = k;
    return CCFV_OK;
}

int autotune_dir(ccfv_ctx *c, Dir dir, int batch_fields, double h)
{
    const SolveTable *tb;
    int rc = get_table(c, dir, h, &tb);
    if (rc) return rc;
    const bool post = post_flag(c);
    const size_t N = c->ncell;
    double *scratch = nullptr;
    HIPCHK(hipMalloc(&scratch, sizeof(double) * N * (size_t)batch_fields));
    for (int f = 0; f < batch_fields; f++) {
        hipError_t e = launch_copy(c->state, scratch + f * N, N, c->stream);
        if (e != hipSuccess) {
            (void)hipFree(scratch);
            return hip_fail(e, "autotune copy");
        }
    }
    static const int ws[] = {1, 2, 4, 8, 16, 32};
    static const int nws[] = {1, 2, 4};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipStream_t keep = c->cur;
    c->cur = c->stream;
    int status = CCFV_OK;
    const int n = dir == DIR_X ? c->P.nx : c->P.ny;
    const int nl = dir == DIR_X ? c->P.ny : c->P.nx;
    for (int b = 0; b < 2 && !status; b++) {
        const int nf = b ? batch_fields : 1;
        const int reps = b ? 3 : 20;
        // x: the launch the solves issue, the last x sub-step of a step
        // fused with the first of the next (one sub-step when fusion is off)
        const bool fused = dir == DIR_X && c->fuse;
        SweepArgs a = make_args(c, dir, dir == DIR_X ? (fused ? 3 : 1) : 2, *tb, h, 0.0, 0.0, 0,
                                0.0);
        a.in = scratch;
        a.out = scratch;
        a.fstride = (long long)N;
        if (fused) {
            a.out2 = scratch;
            a.field0_2 = 0;
            a.t0_2 = 0.0;
            a.jdt2 = h;
        }
        auto sweep = [&](Tile t) {
            return fused ? run_sweep_fused(c, post, EPI_NONE, a, nf, t)
                         : run_sweep(c, dir, false, EPI_NONE, a, nf, t);
        };
        float best = 1e30f;
        Tile bt = c->tile[dir][b];
        for (int w : ws)
            for (int nw : nws) {
                Tile t{w, nw};
                if (!tile_ok(t, n, nl, dir, post && dir == DIR_X, c->P.src_kind != CCFV_SRC_NONE,
                             c->P.bc))
                    continue;
                float tbest = 1e30f;
                for (int trial = 0; trial < 2; trial++) {
                    hipError_t e = sweep(t);
                    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
                    for (int r = 0; r < reps && e == hipSuccess; r++) e = sweep(t);
                    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
                    if (e == hipSuccess) e = hipEventSynchronize(e1);
                    float ms = 0.f;
                    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                    if (e != hipSuccess) {
                        status = hip_fail(e, "autotune sweep");
                        break;
                    }
                    tbest = std::min(tbest, ms / reps);
                }
                if (status) break;
                if (tbest < best) {
                    best = tbest;
                    bt = t;
                }
            }
        c->tile[dir][b] = bt;
    }
    c->cur = keep;
    clear_graphs(c);
    c->tiles_version++;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(scratch);
    return status;
}

hipEvent_t sync_event(ccfv_ctx *c, size_t i)
{
    while (c->sync_ev.size() <= i) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return nullptr;
        c->sync_ev.push_back(e);
    }
    return c->sync_ev[i];
}

// Pipelined Parareal.  Slices form groups of pipe_group.  The coarse chain
// runs on the context's stream, the fine propagators on a second stream:
// the fine propagators of iteration k+1 for a group start as soon as the
// chain of iteration k has written that group's slots (and consumed its F),
// so the latency-bound chain overlaps the bandwidth-bound fine batches.
// Fine work of iteration k+1 is issued before iteration k's increment is
// known; it only writes F, so it is discarded if k converged.  The kernels
// and their inputs are those of parareal_sequential: results are identical.
//
// With graphs on, each phase {chain(k) || fine(k+1)} (and the first phase
// {k=0 coarse sweep || fine(0)}) is captured once as one hipGraph with a
// fork/join between the streams and replayed on later solves with the same
// plan; the dependencies of a phase on the previous one follow from the
// graph launches' order on the chain stream.
int parareal_pipelined(ccfv_ctx *c, const ParaPlan &pl, ccfv_parareal_report &R,
                       std::vector<double> &slot_u2, bool &converged, int &kdone)
{
    const size_t N = c->ncell;
    const int ns = pl.ns, J = pl.J, nb = c->chain_blocks;
    const int gs = std::min(c->pipe_group, ns);
    const int G = (ns + gs - 1) / gs;
    const bool graphs = c->graphs && !c->prof;
    auto g0 = [&](int g) { return g * gs; };
    auto g1 = [&](int g) { return std::min(ns, (g + 1) * gs); };
    auto grp = [&](int n) { return n / gs; };
    // events: [par][kind][g], kind 0 = chain passed g, 1 = fine read U of g,
    // 2 = fine done with g; plus fork / join events
    auto ev = [&](int par, int kind, int g) { return c->sync_ev[(size_t)((par * 3 + kind) * G + g)]; };
    for (int i = 0; i < 6 * G + 2; i++)
        if (!sync_event(c, (size_t)i)) return fail(CCFV_EDEVICE, "event creation failed");
    hipEvent_t fork = c->sync_ev[(size_t)6 * G], join = c->sync_ev[(size_t)6 * G + 1];
    hipStream_t sC = c->stream, sF = c->stream2;
    int rc;

    // groups [gb, ge) of the fine propagators of iteration k
    auto enqueue_fine = [&](int k, int gb = 0, int ge = -1) -> int {
        const int par = k & 1;
        if (ge < 0) ge = G;
        c->cur = sF;
        for (int g = gb; g < ge; g++) {
            const int lo = std::max(g0(g), k);
            if (lo >= g1(g)) continue;
            HIPCHK(hipStreamWaitEvent(sF, ev(par, 0, g), 0));
            int r2 = run_steps(
                c, J, [&](int j) { return fine_step(c, pl, lo, g1(g) - lo, j); }, {},
                [&](int j) -> int {
                    // U of the group has been read once step 0 is done
                    if (j == 0 && !graphs) HIPCHK(hipEventRecord(ev(par, 1, g), sF));
                    return CCFV_OK;
                });
            if (r2) return r2;
            if (!graphs) HIPCHK(hipEventRecord(ev(par, 2, g), sF));
        }
        c->cur = sC;
        return CCFV_OK;
    };
    auto fine_count = [&](int k, int gb = 0, int ge = -1) {
        double s = 0;
        if (ge < 0) ge = G;
        for (int g = gb; g < ge; g++) {
            const int lo = std::max(g0(g), k);
            if (lo < g1(g)) s += (double)(g1(g) - lo) * J;
        }
        return s;
    };
    // k = 0 coarse sweep; group g is ready for fine(0) after step g1-1
    auto enqueue_init = [&]() -> int {
        c->cur = sC;
        return run_steps(
            c, ns, [&](int n) { return init_step(c, pl, n); }, {},
            [&](int n) -> int {
                if (n == g1(grp(n)) - 1) HIPCHK(hipEventRecord(ev(0, 0, grp(n)), sC));
                return CCFV_OK;
            });
    };
    auto enqueue_chain = [&](int k) -> int {
        const int par = k & 1, npar = (k + 1) & 1;
        c->cur = sC;
        // slot k+1: needs F[k], and fine(k) of the next group must have read
        // U[k+1] before it is overwritten (implied by the phase order when
        // graphs are on)
        if (!graphs) {
            HIPCHK(hipStreamWaitEvent(sC, ev(par, 2, grp(k)), 0));
            if (k + 1 < ns) HIPCHK(hipStreamWaitEvent(sC, ev(par, 1, grp(k + 1)), 0));
        }
        HIPCHK(launch_correction(c->G + k * N, c->G + k * N, c->F + k * N,
                                 c->U + (k + 1) * N, N,
                                 c->partial + 2 * (size_t)nb * (k + 1), nb, sC));
        // waits come before the launch that starts step n (which also ends
        // step n-1), records after the launch that ends step n
        int r2 = run_steps(
            c, ns - k - 1, [&](int i) { return chain_step(c, pl, k + 1 + i); },
            [&](int i) -> int {
                const int n = k + 1 + i, g = grp(n);
                if (!graphs && (n == k + 1 || n == g0(g))) {
                    HIPCHK(hipStreamWaitEvent(sC, ev(par, 2, g), 0));
                    if (g + 1 < G) HIPCHK(hipStreamWaitEvent(sC, ev(par, 1, g + 1), 0));
                }
                return CCFV_OK;
            },
            [&](int i) -> int {
                const int n = k + 1 + i, g = grp(n);
                if (n == g1(g) - 1) HIPCHK(hipEventRecord(ev(npar, 0, g), sC));
                return CCFV_OK;
            });
        if (r2) return r2;
        HIPCHK(keep_iterate(c, ns, k, sC));
        double *sums = c->slot_sums + (size_t)k * 2 * (ns + 1);
        HIPCHK(launch_reduce_segments(c->partial, nb, k + 1, ns - k, sums, sC));
        HIPCHK(hipMemcpyAsync(c->h_sums + (size_t)k * 2 * (ns + 1), sums,
                              sizeof(double) * 2 * (ns + 1), hipMemcpyDeviceToHost, sC));
        return CCFV_OK;
    };
    // one phase: captured and cached as a graph, or issued directly
    auto phase = [&](int k, auto &&body) -> int {
        if (!graphs) return body();
        char key[256];
        snprintf(key, sizeof(key), "pa|%a|%a|%a|%d|%d|%d|%d|%d|%d", pl.t0, pl.DT, pl.dt,
                 J, ns, k, (k + 1 < pl.kmax) ? 1 : 0, gs, c->tiles_version);
        auto it = c->gcache.find(key);
        if (it == c->gcache.end()) {
            HIPCHK(hipStreamBeginCapture(sC, hipStreamCaptureModeRelaxed));
            hipError_t e = hipEventRecord(fork, sC);
            if (e == hipSuccess) e = hipStreamWaitEvent(sF, fork, 0);
            int r2 = e == hipSuccess ? body() : CCFV_EDEVICE;
            if (r2 == CCFV_OK) {
                e = hipEventRecord(join, sF);
                if (e == hipSuccess) e = hipStreamWaitEvent(sC, join, 0);
            }
            hipGraph_t gr = nullptr;
            hipError_t e2 = hipStreamEndCapture(sC, &gr);
            if (r2 || e != hipSuccess || e2 != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                if (r2) return r2;
                return hip_fail(e != hipSuccess ? e : e2, "Parareal graph capture");
            }
            hipGraphExec_t x = nullptr;
            e = hipGraphInstantiate(&x, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
            it = c->gcache.emplace(key, x).first;
        }
        HIPCHK(hipGraphLaunch(it->second, sC));
        return CCFV_OK;
    };

    if (!graphs) {
        // the fine stream starts after everything queued on the chain stream
        HIPCHK(hipEventRecord(join, sC));
        HIPCHK(hipStreamWaitEvent(sF, join, 0));
    }
    rc = phase(-1, [&]() -> int {
        int r2 = enqueue_init();
        if (r2 == CCFV_OK && pl.kmax > 0) r2 = enqueue_fine(0);
        return r2;
    });
    if (rc) return rc;
    R.coarse_steps_run = ns;
    if (pl.kmax > 0) R.fine_steps_run += fine_count(0);

    // With a tolerance (and no graphs) the last group of iteration k+1's
    // fine propagators is enqueued only once the host has read iteration
    // k's increment and found it above tol: that group can start only when
    // the chain of k has passed it (its end) anyway, and a converged solve
    // no longer runs it (half of an iteration's fine work at 64 slices of
    // 32-slice groups).  The other groups start behind the chain as before.
    const bool defer = pl.tol > 0 && !graphs;
    const int gq = defer ? G - 1 : G;
    int k;
    bool pending = false;  // increments still to be read (tol <= 0)
    for (k = 0; k < pl.kmax; k++) {
        char label[48];
        snprintf(label, sizeof(label), "parareal iteration %d", k);
        TraceRange tr(label);
        rc = phase(k, [&]() -> int {
            int r2 = enqueue_chain(k);
            if (r2 == CCFV_OK && k + 1 < pl.kmax) r2 = enqueue_fine(k + 1, 0, gq);
            return r2;
        });
        if (rc) return rc;
        R.coarse_steps_run += ns - k - 1;
        if (k + 1 < pl.kmax) R.fine_steps_run += fine_count(k + 1, 0, gq);
        if (pl.tol > 0) {
            HIPCHK(hipStreamSynchronize(sC));
            double inc = increment_of(c, ns, k, slot_u2);
            R.increment[k] = inc;
            if (!std::isfinite(inc)) {
                (void)hipStreamSynchronize(sF);
                kdone = k + 1;
                return fail(CCFV_ENONFINITE, "Parareal iterate is not finite");
            }
            if (inc <= pl.tol) {
                converged = true;
                k++;
                break;
            }
            if (defer && k + 1 < pl.kmax) {
                TraceRange td("deferred fine group");
                if ((rc = enqueue_fine(k + 1, gq, G))) return rc;
                R.fine_steps_run += fine_count(k + 1, gq, G);
            }
        } else {
            pending = true;
        }
    }
    kdone = k;
    if (!graphs) {
        // join the fine stream (speculative work included) into the chain stream
        HIPCHK(hipEventRecord(join, sF));
        HIPCHK(hipStreamWaitEvent(sC, join, 0));
    }
    if (pending) {
        HIPCHK(hipStreamSynchronize(sC));
        for (int i = 0; i < kdone; i++) {
            R.increment[i] = increment_of(c, ns, i, slot_u2);
            if (!std::isfinite(R.increment[i]))
                return fail(CCFV_ENONFINITE, "Parareal iterate is not finite");
        }
    }
    return CCFV_OK;
}

// Overlapped Parareal (default when not capturing graphs).  As the
// pipelined schedule, the fine propagators of iteration k+1 run on a second
// stream group by group once the chain of iteration k has passed the group;
// in addition the coarse chains of consecutive iterations run on two streams
// (even k on the context's stream, odd k on a third) with their own work
// field and increment partials, so the chain of iteration k+1 starts as soon
// as its first group's fine propagators are done instead of after the whole
// chain of iteration k.  Events per iteration parity and group order every
// read-after-write and write-after-read on U, G and F:
//   pass[p][g]  chain(k-1) (or the k=0 coarse sweep) has written U^k of
//               group g and read its F and G  (p = k & 1)
//   read[p][g]  fine(k) has read U^k of group g
//   done[p][g]  fine(k) has written F of group g
// chain(k+1) step n also overwrites U[n+1] and G[n], which chain(k) reads
// at steps n+1 and n; the waits on done[][g] and read[][g+1] order it after
// chain(k) has passed group g+1.  The chain of iteration k+1 is issued
// before k's increment is known; each chain copies its U[ns] to a copy of
// its parity, and the accepted iteration's copy becomes U[ns] at the end.
int parareal_overlapped(ccfv_ctx *c, const ParaPlan &pl, ccfv_parareal_report &R,
                        std::vector<double> &slot_u2, bool &converged, int &kdone)
{
    const size_t N = c->ncell;
    const int ns = pl.ns, J = pl.J, nb = c->chain_blocks;
    const int gs = std::min(c->pipe_group, ns);
    const int G = (ns + gs - 1) / gs;
    auto g0 = [&](int g) { return g * gs; };
    auto g1 = [&](int g) { return std::min(ns, (g + 1) * gs); };
    auto grp = [&](int n) { return n / gs; };
    // events: [par][kind][g] (kind 0 pass, 1 read, 2 done) + end[par] + join
    auto ev = [&](int par, int kind, int g) { return c->sync_ev[(size_t)((par * 3 + kind) * G + g)]; };
    for (int i = 0; i < 6 * G + 3; i++)
        if (!sync_event(c, (size_t)i)) return fail(CCFV_EDEVICE, "event creation failed");
    hipEvent_t end_ev[2] = {c->sync_ev[(size_t)6 * G], c->sync_ev[(size_t)6 * G + 1]};
    hipEvent_t join = c->sync_ev[(size_t)6 * G + 2];
    hipStream_t sCh[2] = {c->stream, c->stream3}, sF = c->stream2;
    double *wk[2] = {c->Wk, c->Wk2}, *part[2] = {c->partial, c->partial2};
    double *res[2] = {c->res2, c->res2 + N};
    int rc;

    // groups [gb, ge) of the fine propagators of iteration k
    auto enqueue_fine = [&](int k, int gb = 0, int ge = -1) -> int {
        const int par = k & 1;
        if (ge < 0) ge = G;
        c->cur = sF;
        for (int g = gb; g < ge; g++) {
            const int lo = std::max(g0(g), k);
            if (lo >= g1(g)) continue;
            HIPCHK(hipStreamWaitEvent(sF, ev(par, 0, g), 0));
            int r2 = run_steps(
                c, J, [&](int j) { return fine_step(c, pl, lo, g1(g) - lo, j); }, {},
                [&](int j) -> int {
                    if (j == 0) HIPCHK(hipEventRecord(ev(par, 1, g), sF));
                    return CCFV_OK;
                });
            if (r2) return r2;
            HIPCHK(hipEventRecord(ev(par, 2, g), sF));
        }
        return CCFV_OK;
    };
    auto fine_count = [&](int k) {
        double t = 0;
        for (int g = 0; g < G; g++) {
            const int lo = std::max(g0(g), k);
            if (lo < g1(g)) t += (double)(g1(g) - lo) * J;
        }
        return t;
    };
    // chain(k) on sCh[k & 1] with that parity's work field and partials
    auto enqueue_chain = [&](int k) -> int {
        const int par = k & 1, npar = (k + 1) & 1;
        const hipStream_t sC = sCh[par];
        c->cur = sC;
        // slot k+1 needs F[k]; it overwrites U[k+1], which fine(k) of the
        // next group reads
        HIPCHK(hipStreamWaitEvent(sC, ev(par, 2, grp(k)), 0));
        if (k + 1 < ns) HIPCHK(hipStreamWaitEvent(sC, ev(par, 1, grp(k + 1)), 0));
        HIPCHK(launch_correction(c->G + k * N, c->G + k * N, c->F + k * N,
                                 c->U + (k + 1) * N, N,
                                 part[par] + 2 * (size_t)nb * (k + 1), nb, sC));
        auto step = [&](int i) {
            const int n = k + 1 + i;
            return StepDesc{c->U + n * N, wk[par], 1, pl.t0, pl.DT, n, 0.0, pl.DT,
                            EPI_CORRECT, c->G + n * N, c->F + n * N, c->U + (n + 1) * N,
                            part[par] + 2 * (size_t)nb * (n + 1)};
        };
        int r2 = run_steps(
            c, ns - k - 1, step,
            [&](int i) -> int {
                const int n = k + 1 + i, g = grp(n);
                if (n == k + 1 || n == g0(g)) {
                    HIPCHK(hipStreamWaitEvent(sC, ev(par, 2, g), 0));
                    if (g + 1 < G) HIPCHK(hipStreamWaitEvent(sC, ev(par, 1, g + 1), 0));
                }
                return CCFV_OK;
            },
            [&](int i) -> int {
                const int n = k + 1 + i, g = grp(n);
                // the last group's pass follows the copy of U[ns] below
                if (n == g1(g) - 1 && g + 1 < G) HIPCHK(hipEventRecord(ev(npar, 0, g), sC));
                return CCFV_OK;
            });
        if (r2) return r2;
        HIPCHK(launch_copy(c->U + (size_t)ns * N, res[par], N, sC));
        HIPCHK(keep_iterate(c, ns, k, sC));
        if (k + 1 < ns) HIPCHK(hipEventRecord(ev(npar, 0, G - 1), sC));
        double *sums = c->slot_sums + (size_t)k * 2 * (ns + 1);
        HIPCHK(launch_reduce_segments(part[par], nb, k + 1, ns - k, sums, sC));
        HIPCHK(hipMemcpyAsync(c->h_sums + (size_t)k * 2 * (ns + 1), sums,
                              sizeof(double) * 2 * (ns + 1), hipMemcpyDeviceToHost, sC));
        HIPCHK(hipEventRecord(end_ev[par], sC));
        return CCFV_OK;
    };
    // one finished iteration's increment and the stop decision
    auto check = [&](int k, bool &stop) -> int {
        HIPCHK(hipEventSynchronize(end_ev[k & 1]));
        const double inc = increment_of(c, ns, k, slot_u2);
        R.increment[k] = inc;
        stop = false;
        if (!std::isfinite(inc)) return fail(CCFV_ENONFINITE, "Parareal iterate is not finite");
        if (pl.tol > 0 && inc <= pl.tol) {
            converged = true;
            stop = true;
        }
        return CCFV_OK;
    };

    // everything queued on the context's stream before (the initial copy)
    // comes first on every stream
    HIPCHK(hipEventRecord(join, c->stream));
    HIPCHK(hipStreamWaitEvent(sF, join, 0));
    HIPCHK(hipStreamWaitEvent(c->stream3, join, 0));
    // k = 0 coarse sweep on the even chain stream
    c->cur = sCh[0];
    rc = run_steps(
        c, ns, [&](int n) { return init_step(c, pl, n); }, {},
        [&](int n) -> int {
            if (n == g1(grp(n)) - 1) HIPCHK(hipEventRecord(ev(0, 0, grp(n)), sCh[0]));
            return CCFV_OK;
        });
    if (rc) return rc;
    R.coarse_steps_run = ns;
    if (pl.kmax > 0) {
        if ((rc = enqueue_fine(0))) return rc;
        R.fine_steps_run += fine_count(0);
    }
    int accepted = -1;  // last iteration whose increment has been read
    bool stop = false;
    int k;
    for (k = 0; k < pl.kmax && !stop; k++) {
        if ((rc = enqueue_chain(k))) return rc;
        R.coarse_steps_run += ns - k - 1;
        if (k + 1 < pl.kmax) {
            if ((rc = enqueue_fine(k + 1))) return rc;
            R.fine_steps_run += fine_count(k + 1);
        }
        // read iteration k-1 while chain(k) runs
        if (pl.tol > 0 && k >= 1) {
            if ((rc = check(k - 1, stop))) {
                kdone = k;
                return rc;
            }
            accepted = k - 1;
        }
    }
    int last = stop ? accepted : k - 1;  // the iteration whose U[ns] is the result
    if (!stop && last >= 0 && pl.tol > 0) {
        if ((rc = check(last, stop))) {
            kdone = last + 1;
            return rc;
        }
    }
    kdone = last + 1;
    // join every stream into the context's stream (speculative chains too)
    for (hipStream_t s : {sF, c->stream3}) {
        HIPCHK(hipEventRecord(join, s));
        HIPCHK(hipStreamWaitEvent(c->stream, join, 0));
    }
    if (last >= 0)
        HIPCHK(launch_copy(res[last & 1], c->U + (size_t)ns * N, N, c->stream));
    c->cur = c->stream;
    if (pl.tol <= 0) {
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < kdone; i++) {
            R.increment[i] = increment_of(c, ns, i, slot_u2);
            if (!std::isfinite(R.increment[i]))
                return fail(CCFV_ENONFINITE, "Parareal iterate is not finite");
        }
    }
    return CCFV_OK;
}

}  // namespace rt
}  // namespace ccfv

extern "C" {

int ccfv_parareal_solve(ccfv_ctx *c, double t0, double t1, int32_t nt_fine,
                        int32_t ns, int32_t kmax, double tol,
                        ccfv_parareal_report *rep)
{
    if (!c) return fail(CCFV_EINVAL, "NULL context");
    if (ns < 1 || nt_fine < ns || nt_fine % ns)
        return fail(CCFV_EINVAL, "nt_fine must be a positive multiple of n_slices");
    if (kmax < 0 || kmax > CCFV_MAX_ITERS)
        return fail(CCFV_EINVAL, "kmax must be in [0, 64]");
    if (!(t1 > t0)) return fail(CCFV_EINVAL, "need t1 > t0");
    ccfv::rt::TraceRange tr("ccfv_parareal_solve");
    int rc = trim_tables(c);
    if (rc) return rc;
    rc = ensure_para(c, ns);
    if (rc) return rc;
    const size_t N = c->ncell;
    ParaPlan pl;
    pl.t0 = t0;
    pl.DT = (t1 - t0) / ns;       // :1251
    pl.J = nt_fine / ns;          // :1252
    pl.dt = pl.DT / (double)pl.J; // :979
    pl.ns = ns;
    // iteration k fixes slot k+1; past n_slices iterations nothing changes
    pl.kmax = std::min((int)kmax, (int)ns);
    pl.tol = tol;
    // tables up front (host-side work stays out of the timed region); the
    // serial finish steps by (t1 - t0) / nt_fine as ccfv_plain_solve
    const bool finish = c->serial_finish && tol > 0;
    const double dt_plain = (t1 - t0) / nt_fine;
    const SolveTable *tt;
    if ((rc = get_table(c, DIR_X, pl.DT, &tt)) || (rc = get_table(c, DIR_Y, pl.DT, &tt)) ||
        (rc = get_table(c, DIR_X, pl.dt, &tt)) || (rc = get_table(c, DIR_Y, pl.dt, &tt)) ||
        (finish && ((rc = get_table(c, DIR_X, dt_plain, &tt)) ||
                    (rc = get_table(c, DIR_Y, dt_plain, &tt)))) ||
        (rc = ensure_panel(c, ns, pl.dt)))
        return rc;
    c->iter_valid = 0;
    if (c->keep_iterates && c->iter_cap < pl.kmax) {
        HIPCHK(hipStreamSynchronize(c->stream));
        (void)hipFree(c->iter_buf);
        c->iter_buf = nullptr;
        c->iter_cap = 0;
        clear_graphs(c);  // captured phases hold the old buffer's address
        if (hipMalloc(&c->iter_buf, sizeof(double) * N * (size_t)pl.kmax) != hipSuccess) {
            (void)hipGetLastError();
            return fail(CCFV_ENOMEM, "buffer of kept Parareal iterates");
        }
        c->iter_cap = pl.kmax;
    }
    ccfv_parareal_report R;
    memset(&R, 0, sizeof(R));
    R.n_slices = ns;
    R.j_fine = pl.J;
    std::vector<double> slot_u2(ns + 1, 0.0);
    bool converged = false;
    int kdone = 0;

    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(launch_copy(c->state, c->U, N, c->stream));
    if (!c->pipeline) rc = parareal_sequential(c, pl, R, slot_u2, converged, kdone);
    else if (c->pipeline == 2 && !(c->graphs && !c->prof))
        rc = parareal_overlapped(c, pl, R, slot_u2, converged, kdone);
    else rc = parareal_pipelined(c, pl, R, slot_u2, converged, kdone);
    c->cur = c->stream;
    if (rc == CCFV_OK && kdone == ns) converged = true;  // every slice exact
    R.iterations = kdone;
    if (c->keep_iterates) c->iter_valid = std::min(kdone, c->iter_cap);
    R.converged = converged ? 1 : 0;
    if (rc) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->stream2);
        (void)hipStreamSynchronize(c->stream3);
        if (rep) *rep = R;
        return rc;
    }
    if (finish && !converged) {
        // kmax reached above tol: the serial fine solve from the initial
        // state, still in U[0] (the steps of ccfv_plain_solve)
        HIPCHK(launch_copy(c->U, c->state, N, c->stream));
        rc = run_steps(c, nt_fine, [&](int n) {
            return StepDesc{c->state, c->state, 1, t0 + n * dt_plain, 0.0, 0, 0.0, dt_plain};
        });
        if (rc) {
            (void)hipStreamSynchronize(c->stream);
            if (rep) *rep = R;
            return rc;
        }
        R.fine_steps_run += nt_fine;
        R.serial_finish = 1;
        R.converged = 1;
        converged = true;
    } else {
        HIPCHK(launch_copy(c->U + (size_t)ns * N, c->state, N, c->stream));
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    rc = check_flag(c);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    R.wall_ms = ms;
    if (rep) *rep = R;
    if (rc) return rc;
    if (tol > 0 && !converged)
        return fail(CCFV_ENOTCONV, "Parareal did not reach the tolerance within kmax");
    return CCFV_OK;
}

int ccfv_set_option(ccfv_ctx *c, int32_t key, int64_t value)
{
    if (!c) return fail(CCFV_EINVAL, "NULL context");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamSynchronize(c->stream2));
    HIPCHK(hipStreamSynchronize(c->stream3));
    switch (key) {
    case CCFV_OPT_PIPELINE:
        if (value < 0 || value > 2) return fail(CCFV_EINVAL, "pipeline must be 0, 1 or 2");
        c->pipeline = (int)value;
        return CCFV_OK;
    case CCFV_OPT_PIPE_GROUP:
        if (value < 1 || value > 4096) return fail(CCFV_EINVAL, "group size out of range");
        c->pipe_group = (int)value;
        return CCFV_OK;
    case CCFV_OPT_GRAPHS:
        c->graphs = value != 0;
        clear_graphs(c);
        return CCFV_OK;
    case CCFV_OPT_PLAIN_GRAPHS:
        c->plain_graphs = value != 0;
        clear_graphs(c);
        return CCFV_OK;
    case CCFV_OPT_FUSE:
        c->fuse = value != 0;
        clear_graphs(c);
        return CCFV_OK;
    case CCFV_OPT_PANEL:
        c->panel_on = value != 0;
        clear_graphs(c);
        return CCFV_OK;
    case CCFV_OPT_SERIAL_FINISH:
        if (value != 0 && value != 1) return fail(CCFV_EINVAL, "serial finish must be 0 or 1");
        c->serial_finish = value != 0;
        return CCFV_OK;
    case CCFV_OPT_KEEP_ITERATES:
        if (value != 0 && value != 1) return fail(CCFV_EINVAL, "keep iterates must be 0 or 1");
        c->keep_iterates = value != 0;
        c->iter_valid = 0;
        clear_graphs(c);
        return CCFV_OK;
    }
    return fail(CCFV_EINVAL, "unknown option");
}

int ccfv_get_option(ccfv_ctx *c, int32_t key, int64_t *value)
{
    if (!c || !value) return fail(CCFV_EINVAL, "NULL argument");
    switch (key) {
    case CCFV_OPT_PIPELINE: *value = c->pipeline; return CCFV_OK;
    case CCFV_OPT_PIPE_GROUP: *value = c->pipe_group; return CCFV_OK;
    case CCFV_OPT_GRAPHS: *value = c->graphs ? 1 : 0; return CCFV_OK;
    case CCFV_OPT_PLAIN_GRAPHS: *value = c->plain_graphs ? 1 : 0; return CCFV_OK;
    case CCFV_OPT_FUSE: *value = c->fuse ? 1 : 0; return CCFV_OK;
    case CCFV_OPT_PANEL: *value = c->panel_on ? 1 : 0; return CCFV_OK;
    case CCFV_OPT_SERIAL_FINISH: *value = c->serial_finish ? 1 : 0; return CCFV_OK;
    case CCFV_OPT_KEEP_ITERATES: *value = c->keep_iterates ? 1 : 0; return CCFV_OK;
    }
    return fail(CCFV_EINVAL, "unknown option");
}

}  // extern "C"
