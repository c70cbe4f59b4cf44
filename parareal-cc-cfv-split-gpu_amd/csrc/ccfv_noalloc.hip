// ccfv_noalloc.hip — ccfv_plain_solve_noalloc: the plain serial solve of
// ccfv_plain_solve for callers that cannot allocate device memory (code
// under stream capture, tight serving loops, harnesses that forbid
// hipMalloc).  The caller passes the field; everything else the sweeps need
// beyond it — the crossed-feet flag, both directions' solve tables, the
// per-line shift and weights of the uniform-offset sweep, the swirl's face
// table — lives in one static device array per GPU, built there by kernels
// from the same host/device functions get_table uses (bitwise the same
// tables).  No device memory is allocated or freed.
//
// Every call on a GPU uses that GPU's array and leaves its contents behind:
// calls that overlap on one GPU (two host threads, or two streams in
// flight) race on it.
#include "ccfv_ctx.h"
#include "ccfv_tables.h"

#include <cmath>
#include <mutex>

using namespace ccfv;
using namespace ccfv::rt;

namespace {

// 64 Ki doubles (512 KiB): the layout below for grids up to 2048 x 2048
constexpr size_t SCRATCH_DOUBLES = (size_t)1 << 16;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_ptr[MAX_DEVICES] = {};

// this device's address of `scratch`, looked up once per device (with that
// device current); host code uses it, never the array's name
int scratch_ptr(double **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(CCFV_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_ptr[dev];
    return CCFV_OK;
}

__global__ void clear_flag_kernel(int *flag)
{
    if (threadIdx.x == 0) *flag = 0;
}

__global__ void thomas_kernel(int n, double d, double theta, double nu, double h, int per,
                              int ch, double *tab, double *work)
{
    if (threadIdx.x == 0) (void)thomas_fill(n, d, theta, nu, h, per != 0, ch, tab, work);
}

__global__ void lines_kernel(ccfv_problem P, int dir, double h, double coef, int nl,
                             double *lines)
{
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= nl) return;
    double S, w[6];
    if (line_fill(P, dir, h, coef, L, &S, w)) {
        lines[L] = S;
        for (int k = 0; k < 6; k++) lines[nl + 6 * (size_t)L + k] = w[k];
    }
}

// offsets (doubles) of the regions, each 256-byte aligned
struct Layout {
    size_t flag = 0, tab[2] = {0, 0}, work = 0, lines[2] = {0, 0}, trig = 0, end = 0;
};

size_t up(size_t v) { return (v + 31) & ~(size_t)31; }

Layout layout_of(const ccfv_problem &P, bool fast)
{
    Layout L;
    size_t o = 32;  // the flag's 256 bytes
    L.tab[0] = o;
    o = up(o + (size_t)tab_doubles(P.nx));
    L.tab[1] = o;
    o = up(o + (size_t)tab_doubles(P.ny));
    L.work = o;
    o = up(o + 2 * (size_t)std::max(P.nx, P.ny));
    if (fast) {
        L.lines[0] = o;
        o = up(o + 7 * (size_t)P.ny);
        L.lines[1] = o;
        o = up(o + 7 * (size_t)P.nx);
    }
    if (P.vel_kind == CCFV_VEL_SWIRL) {
        L.trig = o;
        o = up(o + 2 * (size_t)(P.nx + 1) + 2 * (size_t)(P.ny + 1));
    }
    L.end = o;
    return L;
}

// the host's share of a table: its scalars and scan steps (the device
// builds the arrays with the same function), and whether every line takes
// the uniform-offset sweep; fixed buffers, no allocation per call
thread_local double h_tab[tab_doubles(2048)], h_work[2 * 2048];

bool host_table(const ccfv_problem &P, Dir dir, double h, SolveTable &T)
{
    const int n = dir == DIR_X ? P.nx : P.ny;
    const int nl = dir == DIR_X ? P.ny : P.nx;
    const TabScalars S =
        thomas_fill(n, cell_width(P, dir), dir == DIR_X ? 0.25 : 0.5, P.nu, h,
                    P.bc == CCFV_BC_PERIODIC, chunk_for(n), h_tab, h_work);
    T.ca = S.ca;
    T.cb = S.cb;
    T.coef = S.coef;
    T.vl = S.vl;
    T.diag0 = S.diag0;
    T.smden = S.smden;
    scan_steps(h_tab, n, T);
    if (!uniform_problem(P) || getenv("CCFV_NO_FAST")) return false;
    for (int L = 0; L < nl; L++) {
        double s, w[6];
        if (!line_fill(P, (int)dir, h, T.coef, L, &s, w)) return false;
    }
    return true;
}

struct PlainSteps {
    double *u;
    double t0, dt;
};

}  // namespace

extern "C" {

int ccfv_plain_solve_noalloc(const ccfv_problem *p, double *u, double t0, double t1,
                             int32_t nt, void *stream)
{
    int rc = validate_problem(p);
    if (rc) return rc;
    if (!u) return fail(CCFV_EINVAL, "NULL field");
    if (nt < 1) return fail(CCFV_EINVAL, "nt must be >= 1");
    if (!(t1 > t0) || !std::isfinite(t0) || !std::isfinite(t1))
        return fail(CCFV_EINVAL, "need finite t1 > t0");
    const ccfv_problem &P = *p;
    const double dt = (t1 - t0) / nt;  // as ccfv_plain_solve
    SolveTable T[2];
    // each direction takes the uniform-offset sweep when all its lines do
    const bool fx = host_table(P, DIR_X, dt, T[0]);
    const bool fy = host_table(P, DIR_Y, dt, T[1]);
    const Layout L = layout_of(P, fx || fy);
    if (L.end > SCRATCH_DOUBLES)
        return fail(CCFV_EINVAL, "grid too large for the static scratch of the "
                                 "allocation-free solve (512 KiB: nx, ny <= 2048)");
    double *base = nullptr;
    if ((rc = scratch_ptr(&base))) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);

    // the flag starts at zero; the tables are built in place
    hipLaunchKernelGGL(clear_flag_kernel, dim3(1), dim3(64), 0, s,
                       reinterpret_cast<int *>(base + L.flag));
    for (int d = 0; d < 2; d++) {
        const Dir dir = (Dir)d;
        const int n = d == 0 ? P.nx : P.ny, nl = d == 0 ? P.ny : P.nx;
        hipLaunchKernelGGL(thomas_kernel, dim3(1), dim3(64), 0, s, n, cell_width(P, dir),
                           d == 0 ? 0.25 : 0.5, P.nu, dt, P.bc == CCFV_BC_PERIODIC ? 1 : 0,
                           chunk_for(n), base + L.tab[d], base + L.work);
        T[d].dev = base + L.tab[d];
        if (d == 0 ? fx : fy) {
            hipLaunchKernelGGL(lines_kernel, dim3((nl + 255) / 256), dim3(256), 0, s, P, d, dt,
                               T[d].coef, nl, base + L.lines[d]);
            T[d].line_dev = base + L.lines[d];
            T[d].fast_ok = true;
        }
    }
    HIPCHK(hipGetLastError());

    // a context on the stack: the caller's field and stream, the static
    // tables, no allocation anywhere
    ccfv_ctx c;
    c.P = P;
    HIPCHK(hipGetDevice(&c.device));
    c.stream = c.cur = s;
    c.ncell = (size_t)P.nx * P.ny;
    c.state = u;
    c.d_flag = reinterpret_cast<int *>(base + L.flag);
    const bool post = (P.flags & CCFV_FLAG_STEP3_POSTREMAP) != 0;
    const bool src = P.src_kind != CCFV_SRC_NONE;
    for (int b = 0; b < 2; b++) {
        c.tile[DIR_X][b] = pick_tile(DIR_X, P.nx, P.ny, b, post, src);
        c.tile[DIR_Y][b] = pick_tile(DIR_Y, P.ny, P.nx, b, false, src);
    }
    if (P.vel_kind == CCFV_VEL_SWIRL) {
        c.trig = base + L.trig;
        HIPCHK(launch_fill_trig(P.x_lo, cell_width(P, DIR_X), P.nx, c.trig, s));
        HIPCHK(launch_fill_trig(P.y_lo, cell_width(P, DIR_Y), P.ny,
                                c.trig + 2 * (size_t)(P.nx + 1), s));
    }
    c.preset[DIR_X] = &T[0];
    c.preset[DIR_Y] = &T[1];
    c.preset_h = dt;
    c.no_alloc = true;
    c.panel_on = false;
    const PlainSteps ps{u, t0, dt};
    const PlainSteps *pp = &ps;
    // steps n = 0 .. nt-1 from t0 + n dt, in place (T2S4_Parareal.cu:1234-1236)
    rc = run_steps(&c, nt, [pp](int n) {
        return StepDesc{pp->u, pp->u, 1, pp->t0 + n * pp->dt, 0.0, 0, 0.0, pp->dt};
    });
    c.trig = nullptr;  // static memory: nothing for the context to free
    if (rc) return rc;
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, c.d_flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (flag)
        return fail(CCFV_ECFL, "characteristic feet crossed: the time step is too large "
                               "for this velocity field");
    return CCFV_OK;
}

}  // extern "C"
