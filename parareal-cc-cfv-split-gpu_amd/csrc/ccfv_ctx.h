// ccfv_ctx.h — the solver context and the host runtime helpers shared by
// the single-problem entry points (ccfv_runtime.hip) and the batch / serving
// layer (ccfv_batch.hip).
#pragma once

#include <functional>
#include <map>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "ccfv_internal.h"

struct ccfv_ctx {
    ccfv_problem P{};
    int device = 0;
    hipStream_t stream = nullptr;   // the context's stream (coarse chain)
    hipStream_t stream2 = nullptr;  // fine propagators when pipelined
    hipStream_t stream3 = nullptr;  // odd iterations' coarse chain (overlapped)
    hipStream_t cur = nullptr;      // stream the next sweep launches go to
    size_t ncell = 0;
    double *state = nullptr;
    int *d_flag = nullptr;   // crossed-feet / non-finite flag
    int *h_flag = nullptr;   // pinned mirror
    std::map<std::pair<int, double>, ccfv::SolveTable> tables;
    // tiles: [dir][0 = single field, 1 = batch]
    ccfv::Tile tile[2][2];
    // Parareal buffers
    int para_ns = 0;
    double *U = nullptr, *G = nullptr, *F = nullptr, *Wk = nullptr;
    double *partial = nullptr, *slot_sums = nullptr, *h_sums = nullptr;
    // overlapped schedule: the odd iterations' work field and partials, and
    // a copy of U[ns] per iteration parity
    double *Wk2 = nullptr, *partial2 = nullptr, *res2 = nullptr;
    // x launches of the coarse chain (Parareal epilogues): wider tiles than
    // the one-field sweeps, so that fewer workgroups share the chip with the
    // concurrent fine batches; chain_blocks = ny / chain_tile.w
    ccfv::Tile chain_tile{0, 0};
    int chain_blocks = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // pipelined Parareal: fine propagators of iteration k+1 overlap the
    // coarse chain of iteration k, group by group
    int pipeline = 1;  // 0 one stream, 1 pipelined, 2 overlapped chains
    int pipe_group = 32;
    // captured Parareal phases, keyed by plan; cleared when buffers or
    // launch shapes change
    bool graphs = false;
    bool plain_graphs = false;
    // one launch for the last x sweep of a step and the first of the next
    bool fuse = true;
    // batched fine propagators keep their intermediate states in the panel
    // layout (SweepArgs::in_panel) in this work buffer of panel_fields
    // fields; the run's first x sweep reads the row-major input and its last
    // x sweep writes the row-major result
    bool panel_on = true;
    // a Parareal solve that reaches kmax above tol finishes with the serial
    // fine solve from its initial state (CCFV_OPT_SERIAL_FINISH)
    bool serial_finish = false;
    // every iteration's end state U^k(t1) kept in iter_buf (iter_cap fields;
    // iter_valid of them from the last solve) (CCFV_OPT_KEEP_ITERATES)
    bool keep_iterates = false;
    double *iter_buf = nullptr;
    int iter_cap = 0, iter_valid = 0;
    double *panel = nullptr;
    int panel_fields = 0;
    // swirl: sin(pi s), cos(pi s) of the faces of x lines (2 (nx + 1)) then
    // of y lines (2 (ny + 1)) (ensure_trig)
    double *trig = nullptr;
    // tables placed by the caller for one step size (the allocation-free
    // entry point: its tables live in static device memory); with no_alloc
    // set, get_table fails rather than allocate a table for another step
    const ccfv::SolveTable *preset[2] = {nullptr, nullptr};
    double preset_h = 0.0;
    bool no_alloc = false;
    int tiles_version = 0;
    // batches and the serving engine: the phase lane (stream) of this
    // problem's launches
    int lane = 0;
    std::map<std::string, hipGraphExec_t> gcache;
    std::vector<hipEvent_t> sync_ev;  // cross-stream events (no timing)
    // kernel timing
    bool prof = false;
    struct Rec {
        int kind;
        double bytes;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> ev_pool;
    ccfv_kernel_stat stats[CCFV_KSTAT_COUNT];
};

namespace ccfv {
namespace rt {

int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define HIPCHK(expr)                                      \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return ccfv::rt::hip_fail(e_, #expr); \
    } while (0)

// bc: the boundary mode the tile is for; the closed box (default) is the
// largest general-sweep layout, valid for either mode
bool tile_ok(Tile t, int n, int nl, Dir dir, bool post, bool src = false,
             int bc = CCFV_BC_REFERENCE);

// Host range for rocprofv3 --marker-trace (a solve, a Parareal iteration, a
// batch or serving step); without a tool attached the calls return at once.
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};
// launch shape of the coarse chain's x launches (ccfv_ctx::chain_tile)
Tile chain_tile_for(const ccfv_problem &p, Tile fallback);
// default launch shape of one sweep direction (one field or a batch)
Tile pick_tile(Dir dir, int n, int nl, bool batch, bool post, bool src);
// times every valid launch shape of direction d for one field and for
// batch_fields fields at step size h (copies of the state in scratch) and
// keeps the fastest in c->tile[d][0] and c->tile[d][1]
int autotune_dir(ccfv_ctx *c, Dir d, int batch_fields, double h);
// lane-scan doubling steps of a table (its mneg and pneg parts): sf / sb
// for the lanes of chunk_for(n) cells, sfs / sbs within the waves of a split
// line (8-cell lanes)
void scan_steps(const double *tab, int n, SolveTable &T);
double cell_width(const ccfv_problem &P, Dir dir);
int get_table(ccfv_ctx *c, Dir dir, double h, const SolveTable **out);
SweepArgs make_args(ccfv_ctx *c, Dir dir, int substep, const SolveTable &T,
                    double h, double t0, double tDT, int field0, double jdt);
bool post_flag(const ccfv_ctx *c);
// One Strang step of nf fields (in -> out) with an optional Parareal
// epilogue on its last x sweep.
struct StepDesc {
    const double *in;
    double *out;
    int nf;
    double t0, tDT;
    int field0;
    double jdt, h;
    Epi epi = EPI_NONE;
    double *g_cache = nullptr;
    const double *f_in = nullptr;
    double *u_next = nullptr;
    double *partial = nullptr;
};
// `count` steps in order; where step i+1 reads step i's result (out, or
// u_next with an epilogue) the last x sweep of i and the first of i+1 are
// one launch.  pre(i) runs before the launch that starts step i, post(i)
// after the launch that ends it.
int run_steps(ccfv_ctx *c, int count, const std::function<StepDesc(int)> &desc,
              const std::function<int(int)> &pre = {},
              const std::function<int(int)> &post = {});
int step_fields(ccfv_ctx *c, const double *in, double *out, int nf, double t0,
                double tDT, int field0, double jdt, double h, Epi epi = EPI_NONE,
                double *g_cache = nullptr, const double *f_in = nullptr,
                double *u_next = nullptr, double *partial = nullptr);
int check_flag(ccfv_ctx *c);
int validate_problem(const ccfv_problem *p);
int ensure_para(ccfv_ctx *c, int ns);
// the face sine / cosine table of a swirl problem (no-op for other
// velocities), filled on stream s; the caller frees c->trig
int ensure_trig(ccfv_ctx *c, hipStream_t s);
void clear_graphs(ccfv_ctx *c);
double increment_of(const ccfv_ctx *c, int ns, int k, std::vector<double> &slot_u2);

}  // namespace rt
}  // namespace ccfv
