// ccfv_batch.h — many problems through shared ragged launches: the phase
// builder and runner (ccfv_phase.hip), the per-problem Parareal job used by
// the batch (ccfv_batch.hip) and the serving engine (ccfv_server.hip).
#pragma once

#include <functional>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "ccfv_ctx.h"

namespace ccfv {
namespace rt {

// one problem's contribution to one sweep stage
struct StageItem {
    int lane = 0;  // the problem's phase lane (ccfv_ctx::lane)
    Dir dir;
    int bc;
    bool post;
    Epi epi;
    bool fuse;  // x: also the next step's first x sub-step (out2)
    Tile tile;
    RaggedItem item;
};

// one ragged launch of a phase; items / prefix are byte offsets into the
// phase's descriptor arena
struct Launch {
    enum Kind { SWEEP, CORR, REDUCE, COPY } kind;
    int lane = 0;  // the stream it goes to (PhaseRunner::lanes)
    Dir dir = DIR_X;
    Tile tile{0, 0};
    int ch = 1, wpl = 1, bc = 0;  // wpl: waves per line (split lines)
    bool fast = false, post = false, fuse = false;
    Epi epi = EPI_NONE;
    int nitems = 0, nblocks = 0;
    size_t lds = 0;
    size_t items = 0, prefix = 0;
};

// descriptors of a run of launches, uploaded with one copy
struct Phase {
    std::vector<char> host;
    std::vector<Launch> launches;
    void clear();
    size_t append(const void *p, size_t bytes);
    void add_sweep_stage(std::vector<StageItem> &items);
    // lanes: the lane of each item's problem; one launch per lane
    void add_corrections(const std::vector<RaggedCorr> &c, const std::vector<int> &blocks,
                         const std::vector<int> &lanes);
    void add_reductions(const std::vector<RaggedReduce> &r, const std::vector<int> &lanes);
    void add_copies(const std::vector<RaggedCopy> &c, const std::vector<int> &lanes);
    bool empty() const { return launches.empty(); }
};

// Issues phases.  The descriptors go through two pinned host + device
// arenas used alternately, so building phase p+1 on the host overlaps the
// device running phase p.
// Lanes: every problem belongs to one of LANES lanes (ccfv_ctx::lane, by
// grid class) and every launch holds the items of one lane, so launches of
// different lanes touch different problems.  Each lane issues on its own
// stream (lane 0 on `stream`): after the descriptor upload the other lanes
// wait for `stream`, and at the end of the phase `stream` waits for them, so
// the grid classes of a batch run side by side instead of one after another
// within a phase, and everything after the phase on `stream` sees all of it.
// With graphs on, the launches of a phase (not its descriptor upload) are
// one hipGraph: a launch list seen a second time — same kernels, shapes,
// lanes, item counts and descriptor offsets in the same arena — is captured
// (with the lanes' fork and join) and from then on replayed with one graph
// launch; the kernels read the new descriptors from the arena, so only the
// launch list has to repeat.
struct PhaseRunner {
    static constexpr int LANES = 4;
    hipStream_t stream = nullptr;
    hipStream_t lanes[LANES] = {};  // lanes[0] == stream
    hipEvent_t fork_ev = nullptr, join_ev[LANES] = {};
    // grid classes (nx, ny) and their lanes, and the cells given to each
    // lane so far (lane_for: a new class goes to the lane with the fewest)
    std::map<std::pair<int, int>, int> class_lane;
    double lane_cells[LANES] = {};
    int lane_for(int nx, int ny);
    char *hbuf[2] = {nullptr, nullptr};
    char *dbuf[2] = {nullptr, nullptr};
    size_t cap[2] = {0, 0};
    hipEvent_t done[2] = {nullptr, nullptr};
    int slot = 0;
    bool graphs = false;
    std::map<std::string, hipGraphExec_t> cache;
    std::map<std::string, int> seen;
    long long graph_launches = 0, eager_launches = 0;  // phases issued each way
    int init();
    void release();
    void clear_cache();
    int run(Phase &ph);
    int issue(const Phase &ph, const char *d);
};

// A run of chained Strang steps of one problem (as run_steps), emitted into
// a phase round by round: round r holds the x launch that ends step r-1 and
// starts step r (one fused launch) and the y sweep of step r.
struct StepRun {
    ccfv_ctx *c;
    int count;
    int shape;  // 1: batch launch shapes (fine propagators), 0: one-field
    std::function<StepDesc(int)> desc;
};
// emits rounds [r0, r1) of every run (round `count` is a run's last x
// sweep); returns CCFV_OK or the table lookup's error
int emit_runs(Phase &ph, const std::vector<StepRun> &runs, int r0, int r1);
// rounds a set of runs takes
int run_rounds(const std::vector<StepRun> &runs);

// launch shapes every grid of a batch can use (multiples of 8 cells)
void shared_tiles(ccfv_ctx *c);

// One problem's Parareal run, advanced one iteration per phase.  The work
// and its order are those of ccfv_parareal_solve: the k = 0 coarse sweep
// (cached G), then per iteration the fine propagators of slices k..ns-1,
// the correction of slot k+1, the coarse chain with the fused correction and
// the per-slot increment sums.
struct ParaJob {
    ccfv_ctx *c = nullptr;
    int ns = 0, J = 0, nb = 0, kmax = 0;
    double t0 = 0, t1 = 0, DT = 0, dt = 0, tol = 0;
    int nt_fine = 0;
    // reaching kmax above tol switches the job to its serial fine solve
    // from U[0] (the initial state, never written by the iterations):
    // plain = true from then on, fine steps counted after para_fine
    bool serial_finish = false;
    double para_fine = 0;
    const SolveTable *fx = nullptr, *fy = nullptr, *cx = nullptr, *cy = nullptr;
    int k = 0;            // next iteration
    bool started = false; // k = 0 coarse sweep issued
    bool running = true;
    int status = CCFV_OK;
    size_t sums_off = 0;  // this phase's slot sums in the shared sums buffer
    std::vector<double> slot_u2;
    ccfv_parareal_report R{};
    // plain serial solve instead of Parareal: nt steps of dt from t0 on
    // U[0] in place (the steps of ccfv_plain_solve), PLAIN_CHUNK per phase
    bool plain = false;
    int nt = 0, done = 0;
    const double *result() const { return plain ? c->U : c->U + (size_t)ns * c->ncell; }
};
constexpr int PLAIN_CHUNK = 64;
// validates the plan, sets up the tables and the report (c->U etc. must be
// in place: ensure_para or the caller's own buffers)
int job_setup(ParaJob &j, ccfv_ctx *c, double t0, double t1, int nt_fine, int ns,
              int kmax, double tol);
// a plain serial solve of nt steps over [t0, t1] (c->U: one field)
int job_setup_plain(ParaJob &j, ccfv_ctx *c, double t0, double t1, int nt);
// Appends, for every job, its k = 0 coarse sweep if not yet issued and then
// (if it has iterations left) one iteration; for a plain job its next
// PLAIN_CHUNK steps.  Jobs advance in lock-step
// rounds; slot sums go to sums_dev at each job's sums_off (2 (ns+1) doubles),
// assigned here; *used: the doubles of sums_dev used.
int append_iteration(Phase &ph, std::vector<ParaJob *> &jobs, double *sums_dev,
                     size_t *used);
// reads an iteration's slot sums (host copy of sums_dev) into the job; for
// a plain job counts the steps of the phase.  A Parareal job that stops at
// kmax above tol with serial_finish set becomes a plain job of nt_fine steps
// of (t1 - t0) / nt_fine on U[0] (ccfv_plain_solve's steps) and keeps
// running.
void finish_iteration(ParaJob &j, const double *sums_host);

}  // namespace rt
}  // namespace ccfv

struct ccfv_batch {
    std::vector<ccfv_ctx *> ctx;  // lane of each set by PhaseRunner::lane_for
    ccfv::rt::PhaseRunner run;
    // per-slot increment sums of every problem
    double *sums_dev = nullptr, *sums_host = nullptr;
    size_t sums_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};
