// ccfv_batch.hip — many independent problems of mixed sizes solved together.
//
// Every problem keeps its own context (state, tables, Parareal buffers); the
// work of all of them goes through shared ragged launches (ccfv_phase.hip),
// so each problem's result is bitwise the one it gets alone.
#include "ccfv_batch.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

using namespace ccfv;
using namespace ccfv::rt;

namespace {

// work queued on a member context's own streams (ccfv_step, ...) finishes
// before the batch stream touches its fields
int join_contexts(ccfv_batch *b)
{
    for (ccfv_ctx *c : b->ctx) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipStreamSynchronize(c->stream2));
    }
    return CCFV_OK;
}

int ensure_sums(ccfv_batch *b, size_t n)
{
    if (n <= b->sums_cap) return CCFV_OK;
    (void)hipFree(b->sums_dev);
    if (b->sums_host) (void)hipHostFree(b->sums_host);
    b->sums_dev = nullptr;
    b->sums_host = nullptr;
    b->sums_cap = 0;
    HIPCHK(hipMalloc((void **)&b->sums_dev, sizeof(double) * n));
    HIPCHK(hipHostMalloc((void **)&b->sums_host, sizeof(double) * n));
    b->sums_cap = n;
    return CCFV_OK;
}

}  // namespace

extern "C" {

int ccfv_batch_create(const ccfv_problem *probs, int32_t n, ccfv_batch **out)
{
    if (!out) return fail(CCFV_EINVAL, "NULL out");
    *out = nullptr;
    if (!probs || n < 1 || n > 65536) return fail(CCFV_EINVAL, "need 1..65536 problems");
    for (int i = 0; i < n; i++) {
        int rc = validate_problem(&probs[i]);
        if (rc) return fail(rc, "problem " + std::to_string(i) + ": " + ccfv_last_error());
        if (probs[i].ny % 8 || probs[i].nx % 8)
            return fail(CCFV_EINVAL, "batch grids must be multiples of 8 cells");
    }
    ccfv_batch *b = new ccfv_batch();
    for (int i = 0; i < n; i++) {
        ccfv_ctx *c = nullptr;
        int rc = ccfv_create(&probs[i], &c);
        if (rc) {
            ccfv_batch_destroy(b);
            return rc;
        }
        shared_tiles(c);
        b->ctx.push_back(c);
    }
    int rc = b->run.init();
    if (rc == CCFV_OK) {
        // lanes by grid class, the largest classes first
        std::vector<int> order(n);
        for (int i = 0; i < n; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
            return (double)probs[a].nx * probs[a].ny > (double)probs[c].nx * probs[c].ny;
        });
        for (int i : order) b->ctx[i]->lane = b->run.lane_for(probs[i].nx, probs[i].ny);
    }
    hipError_t e = hipSuccess;
    if (rc == CCFV_OK && ((e = hipEventCreate(&b->ev0)) != hipSuccess ||
                          (e = hipEventCreate(&b->ev1)) != hipSuccess))
        rc = hip_fail(e, "batch events");
    if (rc) {
        ccfv_batch_destroy(b);
        return rc;
    }
    *out = b;
    return CCFV_OK;
}

void ccfv_batch_destroy(ccfv_batch *b)
{
    if (!b) return;
    b->run.release();
    for (ccfv_ctx *c : b->ctx) ccfv_destroy(c);
    (void)hipFree(b->sums_dev);
    if (b->sums_host) (void)hipHostFree(b->sums_host);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    delete b;
}

int ccfv_batch_size(ccfv_batch *b, int32_t *n)
{
    if (!b || !n) return fail(CCFV_EINVAL, "NULL argument");
    *n = (int32_t)b->ctx.size();
    return CCFV_OK;
}

int ccfv_batch_context(ccfv_batch *b, int32_t i, ccfv_ctx **ctx)
{
    if (!b || !ctx) return fail(CCFV_EINVAL, "NULL argument");
    if (i < 0 || i >= (int32_t)b->ctx.size()) return fail(CCFV_EINVAL, "index out of range");
    *ctx = b->ctx[i];
    return CCFV_OK;
}

int ccfv_batch_plain(ccfv_batch *b, const double *t0, const double *t1,
                     const int32_t *nt)
{
    ccfv::rt::TraceRange tr("ccfv_batch_plain");
    if (!b || !t0 || !t1 || !nt) return fail(CCFV_EINVAL, "NULL argument");
    const int np = (int)b->ctx.size();
    int rc;
    if ((rc = join_contexts(b))) return rc;
    std::vector<double> dt(np);
    for (int p = 0; p < np; p++) {
        if (nt[p] < 1 || !(t1[p] > t0[p])) return fail(CCFV_EINVAL, "bad time interval");
        dt[p] = (t1[p] - t0[p]) / nt[p];
        // tables up front: host work stays out of the launch sequence
        const SolveTable *t;
        if ((rc = get_table(b->ctx[p], DIR_X, dt[p], &t)) ||
            (rc = get_table(b->ctx[p], DIR_Y, dt[p], &t)))
            return rc;
    }
    std::vector<StepRun> runs;
    for (int p = 0; p < np; p++) {
        ccfv_ctx *c = b->ctx[p];
        const double a = t0[p], h = dt[p];
        runs.push_back(StepRun{c, (int)nt[p], 0, [c, a, h](int n) {
                                   return StepDesc{c->state, c->state, 1, a + n * h, 0.0, 0,
                                                   0.0, h};
                               }});
    }
    // phases of 64 rounds keep the descriptor upload small
    Phase ph;
    const int rounds = run_rounds(runs);
    for (int r0 = 0; r0 < rounds; r0 += 64) {
        ph.clear();
        if ((rc = emit_runs(ph, runs, r0, std::min(rounds, r0 + 64)))) return rc;
        if ((rc = b->run.run(ph))) return rc;
    }
    HIPCHK(hipStreamSynchronize(b->run.stream));
    for (ccfv_ctx *c : b->ctx)
        if ((rc = check_flag(c))) return rc;
    return CCFV_OK;
}

int ccfv_batch_parareal(ccfv_batch *b, const double *t0, const double *t1,
                        const int32_t *nt_fine, const int32_t *n_slices, int32_t kmax,
                        double tol, ccfv_parareal_report *reports)
{
    ccfv::rt::TraceRange tr("ccfv_batch_parareal");
    if (!b || !t0 || !t1 || !nt_fine || !n_slices)
        return fail(CCFV_EINVAL, "NULL argument");
    if (kmax < 0 || kmax > CCFV_MAX_ITERS) return fail(CCFV_EINVAL, "kmax must be in [0, 64]");
    const int np = (int)b->ctx.size();
    int rc;
    if ((rc = join_contexts(b))) return rc;
    std::vector<ParaJob> jobs(np);
    size_t sums_total = 0;
    for (int i = 0; i < np; i++) {
        ccfv_ctx *c = b->ctx[i];
        if (n_slices[i] < 1) return fail(CCFV_EINVAL, "problem " + std::to_string(i) +
                                                          ": n_slices must be positive");
        if ((rc = ensure_para(c, n_slices[i])) ||
            (rc = job_setup(jobs[i], c, t0[i], t1[i], nt_fine[i], n_slices[i], kmax, tol)))
            return fail(rc, "problem " + std::to_string(i) + ": " + ccfv_last_error());
        sums_total += 2 * (size_t)(n_slices[i] + 1);
    }
    if ((rc = ensure_sums(b, sums_total))) return rc;
    const hipStream_t s = b->run.stream;
    HIPCHK(hipEventRecord(b->ev0, s));
    Phase ph;
    {
        std::vector<RaggedCopy> cp;
        std::vector<int> lanes;
        for (ParaJob &j : jobs) {
            cp.push_back(RaggedCopy{j.c->state, j.c->U, j.c->ncell});
            lanes.push_back(j.c->lane);
        }
        ph.add_copies(cp, lanes);
    }
    std::vector<ParaJob *> live;
    for (ParaJob &j : jobs) live.push_back(&j);
    while (!live.empty()) {
        size_t used = 0;
        if ((rc = append_iteration(ph, live, b->sums_dev, &used))) return rc;
        if ((rc = b->run.run(ph))) return rc;
        ph.clear();
        if (used) {
            HIPCHK(hipMemcpyAsync(b->sums_host, b->sums_dev, sizeof(double) * used,
                                  hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        }
        std::vector<ParaJob *> next;
        for (ParaJob *j : live) {
            if (j->running && (j->plain || j->k < j->kmax)) finish_iteration(*j, b->sums_host);
            else j->running = false;
            if (j->running) next.push_back(j);
        }
        live.swap(next);
    }
    {
        std::vector<RaggedCopy> cp;
        std::vector<int> lanes;
        for (ParaJob &j : jobs) {
            cp.push_back(RaggedCopy{j.result(), j.c->state, j.c->ncell});
            lanes.push_back(j.c->lane);
        }
        ph.add_copies(cp, lanes);
        if ((rc = b->run.run(ph))) return rc;
    }
    HIPCHK(hipEventRecord(b->ev1, s));
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    int status = CCFV_OK;
    for (int i = 0; i < np; i++) {
        ParaJob &j = jobs[i];
        j.R.wall_ms = ms;
        if (reports) reports[i] = j.R;
        if ((rc = check_flag(j.c)) && status == CCFV_OK) status = rc;
        if (j.status != CCFV_OK && status == CCFV_OK)
            status = fail(j.status, "problem " + std::to_string(i) +
                                        ": Parareal iterate is not finite");
    }
    if (status) return status;
    if (tol > 0)
        for (int i = 0; i < np; i++)
            if (!jobs[i].R.converged)
                return fail(CCFV_ENOTCONV, "problem " + std::to_string(i) +
                                               " did not reach the tolerance within kmax");
    return CCFV_OK;
}

int ccfv_batch_set_option(ccfv_batch *b, int32_t key, int64_t value)
{
    if (!b) return fail(CCFV_EINVAL, "NULL batch");
    for (ccfv_ctx *c : b->ctx) {
        int rc = ccfv_set_option(c, key, value);
        if (rc) return rc;
    }
    if (key == CCFV_OPT_GRAPHS) {
        HIPCHK(hipStreamSynchronize(b->run.stream));
        b->run.graphs = value != 0;
        b->run.clear_cache();
    }
    return CCFV_OK;
}

int ccfv_batch_autotune(ccfv_batch *b, const double *t0, const double *t1, const int32_t *nt,
                        int32_t batch_fields)
{
    if (!b || !t0 || !t1 || !nt) return fail(CCFV_EINVAL, "NULL argument");
    if (batch_fields < 1 || batch_fields > 4096)
        return fail(CCFV_EINVAL, "batch_fields must be in [1, 4096]");
    int rc;
    if ((rc = join_contexts(b))) return rc;
    const int np = (int)b->ctx.size();
    // a line class: direction, cells per line, sweep path, boundary mode,
    // source; its representative is the member with the most lines
    struct Class {
        int rep = -1, lines = 0;
        Tile single{0, 0}, batch{0, 0};
    };
    std::map<std::tuple<int, int, int, int, int>, Class> classes;
    auto key_of = [&](ccfv_ctx *c, Dir d, double h) {
        const SolveTable *t = nullptr;
        int r = get_table(c, d, h, &t);
        const int fast = r == CCFV_OK && t->line_dev ? 1 : 0;
        return std::make_tuple((int)d, d == DIR_X ? c->P.nx : c->P.ny, fast, c->P.bc,
                               c->P.src_kind != CCFV_SRC_NONE ? 1 : 0);
    };
    std::vector<double> hs(np);
    for (int i = 0; i < np; i++) {
        if (nt[i] < 1 || !(t1[i] > t0[i])) return fail(CCFV_EINVAL, "bad time interval");
        hs[i] = (t1[i] - t0[i]) / nt[i];
        for (int d = 0; d < 2; d++) {
            ccfv_ctx *c = b->ctx[i];
            Class &k = classes[key_of(c, (Dir)d, hs[i])];
            const int lines = d == 0 ? c->P.ny : c->P.nx;
            if (lines > k.lines) {
                k.lines = lines;
                k.rep = i;
            }
        }
    }
    for (auto &kv : classes) {
        Class &k = kv.second;
        const Dir d = (Dir)std::get<0>(kv.first);
        if ((rc = autotune_dir(b->ctx[k.rep], d, batch_fields, hs[k.rep]))) return rc;
        k.single = b->ctx[k.rep]->tile[d][0];
        k.batch = b->ctx[k.rep]->tile[d][1];
    }
    // every member of a class takes its shapes (when its grid admits them),
    // so the members still share one launch per stage and class; the coarse
    // chain's shape stays fixed (the increments' block sums depend on it)
    for (int i = 0; i < np; i++) {
        ccfv_ctx *c = b->ctx[i];
        const bool post = post_flag(c), src = c->P.src_kind != CCFV_SRC_NONE;
        for (int d = 0; d < 2; d++) {
            const Class &k = classes[key_of(c, (Dir)d, hs[i])];
            const int n = d == 0 ? c->P.nx : c->P.ny, nl = d == 0 ? c->P.ny : c->P.nx;
            if (tile_ok(k.single, n, nl, (Dir)d, post && d == 0, src, c->P.bc))
                c->tile[d][0] = k.single;
            if (tile_ok(k.batch, n, nl, (Dir)d, post && d == 0, src, c->P.bc))
                c->tile[d][1] = k.batch;
        }
        c->tiles_version++;
        clear_graphs(c);
    }
    HIPCHK(hipStreamSynchronize(b->run.stream));
    b->run.clear_cache();
    return CCFV_OK;
}

}  // extern "C"
