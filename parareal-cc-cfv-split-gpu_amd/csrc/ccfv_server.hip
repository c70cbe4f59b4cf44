// ccfv_server.hip — the serving engine: Parareal solve requests of mixed
// sizes, admitted by a device-memory budget and advanced together.
//
// The engine owns one device arena (the memory budget, allocated once) and a
// fixed number of slots.  A request waits in arrival order until its arena
// block and a slot are free (ccfv_sched.h).  Once admitted its whole
// Parareal state (U, G, F, the coarse work field, increment partials) stays
// in its block until it finishes.  Each engine step advances every resident
// request by one Parareal iteration (a newly admitted request also runs its
// k = 0 coarse sweep) through shared ragged launches, then reads the slot
// sums of all of them with one copy and decides, per request, whether to go
// on.  Requests at different iterations and of different grids share
// launches; each one's kernels, inputs and order are those of
// ccfv_parareal_solve, so its result is bitwise the one it gets alone.
#include <algorithm>
#include <cstring>
#include <deque>
#include <memory>
#include <unordered_map>

#include "ccfv_batch.h"
#include "ccfv_sched.h"

using namespace ccfv;
using namespace ccfv::rt;

namespace {

struct Request {
    int64_t id = 0;
    ccfv_request r{};
    bool plain = false;      // run as the plain serial solve
    std::vector<double> u0;  // empty: initial data from r.ic
    size_t bytes = 0, off = 0;
    int slot = -1;
    ccfv_ctx ctx;            // no streams of its own: pointers into the arena
    ParaJob job;
    bool done = false;
    int status = CCFV_OK;
    std::vector<double> result;
};

}  // namespace

struct ccfv_server {
    ccfv_server_config cfg{};
    int device = 0;
    char *arena = nullptr;
    sched::ArenaAlloc alloc;
    sched::AdmissionQueue queue;
    PhaseRunner run;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::vector<Request *> active;
    std::vector<int> free_slots;
    std::deque<int64_t> finished;
    int64_t next_id = 1;
    int *d_flags = nullptr, *h_flags = nullptr;
    double *sums_dev = nullptr, *sums_host = nullptr;
    size_t sums_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t steps = 0;
    double last_step_ms = 0;
};

namespace {

int request_footprint(const ccfv_request *r, bool plain, size_t *bytes)
{
    if (plain) {  // the state field alone
        *bytes = sched::round_up(sizeof(double) * (size_t)r->problem.nx * r->problem.ny, 256);
        return CCFV_OK;
    }
    // coarse-chain blocks: one per pair of x lines (the one-field x shape)
    const int nb = r->problem.ny / 2;
    *bytes = sched::request_bytes(r->problem.nx, r->problem.ny, r->n_slices, nb);
    return CCFV_OK;
}

int validate_request(const ccfv_server *s, const ccfv_request *r)
{
    int rc = validate_problem(&r->problem);
    if (rc) return rc;
    if (r->problem.nx % 8 || r->problem.ny % 8)
        return fail(CCFV_EINVAL, "served grids must be multiples of 8 cells");
    if (r->n_slices < 0 || (s && r->n_slices > s->cfg.max_slices))
        return fail(CCFV_EINVAL, "n_slices out of range for this engine");
    if (r->n_slices == 0 ? r->nt_fine < 1
                         : (r->nt_fine < r->n_slices || r->nt_fine % r->n_slices))
        return fail(CCFV_EINVAL, "nt_fine must be a positive multiple of n_slices");
    if (r->kmax < 0 || r->kmax > CCFV_MAX_ITERS) return fail(CCFV_EINVAL, "kmax must be in [0, 64]");
    if (!(r->t1 > r->t0)) return fail(CCFV_EINVAL, "need t1 > t0");
    if (r->ic.kind < 0 || r->ic.kind > CCFV_IC_COSINE_BELL)
        return fail(CCFV_EINVAL, "unknown initial condition");
    return CCFV_OK;
}

// arena block -> the request's light context
int place(ccfv_server *s, Request &q)
{
    ccfv_ctx &c = q.ctx;
    c.P = q.r.problem;
    c.device = s->device;
    c.stream = c.stream2 = c.cur = s->run.stream;
    c.ncell = (size_t)c.P.nx * c.P.ny;
    c.lane = s->run.lane_for(c.P.nx, c.P.ny);
    shared_tiles(&c);
    const int ns = q.r.n_slices;
    const size_t f = sizeof(double) * c.ncell;
    char *p = s->arena + q.off;
    auto take = [&](size_t b) {
        char *r = p;
        p += sched::round_up(b, 256);
        return (double *)r;
    };
    c.d_flag = s->d_flags + q.slot;
    c.h_flag = s->h_flags + q.slot;
    int rc;
    if (q.plain) {
        c.U = take(f);
        c.state = nullptr;
        if ((size_t)(p - (s->arena + q.off)) > q.bytes)
            return fail(CCFV_EDEVICE, "request layout exceeds its arena block");
        if ((rc = job_setup_plain(q.job, &c, q.r.t0, q.r.t1, q.r.nt_fine))) return rc;
    } else {
    c.U = take(f * (ns + 1));
    c.G = take(f * ns);
    c.F = take(f * ns);
    c.Wk = take(f);
    c.chain_blocks = c.P.ny / c.chain_tile.w;
    c.partial = take(sizeof(double) * 2 * (size_t)c.chain_blocks * (ns + 1));
    c.para_ns = ns;
    c.state = nullptr;
    c.serial_finish = s->cfg.serial_finish != 0;
    if ((size_t)(p - (s->arena + q.off)) > q.bytes)
        return fail(CCFV_EDEVICE, "request layout exceeds its arena block");
    if ((rc = job_setup(q.job, &c, q.r.t0, q.r.t1, q.r.nt_fine, ns, q.r.kmax, q.r.tol)))
        return rc;
    }
    if ((rc = ensure_trig(&c, s->run.stream))) return rc;
    HIPCHK(hipMemsetAsync(c.d_flag, 0, sizeof(int), s->run.stream));
    if (!q.u0.empty()) {
        HIPCHK(hipMemcpyAsync(c.U, q.u0.data(), f, hipMemcpyHostToDevice, s->run.stream));
    } else {
        HIPCHK(launch_fill_exact(c.P, q.r.ic, q.r.t0, c.U, s->run.stream));
    }
    return CCFV_OK;
}

void release_tables(ccfv_ctx &c)
{
    for (auto &kv : c.tables) {
        (void)hipFree(kv.second.dev);
        (void)hipFree(kv.second.line_dev);
    }
    c.tables.clear();
    (void)hipFree(c.trig);
    c.trig = nullptr;
}

// finished: result to the host, arena block and slot back
int retire(ccfv_server *s, Request &q)
{
    const size_t f = sizeof(double) * q.ctx.ncell;
    q.result.resize(q.ctx.ncell);
    if (q.status == CCFV_OK || q.status == CCFV_ENOTCONV)
        HIPCHK(hipMemcpy(q.result.data(), q.job.result(), f, hipMemcpyDeviceToHost));
    release_tables(q.ctx);
    s->alloc.release(q.off, q.bytes);
    s->free_slots.push_back(q.slot);
    q.slot = -1;
    q.done = true;
    q.u0.clear();
    q.u0.shrink_to_fit();
    s->finished.push_back(q.id);
    return CCFV_OK;
}

}  // namespace

extern "C" {

int ccfv_request_bytes(const ccfv_request *r, uint64_t *bytes)
{
    if (!r || !bytes) return fail(CCFV_EINVAL, "NULL argument");
    int rc = validate_request(nullptr, r);
    if (rc) return rc;
    size_t b = 0;
    request_footprint(r, r->n_slices == 0, &b);
    *bytes = b;
    return CCFV_OK;
}

int ccfv_server_create(const ccfv_server_config *cfg, ccfv_server **out)
{
    if (!out) return fail(CCFV_EINVAL, "NULL out");
    *out = nullptr;
    ccfv_server_config c{};
    if (cfg) c = *cfg;
    if (c.max_active == 0) c.max_active = 256;
    if (c.max_slices == 0) c.max_slices = 64;
    if (c.max_bypass == 0) c.max_bypass = 8;
    if (c.max_active < 1 || c.max_active > 4096) return fail(CCFV_EINVAL, "max_active must be in [1, 4096]");
    if (c.max_slices < 1 || c.max_slices > 1024) return fail(CCFV_EINVAL, "max_slices must be in [1, 1024]");
    if (c.max_bypass < 0) return fail(CCFV_EINVAL, "max_bypass must be >= 0");
    if (c.serial_finish != 0 && c.serial_finish != 1)
        return fail(CCFV_EINVAL, "serial_finish must be 0 or 1");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(CCFV_EDEVICE, "no HIP device available");
    if (c.mem_budget == 0) {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        c.mem_budget = fr / 2;
    }
    ccfv_server *s = new ccfv_server();
    s->cfg = c;
    s->queue = sched::AdmissionQueue(c.max_bypass);
    int rc = CCFV_OK;
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipMalloc((void **)&s->arena, c.mem_budget);
    if (e == hipSuccess) e = hipMalloc((void **)&s->d_flags, sizeof(int) * c.max_active);
    if (e == hipSuccess) e = hipHostMalloc((void **)&s->h_flags, sizeof(int) * c.max_active);
    s->sums_cap = 2 * (size_t)(c.max_slices + 1) * c.max_active;
    if (e == hipSuccess) e = hipMalloc((void **)&s->sums_dev, sizeof(double) * s->sums_cap);
    if (e == hipSuccess) e = hipHostMalloc((void **)&s->sums_host, sizeof(double) * s->sums_cap);
    if (e == hipSuccess) e = hipEventCreate(&s->ev0);
    if (e == hipSuccess) e = hipEventCreate(&s->ev1);
    if (e != hipSuccess) rc = fail(CCFV_ENOMEM, std::string("serving engine allocation: ") +
                                                    hipGetErrorString(e));
    if (rc == CCFV_OK) rc = s->run.init();
    if (rc) {
        ccfv_server_destroy(s);
        return rc;
    }
    s->alloc.reset(c.mem_budget, 256);
    for (int i = c.max_active - 1; i >= 0; i--) s->free_slots.push_back(i);
    memset(s->h_flags, 0, sizeof(int) * c.max_active);
    *out = s;
    return CCFV_OK;
}

void ccfv_server_destroy(ccfv_server *s)
{
    if (!s) return;
    s->run.release();
    for (auto &kv : s->reqs) release_tables(kv.second->ctx);
    (void)hipFree(s->arena);
    (void)hipFree(s->d_flags);
    if (s->h_flags) (void)hipHostFree(s->h_flags);
    (void)hipFree(s->sums_dev);
    if (s->sums_host) (void)hipHostFree(s->sums_host);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    delete s;
}

int ccfv_server_submit(ccfv_server *s, const ccfv_request *r, const double *u0, int64_t *id)
{
    if (!s || !r || !id) return fail(CCFV_EINVAL, "NULL argument");
    *id = -1;
    int rc = validate_request(s, r);
    if (rc) return rc;
    // adaptive policy: a Parareal request that arrives while the engine is
    // busy runs as the plain serial solve (the reference of its accuracy,
    // a sixth or less of the device work)
    const bool busy = s->cfg.adaptive_above > 0 &&
                      s->queue.size() + s->active.size() >= (size_t)s->cfg.adaptive_above;
    const bool plain = r->n_slices == 0 || busy;
    size_t bytes = 0;
    request_footprint(r, plain, &bytes);
    if (sched::round_up(bytes, 256) > s->alloc.size())
        return fail(CCFV_EBUSY, "request needs " + std::to_string(bytes) +
                                    " device bytes, more than the engine's budget");
    auto q = std::make_unique<Request>();
    q->id = s->next_id++;
    q->r = *r;
    q->plain = plain;
    q->bytes = bytes;
    if (u0) q->u0.assign(u0, u0 + (size_t)r->problem.nx * r->problem.ny);
    s->queue.push(q->id, bytes);
    *id = q->id;
    s->reqs.emplace(q->id, std::move(q));
    return CCFV_OK;
}

int ccfv_server_step(ccfv_server *s, ccfv_server_status *st)
{
    ccfv::rt::TraceRange tr("ccfv_server_step");
    if (!s) return fail(CCFV_EINVAL, "NULL server");
    int rc;
    const hipStream_t stream = s->run.stream;
    // admit what fits
    for (const sched::AdmissionQueue::Grant &g :
         s->queue.admit(s->alloc, (int)s->free_slots.size())) {
        Request &q = *s->reqs.at(g.id);
        q.off = g.offset;
        q.slot = s->free_slots.back();
        s->free_slots.pop_back();
        if ((rc = place(s, q))) {
            q.status = rc;
            release_tables(q.ctx);
            s->alloc.release(q.off, q.bytes);
            s->free_slots.push_back(q.slot);
            q.slot = -1;
            q.done = true;
            s->finished.push_back(q.id);
            continue;
        }
        s->active.push_back(&q);
    }
    float ms = 0.f;
    if (!s->active.empty()) {
        HIPCHK(hipEventRecord(s->ev0, stream));
        std::vector<ParaJob *> jobs;
        for (Request *q : s->active) jobs.push_back(&q->job);
        Phase ph;
        size_t used = 0;
        if ((rc = append_iteration(ph, jobs, s->sums_dev, &used))) return rc;
        if ((rc = s->run.run(ph))) return rc;
        if (used)
            HIPCHK(hipMemcpyAsync(s->sums_host, s->sums_dev, sizeof(double) * used,
                                  hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(s->h_flags, s->d_flags, sizeof(int) * s->cfg.max_active,
                              hipMemcpyDeviceToHost, stream));
        HIPCHK(hipEventRecord(s->ev1, stream));
        HIPCHK(hipStreamSynchronize(stream));
        (void)hipEventElapsedTime(&ms, s->ev0, s->ev1);
        std::vector<Request *> still;
        for (Request *q : s->active) {
            ParaJob &j = q->job;
            if (j.running && (j.plain || j.k < j.kmax)) finish_iteration(j, s->sums_host);
            else j.running = false;
            if (s->h_flags[q->slot]) {
                q->status = fail(CCFV_ECFL, "characteristic feet crossed: the time step is "
                                            "too large for this velocity field");
                j.running = false;
            } else if (j.status != CCFV_OK) {
                q->status = j.status;
            }
            if (j.running) {
                still.push_back(q);
                continue;
            }
            if (q->status == CCFV_OK && !j.plain && j.tol > 0 && !j.R.converged)
                q->status = CCFV_ENOTCONV;
            if ((rc = retire(s, *q))) return rc;
        }
        s->active.swap(still);
        s->steps++;
    }
    s->last_step_ms = ms;
    if (st) {
        st->active = (int32_t)s->active.size();
        st->queued = (int32_t)s->queue.size();
        st->finished = (int32_t)s->finished.size();
        st->reserved = 0;
        st->mem_used = s->alloc.used();
        st->mem_budget = s->alloc.size();
        st->steps = s->steps;
        st->last_step_ms = ms;
    }
    return CCFV_OK;
}

int ccfv_server_next_done(ccfv_server *s, int64_t *id)
{
    if (!s || !id) return fail(CCFV_EINVAL, "NULL argument");
    *id = -1;
    if (!s->finished.empty()) {
        *id = s->finished.front();
        s->finished.pop_front();
    }
    return CCFV_OK;
}

int ccfv_server_result(ccfv_server *s, int64_t id, double *u, ccfv_parareal_report *rep,
                       int32_t *status)
{
    if (!s) return fail(CCFV_EINVAL, "NULL server");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(CCFV_EINVAL, "unknown request id");
    Request &q = *it->second;
    if (!q.done) return fail(CCFV_EBUSY, "request has not finished");
    if (u && !q.result.empty()) memcpy(u, q.result.data(), sizeof(double) * q.result.size());
    if (rep) {
        *rep = q.job.R;
        rep->wall_ms = 0.0;
    }
    if (status) *status = q.status;
    // a request whose id was not yet popped leaves the finished list too
    auto f = std::find(s->finished.begin(), s->finished.end(), id);
    if (f != s->finished.end()) s->finished.erase(f);
    s->reqs.erase(it);
    return CCFV_OK;
}

}  // extern "C"
