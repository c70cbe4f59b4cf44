// ccfv_phase.hip — phases of ragged launches over many problems, and the
// per-problem Parareal job that the batch and the serving engine advance.
//
// A phase is a list of launches whose work-list descriptors (one RaggedItem,
// RaggedCorr, RaggedReduce or RaggedCopy per problem, plus a block prefix per
// launch) are built on the host and uploaded with one copy.  Problems advance
// in lock-step rounds: in round r every problem that has an r-th sub-step
// contributes its blocks to one launch per sweep stage and line class.
// Every problem sees exactly the kernels, inputs and order it gets when
// solved alone, so its result is bitwise the same.
#include "ccfv_batch.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ccfv {
namespace rt {

void shared_tiles(ccfv_ctx *c)
{
    const bool post = post_flag(c);
    const Tile x1{2, 2}, y1{2, 2}, yb{2, 1};
    Tile xb{8, 4};
    if (!tile_ok(xb, c->P.nx, c->P.ny, DIR_X, post, c->P.src_kind != CCFV_SRC_NONE))
        xb = Tile{2, 2};
    c->tile[DIR_X][0] = x1;
    c->tile[DIR_X][1] = xb;
    c->tile[DIR_Y][0] = y1;
    c->tile[DIR_Y][1] = yb;
    c->chain_tile = chain_tile_for(c->P, x1);
    c->tiles_version++;
    clear_graphs(c);
}

// ---------------------------------------------------------------------------
// phase builder
// ---------------------------------------------------------------------------
void Phase::clear()
{
    host.clear();
    launches.clear();
}

size_t Phase::append(const void *p, size_t bytes)
{
    size_t off = (host.size() + 15) & ~(size_t)15;
    host.resize(off + bytes);
    memcpy(host.data() + off, p, bytes);
    return off;
}

// sweep items of one stage, grouped into launches
void Phase::add_sweep_stage(std::vector<StageItem> &items)
{
    // one launch shares the lane, the kernel instance (line chunk, waves per
    // line, boundary mode, path, step-3 variant, epilogue) and the block shape
    struct Key {
        int lane, ch, wpl, bc, fast, post, epi, fuse, w, nw;
        bool operator<(const Key &o) const
        {
            return std::tie(lane, ch, wpl, bc, fast, post, epi, fuse, w, nw) <
                   std::tie(o.lane, o.ch, o.wpl, o.bc, o.fast, o.post, o.epi, o.fuse, o.w,
                            o.nw);
        }
    };
    std::map<Key, std::vector<const StageItem *>> groups;
    for (const StageItem &s : items) {
        const bool fast = s.item.a.lineS != nullptr;
        Key k{s.lane, chunk_for(s.item.a.n), fast ? fast_wpl(s.item.a.n) : 1, s.bc,
              fast ? 1 : 0, s.post ? 1 : 0, (int)s.epi, s.fuse ? 1 : 0, s.tile.w, s.tile.nw};
        groups[k].push_back(&s);
    }
    for (auto &kv : groups) {
        const Key &k = kv.first;
        std::vector<RaggedItem> it;
        std::vector<int> prefix(1, 0);
        size_t lds = 0;
        for (const StageItem *s : kv.second) {
            it.push_back(s->item);
            prefix.push_back(prefix.back() + s->item.nfields * s->item.ntiles);
            size_t b = k.fast ? sweep_lds_bytes_fast(s->item.a.n, s->tile)
                              : sweep_lds_bytes_generic(
                                    s->item.a.n, s->tile, s->item.a.src_kind != CCFV_SRC_NONE,
                                    (s->dir == DIR_X && k.bc == CCFV_BC_REFERENCE && k.post) ||
                                        !rhs_in_regs(k.ch, k.bc,
                                                     s->item.a.src_kind != CCFV_SRC_NONE,
                                                     false));
            lds = std::max(lds, b);
        }
        Launch L;
        L.kind = Launch::SWEEP;
        L.lane = k.lane;
        L.dir = kv.second.front()->dir;
        L.tile = kv.second.front()->tile;
        L.ch = k.ch;
        L.wpl = k.wpl;
        L.bc = k.bc;
        L.fast = k.fast != 0;
        L.post = k.post != 0;
        L.fuse = k.fuse != 0;
        L.epi = (Epi)k.epi;
        L.nitems = (int)it.size();
        L.nblocks = prefix.back();
        L.lds = lds;
        L.items = append(it.data(), sizeof(RaggedItem) * it.size());
        L.prefix = append(prefix.data(), sizeof(int) * prefix.size());
        launches.push_back(L);
    }
}

void Phase::add_corrections(const std::vector<RaggedCorr> &c, const std::vector<int> &blocks,
                            const std::vector<int> &lanes)
{
    for (int lane = 0; lane < PhaseRunner::LANES; lane++) {
        std::vector<RaggedCorr> it;
        std::vector<int> prefix(1, 0);
        for (size_t i = 0; i < c.size(); i++)
            if (lanes[i] == lane) {
                it.push_back(c[i]);
                prefix.push_back(prefix.back() + blocks[i]);
            }
        if (it.empty()) continue;
        Launch L;
        L.kind = Launch::CORR;
        L.lane = lane;
        L.nitems = (int)it.size();
        L.nblocks = prefix.back();
        L.items = append(it.data(), sizeof(RaggedCorr) * it.size());
        L.prefix = append(prefix.data(), sizeof(int) * prefix.size());
        launches.push_back(L);
    }
}

void Phase::add_reductions(const std::vector<RaggedReduce> &r, const std::vector<int> &lanes)
{
    for (int lane = 0; lane < PhaseRunner::LANES; lane++) {
        std::vector<RaggedReduce> it;
        std::vector<int> prefix(1, 0);
        for (size_t i = 0; i < r.size(); i++)
            if (lanes[i] == lane) {
                it.push_back(r[i]);
                prefix.push_back(prefix.back() + r[i].count);
            }
        if (it.empty()) continue;
        Launch L;
        L.kind = Launch::REDUCE;
        L.lane = lane;
        L.nitems = (int)it.size();
        L.nblocks = prefix.back();
        L.items = append(it.data(), sizeof(RaggedReduce) * it.size());
        L.prefix = append(prefix.data(), sizeof(int) * prefix.size());
        launches.push_back(L);
    }
}

void Phase::add_copies(const std::vector<RaggedCopy> &c, const std::vector<int> &lanes)
{
    for (int lane = 0; lane < PhaseRunner::LANES; lane++) {
        std::vector<RaggedCopy> it;
        std::vector<int> prefix(1, 0);
        for (size_t i = 0; i < c.size(); i++)
            if (lanes[i] == lane) {
                it.push_back(c[i]);
                const size_t b = (c[i].n + 255) / 256;
                prefix.push_back(prefix.back() +
                                 (int)std::min<size_t>(std::max<size_t>(b, 1), 1024));
            }
        if (it.empty()) continue;
        Launch L;
        L.kind = Launch::COPY;
        L.lane = lane;
        L.nitems = (int)it.size();
        L.nblocks = prefix.back();
        L.items = append(it.data(), sizeof(RaggedCopy) * it.size());
        L.prefix = append(prefix.data(), sizeof(int) * prefix.size());
        launches.push_back(L);
    }
}

// ---------------------------------------------------------------------------
// phase runner
// ---------------------------------------------------------------------------
int PhaseRunner::init()
{
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    lanes[0] = stream;
    for (int l = 1; l < LANES; l++) HIPCHK(hipStreamCreateWithFlags(&lanes[l], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    for (int l = 1; l < LANES; l++)
        HIPCHK(hipEventCreateWithFlags(&join_ev[l], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&done[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&done[1], hipEventDisableTiming));
    return CCFV_OK;
}

int PhaseRunner::lane_for(int nx, int ny)
{
    const auto key = std::make_pair(nx, ny);
    auto it = class_lane.find(key);
    if (it != class_lane.end()) return it->second;
    int best = 0;
    for (int l = 1; l < LANES; l++)
        if (lane_cells[l] < lane_cells[best]) best = l;
    lane_cells[best] += (double)nx * ny;
    class_lane.emplace(key, best);
    return best;
}

void PhaseRunner::clear_cache()
{
    for (auto &kv : cache) (void)hipGraphExecDestroy(kv.second);
    cache.clear();
    seen.clear();
}

void PhaseRunner::release()
{
    for (int l = 0; l < LANES; l++)
        if (lanes[l]) (void)hipStreamSynchronize(lanes[l]);
    clear_cache();
    for (int s = 0; s < 2; s++) {
        if (hbuf[s]) (void)hipHostFree(hbuf[s]);
        if (dbuf[s]) (void)hipFree(dbuf[s]);
        if (done[s]) (void)hipEventDestroy(done[s]);
        hbuf[s] = dbuf[s] = nullptr;
        done[s] = nullptr;
        cap[s] = 0;
    }
    for (int l = 1; l < LANES; l++) {
        if (lanes[l]) (void)hipStreamDestroy(lanes[l]);
        if (join_ev[l]) (void)hipEventDestroy(join_ev[l]);
        lanes[l] = nullptr;
        join_ev[l] = nullptr;
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    fork_ev = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = lanes[0] = nullptr;
}

int PhaseRunner::run(Phase &ph)
{
    if (ph.launches.empty()) return CCFV_OK;
    const size_t bytes = ph.host.size();
    const int s = slot;
    slot ^= 1;
    // this arena was last read by the phase before the previous one
    HIPCHK(hipEventSynchronize(done[s]));
    if (bytes > cap[s]) {
        // captured phases hold addresses in the arena that goes away
        HIPCHK(hipStreamSynchronize(stream));
        clear_cache();
        if (hbuf[s]) (void)hipHostFree(hbuf[s]);
        if (dbuf[s]) (void)hipFree(dbuf[s]);
        hbuf[s] = nullptr;
        dbuf[s] = nullptr;
        cap[s] = 0;
        const size_t c = std::max(bytes, (size_t)1 << 16);
        HIPCHK(hipHostMalloc((void **)&hbuf[s], c));
        HIPCHK(hipMalloc((void **)&dbuf[s], c));
        cap[s] = c;
    }
    memcpy(hbuf[s], ph.host.data(), bytes);
    HIPCHK(hipMemcpyAsync(dbuf[s], hbuf[s], bytes, hipMemcpyHostToDevice, stream));
    const char *d = dbuf[s];
    if (graphs) {
        // the launch list and the arena it reads
        std::string key((const char *)&d, sizeof(d));
        for (const Launch &L : ph.launches) {
            const long long f[] = {L.kind, L.lane, L.dir, L.tile.w, L.tile.nw, L.ch, L.wpl,
                                   L.bc, L.fast, L.post, L.fuse, L.epi, L.nitems, L.nblocks,
                                   (long long)L.lds, (long long)L.items, (long long)L.prefix};
            key.append((const char *)f, sizeof(f));
        }
        auto it = cache.find(key);
        if (it == cache.end() && ++seen[key] >= 2) {
            if (cache.size() >= 256) clear_cache();
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
            int r2 = issue(ph, d);
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamEndCapture(stream, &gr);
            hipGraphExec_t x = nullptr;
            if (r2 == CCFV_OK && e == hipSuccess) e = hipGraphInstantiate(&x, gr, nullptr, nullptr, 0);
            if (gr) (void)hipGraphDestroy(gr);
            if (r2) return r2;
            if (e != hipSuccess) return hip_fail(e, "phase graph capture");
            it = cache.emplace(key, x).first;
        }
        if (it != cache.end()) {
            HIPCHK(hipGraphLaunch(it->second, stream));
            graph_launches++;
            HIPCHK(hipEventRecord(done[s], stream));
            return CCFV_OK;
        }
    }
    int rc = issue(ph, d);
    if (rc) return rc;
    eager_launches++;
    HIPCHK(hipEventRecord(done[s], stream));
    return CCFV_OK;
}

int PhaseRunner::issue(const Phase &ph, const char *d)
{
    // fork: the lanes in use start after what `stream` holds (the upload)
    bool used[LANES] = {};
    for (const Launch &L : ph.launches) used[L.lane] = true;
    HIPCHK(hipEventRecord(fork_ev, stream));
    for (int l = 1; l < LANES; l++)
        if (used[l]) HIPCHK(hipStreamWaitEvent(lanes[l], fork_ev, 0));
    int rc = CCFV_OK;
    for (const Launch &L : ph.launches) {
        hipError_t e = hipSuccess;
        const hipStream_t ls = lanes[L.lane];
        const int *pf = (const int *)(d + L.prefix);
        switch (L.kind) {
        case Launch::SWEEP: {
            const RaggedItem *it = (const RaggedItem *)(d + L.items);
            e = L.fast ? launch_ragged_fast(L.dir, L.epi, L.ch, L.wpl, L.fuse, it, pf,
                                            L.nitems, L.nblocks, L.tile, L.lds, ls)
                       : launch_ragged_generic(L.dir, L.bc, L.post, L.epi, L.ch, L.fuse, it,
                                               pf, L.nitems, L.nblocks, L.tile, L.lds,
                                               ls);
            break;
        }
        case Launch::CORR:
            e = launch_ragged_correction((const RaggedCorr *)(d + L.items), pf, L.nitems,
                                         L.nblocks, ls);
            break;
        case Launch::REDUCE:
            e = launch_ragged_reduce((const RaggedReduce *)(d + L.items), pf, L.nitems,
                                     L.nblocks, ls);
            break;
        case Launch::COPY:
            e = launch_ragged_copy((const RaggedCopy *)(d + L.items), pf, L.nitems,
                                   L.nblocks, ls);
            break;
        }
        if (e != hipSuccess) {
            rc = hip_fail(e, "ragged launch");
            break;
        }
    }
    // join: `stream` continues after every lane's launches (also after a
    // failed launch, so that no lane is left outside a capture)
    for (int l = 1; l < LANES; l++)
        if (used[l]) {
            hipError_t e = hipEventRecord(join_ev[l], lanes[l]);
            if (e == hipSuccess) e = hipStreamWaitEvent(stream, join_ev[l], 0);
            if (e != hipSuccess && rc == CCFV_OK) rc = hip_fail(e, "phase lane join");
        }
    return rc;
}

// ---------------------------------------------------------------------------
// runs of Strang steps as stage items
// ---------------------------------------------------------------------------
namespace {

StageItem stage_item(ccfv_ctx *c, Dir dir, int substep, const SolveTable &T,
                     const StepDesc &d, int shape)
{
    StageItem s;
    s.lane = c->lane;
    s.dir = dir;
    s.bc = c->P.bc;
    s.post = dir == DIR_X && substep == 3 && post_flag(c);
    s.epi = substep == 3 ? d.epi : EPI_NONE;
    s.fuse = false;
    s.tile = dir == DIR_X && s.epi != EPI_NONE ? c->chain_tile : c->tile[dir][shape];
    SweepArgs &a = s.item.a;
    a = make_args(c, dir, substep, T, d.h, d.t0, d.tDT, d.field0, d.jdt);
    a.in = substep == 1 ? d.in : d.out;
    a.out = d.out;
    a.fstride = (long long)c->ncell;
    if (substep == 3) {
        a.g_cache = d.g_cache;
        a.f_in = d.f_in;
        a.u_next = d.u_next;
        a.partial = d.partial;
    }
    s.item.nfields = d.nf;
    s.item.ntiles = (dir == DIR_X ? c->P.ny : c->P.nx) / s.tile.w;
    return s;
}

bool fusable(const StepDesc &d, const StepDesc &n)
{
    const double *res = d.epi == EPI_NONE ? d.out : d.u_next;
    return n.in == res && n.h == d.h && n.nf == d.nf && n.tDT == d.tDT;
}

}  // namespace

int run_rounds(const std::vector<StepRun> &runs)
{
    int r = 0;
    for (const StepRun &u : runs)
        if (u.count > 0) r = std::max(r, u.count + 1);
    return r;
}

int emit_runs(Phase &ph, const std::vector<StepRun> &runs, int r0, int r1)
{
    // stage 0: x launches (x1 of step 0, fused x3(r-1)+x1(r), last x3);
    // stage 1: x1 of steps that cannot be fused with their predecessor;
    // stage 2: y sweeps
    std::vector<StageItem> st[3];
    int rc;
    for (int r = r0; r < r1; r++) {
        for (const StepRun &u : runs) {
            if (r > u.count || u.count <= 0) continue;
            ccfv_ctx *c = u.c;
            const SolveTable *tx, *ty;
            if (r == 0) {
                StepDesc d = u.desc(0);
                if ((rc = get_table(c, DIR_X, d.h, &tx))) return rc;
                st[0].push_back(stage_item(c, DIR_X, 1, *tx, d, u.shape));
            } else {
                StepDesc d = u.desc(r - 1);
                if ((rc = get_table(c, DIR_X, d.h, &tx))) return rc;
                StageItem x3 = stage_item(c, DIR_X, 3, *tx, d, u.shape);
                if (r < u.count) {
                    StepDesc n = u.desc(r);
                    if (fusable(d, n)) {
                        x3.fuse = true;
                        x3.item.a.out2 = n.out;
                        x3.item.a.field0_2 = n.field0;
                        x3.item.a.t0_2 = n.t0;
                        x3.item.a.jdt2 = n.jdt;
                    } else {
                        const SolveTable *tn;
                        if ((rc = get_table(c, DIR_X, n.h, &tn))) return rc;
                        st[1].push_back(stage_item(c, DIR_X, 1, *tn, n, u.shape));
                    }
                }
                st[0].push_back(x3);
            }
            if (r < u.count) {
                StepDesc d = u.desc(r);
                if ((rc = get_table(c, DIR_Y, d.h, &ty))) return rc;
                st[2].push_back(stage_item(c, DIR_Y, 2, *ty, d, u.shape));
            }
        }
        for (int i = 0; i < 3; i++) {
            if (!st[i].empty()) ph.add_sweep_stage(st[i]);
            st[i].clear();
        }
    }
    return CCFV_OK;
}

// ---------------------------------------------------------------------------
// Parareal jobs
// ---------------------------------------------------------------------------
int job_setup(ParaJob &j, ccfv_ctx *c, double t0, double t1, int nt_fine, int ns,
              int kmax, double tol)
{
    if (ns < 1 || nt_fine < ns || nt_fine % ns)
        return fail(CCFV_EINVAL, "nt_fine must be a positive multiple of n_slices");
    if (kmax < 0 || kmax > CCFV_MAX_ITERS) return fail(CCFV_EINVAL, "kmax must be in [0, 64]");
    if (!(t1 > t0)) return fail(CCFV_EINVAL, "need t1 > t0");
    j = ParaJob();
    j.c = c;
    j.ns = ns;
    j.J = nt_fine / ns;
    j.t0 = t0;
    j.t1 = t1;
    j.nt_fine = nt_fine;
    j.serial_finish = c->serial_finish;
    j.DT = (t1 - t0) / ns;
    j.dt = j.DT / (double)j.J;
    j.nb = c->chain_blocks;
    // iteration k fixes slot k+1; past n_slices iterations nothing changes
    j.kmax = std::min(kmax, ns);
    j.tol = tol;
    int rc;
    if ((rc = get_table(c, DIR_X, j.dt, &j.fx)) || (rc = get_table(c, DIR_Y, j.dt, &j.fy)) ||
        (rc = get_table(c, DIR_X, j.DT, &j.cx)) || (rc = get_table(c, DIR_Y, j.DT, &j.cy)))
        return rc;
    j.slot_u2.assign(ns + 1, 0.0);
    memset(&j.R, 0, sizeof(j.R));
    j.R.n_slices = ns;
    j.R.j_fine = j.J;
    j.k = 0;
    j.started = false;
    j.running = j.kmax > 0;
    j.status = CCFV_OK;
    return CCFV_OK;
}

int job_setup_plain(ParaJob &j, ccfv_ctx *c, double t0, double t1, int nt)
{
    if (nt < 1) return fail(CCFV_EINVAL, "nt must be >= 1");
    if (!(t1 > t0)) return fail(CCFV_EINVAL, "need t1 > t0");
    j = ParaJob();
    j.c = c;
    j.plain = true;
    j.nt = nt;
    j.t0 = t0;
    j.dt = (t1 - t0) / nt;  // as ccfv_plain_solve
    int rc;
    if ((rc = get_table(c, DIR_X, j.dt, &j.fx)) || (rc = get_table(c, DIR_Y, j.dt, &j.fy)))
        return rc;
    memset(&j.R, 0, sizeof(j.R));
    j.started = true;
    j.running = true;
    return CCFV_OK;
}

namespace {

// the steps of a job (the schedule of ccfv_parareal_solve)
StepDesc job_init_step(const ParaJob &p, int n)
{
    const size_t N = p.c->ncell;
    return StepDesc{p.c->U + n * N, p.c->Wk, 1, p.t0, p.DT, n, 0.0, p.DT, EPI_INIT,
                    p.c->G + n * N, nullptr, p.c->U + (n + 1) * N, nullptr};
}

StepDesc job_fine_step(const ParaJob &p, int j)
{
    const size_t N = p.c->ncell;
    const double *src = j == 0 ? p.c->U + p.k * N : p.c->F + p.k * N;
    return StepDesc{src, p.c->F + p.k * N, p.ns - p.k, p.t0, p.DT, p.k, j * p.dt, p.dt};
}

StepDesc job_chain_step(const ParaJob &p, int n)
{
    const size_t N = p.c->ncell;
    return StepDesc{p.c->U + n * N, p.c->Wk, 1, p.t0, p.DT, n, 0.0, p.DT, EPI_CORRECT,
                    p.c->G + n * N, p.c->F + n * N, p.c->U + (n + 1) * N,
                    p.c->partial + 2 * (size_t)p.nb * (n + 1)};
}

int emit_all(Phase &ph, const std::vector<StepRun> &runs)
{
    return emit_runs(ph, runs, 0, run_rounds(runs));
}

}  // namespace

int append_iteration(Phase &ph, std::vector<ParaJob *> &jobs, double *sums_dev,
                     size_t *used_out)
{
    int rc;
    // plain jobs: their next steps, tn = t0 + n dt (T2S4_Parareal.cu:1234-1236)
    std::vector<StepRun> runs;
    for (ParaJob *p : jobs) {
        if (!p->plain || !p->running) continue;
        const int n0 = p->done, cnt = std::min(PLAIN_CHUNK, p->nt - p->done);
        runs.push_back(StepRun{p->c, cnt, 0, [p, n0](int i) {
                                   const int n = n0 + i;
                                   return StepDesc{p->c->U, p->c->U, 1, p->t0 + n * p->dt, 0.0, 0,
                                                   0.0, p->dt};
                               }});
    }
    if ((rc = emit_all(ph, runs))) return rc;
    runs.clear();
    // k = 0 coarse sweep: U[n+1] = G(U[n]), G cached (T2S4_Parareal.cu:1264-1280)
    for (ParaJob *p : jobs)
        if (!p->started)
            runs.push_back(StepRun{p->c, p->ns, 0, [p](int n) { return job_init_step(*p, n); }});
    if ((rc = emit_all(ph, runs))) return rc;
    for (ParaJob *p : jobs)
        if (!p->started) {
            p->started = true;
            p->R.coarse_steps_run = p->ns;
        }
    // iterating jobs this phase
    std::vector<ParaJob *> it;
    for (ParaJob *p : jobs)
        if (!p->plain && p->running && p->k < p->kmax) it.push_back(p);
    // fine propagators of slices k..ns-1, all slices of a job as one batch
    // of fields (T2S4_Parareal.cu:1319-1324)
    runs.clear();
    for (ParaJob *p : it)
        runs.push_back(StepRun{p->c, p->J, 1, [p](int j) { return job_fine_step(*p, j); }});
    if ((rc = emit_all(ph, runs))) return rc;
    // slot k+1 by the correction alone (T2S4_Parareal.cu:1327-1330)
    std::vector<RaggedCorr> corr;
    std::vector<int> cblocks;
    for (ParaJob *p : it) {
        const size_t N = p->c->ncell;
        const int k = p->k;
        corr.push_back(RaggedCorr{p->c->G + k * N, p->c->G + k * N, p->c->F + k * N,
                                  p->c->U + (k + 1) * N, N,
                                  p->c->partial + 2 * (size_t)p->nb * (k + 1)});
        cblocks.push_back(p->nb);
    }
    std::vector<int> lanes;
    for (ParaJob *p : it) lanes.push_back(p->c->lane);
    ph.add_corrections(corr, cblocks, lanes);
    // coarse chain from slice k+1 with the correction fused into its last
    // sweep
    runs.clear();
    for (ParaJob *p : it)
        runs.push_back(StepRun{p->c, p->ns - p->k - 1, 0,
                               [p](int i) { return job_chain_step(*p, p->k + 1 + i); }});
    if ((rc = emit_all(ph, runs))) return rc;
    // per-slot increment sums of slots k+1..ns
    std::vector<RaggedReduce> red;
    size_t used = 0;
    for (ParaJob *p : it) {
        p->sums_off = used;
        used += 2 * (size_t)(p->ns + 1);
        red.push_back(RaggedReduce{p->c->partial, sums_dev + p->sums_off, p->nb, p->k + 1,
                                   p->ns - p->k});
        p->R.fine_steps_run += (double)(p->ns - p->k) * p->J;
        p->R.coarse_steps_run += p->ns - p->k - 1;
    }
    ph.add_reductions(red, lanes);
    *used_out = used;
    return CCFV_OK;
}

void finish_iteration(ParaJob &p, const double *sums_host)
{
    if (p.plain) {
        p.done = std::min(p.nt, p.done + PLAIN_CHUNK);
        p.R.fine_steps_run = p.para_fine + p.done;
        if (p.done >= p.nt) {
            p.R.converged = 1;  // the serial fine solution itself
            p.running = false;
        }
        return;
    }
    const double *h = sums_host + p.sums_off;
    const int k = p.k;
    double num = 0.0, den = 0.0;
    for (int s = 1; s <= p.ns; s++) {
        if (s >= k + 1) {
            num += h[2 * s];
            p.slot_u2[s] = h[2 * s + 1];
        }
        den += p.slot_u2[s];
    }
    const double inc = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
    p.R.increment[k] = inc;
    p.R.iterations = k + 1;
    p.k = k + 1;
    if (!std::isfinite(inc)) {
        p.running = false;
        p.status = CCFV_ENONFINITE;
        return;
    }
    if (p.tol > 0 && inc <= p.tol) {
        p.R.converged = 1;
        p.running = false;
    }
    if (p.k >= p.ns) {  // every slice exact
        p.R.converged = 1;
        p.running = false;
    }
    if (p.k >= p.kmax) p.running = false;
    if (!p.running && !p.R.converged && p.tol > 0 && p.serial_finish) {
        // the serial fine solve from the initial state in U[0]
        p.plain = true;
        p.nt = p.nt_fine;
        p.done = 0;
        p.dt = (p.t1 - p.t0) / p.nt;  // as ccfv_plain_solve
        p.para_fine = p.R.fine_steps_run;
        p.R.serial_finish = 1;
        p.running = true;
    }
}

}  // namespace rt
}  // namespace ccfv
