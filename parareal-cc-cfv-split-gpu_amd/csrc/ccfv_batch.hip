// ccfv_batch.hip — many independent problems of mixed sizes solved together.
//
// Every problem keeps its own context (state, tables, Parareal buffers), but
// the work is issued in lock-step rounds: in round r every problem that still
// has an r-th sub-step contributes its blocks to one ragged launch per sweep
// stage and line class (ccfv_sweep_*.hip, `ragged_*` kernels), located by a
// block-prefix array.  The per-problem schedule is the one of
// ccfv_parareal_solve (init coarse sweep; per iteration the fine propagators
// of the slices that can still change, the correction of slot k+1 and the
// coarse chain with the fused correction; per-slot increment sums), so each
// problem's result is bitwise the one it gets alone.  Descriptors of a whole
// phase are built on the host and uploaded with one copy.
#include "ccfv_batch.h"

#include <algorithm>
#include <cmath>
#include <cstring>

using namespace ccfv;
using namespace ccfv::rt;

namespace ccfv {
namespace rt {

namespace {

// common launch shapes for a batch (valid for every even grid that is a
// multiple of 8 lines; W for the coarse chain fixes the partial-sum layout)
void batch_tiles(ccfv_ctx *c)
{
    const bool post = post_flag(c);
    const Tile x1{2, 2}, y1{2, 2}, yb{2, 1};
    Tile xb{8, 4};
    if (!tile_ok(xb, c->P.nx, c->P.ny, DIR_X, post)) xb = Tile{2, 2};
    c->tile[DIR_X][0] = x1;
    c->tile[DIR_X][1] = xb;
    c->tile[DIR_Y][0] = y1;
    c->tile[DIR_Y][1] = yb;
    c->tiles_version++;
    clear_graphs(c);
}

}  // namespace

// ---------------------------------------------------------------------------
// phase builder: descriptors on the host, one upload, then the launches
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

// Sweep items of one stage, grouped by line class into launches.
void Phase::add_sweep_stage(std::vector<StageItem> &items)
{
    // one launch shares the kernel instance (line chunk, boundary mode, path,
    // step-3 variant, epilogue) and the block shape
    struct Key {
        int ch, bc, fast, post, epi, w, nw;
        bool operator<(const Key &o) const
        {
            return std::tie(ch, bc, fast, post, epi, w, nw) <
                   std::tie(o.ch, o.bc, o.fast, o.post, o.epi, o.w, o.nw);
        }
    };
    std::map<Key, std::vector<const StageItem *>> groups;
    for (const StageItem &s : items) {
        Key k{chunk_for(s.item.a.n), s.bc, s.item.a.lineS ? 1 : 0, s.post ? 1 : 0,
              (int)s.epi, s.tile.w, s.tile.nw};
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
                              : sweep_lds_bytes_generic(s->item.a.n, s->tile);
            lds = std::max(lds, b);
        }
        Launch L;
        L.kind = Launch::SWEEP;
        L.dir = kv.second.front()->dir;
        L.tile = kv.second.front()->tile;
        L.ch = k.ch;
        L.bc = k.bc;
        L.fast = k.fast != 0;
        L.post = k.post != 0;
        L.epi = (Epi)k.epi;
        L.nitems = (int)it.size();
        L.nblocks = prefix.back();
        L.lds = lds;
        L.items = append(it.data(), sizeof(RaggedItem) * it.size());
        L.prefix = append(prefix.data(), sizeof(int) * prefix.size());
        launches.push_back(L);
    }
}

void Phase::add_corrections(const std::vector<RaggedCorr> &c, const std::vector<int> &blocks)
{
    if (c.empty()) return;
    std::vector<int> prefix(1, 0);
    for (int b : blocks) prefix.push_back(prefix.back() + b);
    Launch L;
    L.kind = Launch::CORR;
    L.nitems = (int)c.size();
    L.nblocks = prefix.back();
    L.items = append(c.data(), sizeof(RaggedCorr) * c.size());
    L.prefix = append(prefix.data(), sizeof(int) * prefix.size());
    launches.push_back(L);
}

void Phase::add_reductions(const std::vector<RaggedReduce> &r)
{
    if (r.empty()) return;
    std::vector<int> prefix(1, 0);
    for (const RaggedReduce &x : r) prefix.push_back(prefix.back() + x.count);
    Launch L;
    L.kind = Launch::REDUCE;
    L.nitems = (int)r.size();
    L.nblocks = prefix.back();
    L.items = append(r.data(), sizeof(RaggedReduce) * r.size());
    L.prefix = append(prefix.data(), sizeof(int) * prefix.size());
    launches.push_back(L);
}

}  // namespace rt
}  // namespace ccfv

namespace {

// upload the phase's descriptors and issue its launches on the batch stream
int run_phase(ccfv_batch *b, Phase &ph)
{
    if (ph.launches.empty()) return CCFV_OK;
    const size_t bytes = ph.host.size();
    int slot = b->slot;
    b->slot ^= 1;
    // the arena of this slot was last used two phases ago
    HIPCHK(hipEventSynchronize(b->arena_ev[slot]));
    if (bytes > b->cap[slot]) {
        (void)hipHostFree(b->hbuf[slot]);
        (void)hipFree(b->dbuf[slot]);
        b->hbuf[slot] = nullptr;
        b->dbuf[slot] = nullptr;
        size_t cap = std::max(bytes, b->cap[slot] * 2);
        HIPCHK(hipHostMalloc(&b->hbuf[slot], cap));
        HIPCHK(hipMalloc(&b->dbuf[slot], cap));
        b->cap[slot] = cap;
    }
    memcpy(b->hbuf[slot], ph.host.data(), bytes);
    HIPCHK(hipMemcpyAsync(b->dbuf[slot], b->hbuf[slot], bytes, hipMemcpyHostToDevice,
                          b->stream));
    char *d = b->dbuf[slot];
    for (const Launch &L : ph.launches) {
        hipError_t e = hipSuccess;
        if (L.kind == Launch::SWEEP) {
            const RaggedItem *it = (const RaggedItem *)(d + L.items);
            const int *pf = (const int *)(d + L.prefix);
            e = L.fast ? launch_ragged_fast(L.dir, L.epi, L.ch, it, pf, L.nitems,
                                            L.nblocks, L.tile, L.lds, b->stream)
                       : launch_ragged_generic(L.dir, L.bc, L.post, L.epi, L.ch, it, pf,
                                               L.nitems, L.nblocks, L.tile, L.lds,
                                               b->stream);
        } else if (L.kind == Launch::CORR) {
            e = launch_ragged_correction((const RaggedCorr *)(d + L.items),
                                         (const int *)(d + L.prefix), L.nitems,
                                         L.nblocks, b->stream);
        } else {
            e = launch_ragged_reduce((const RaggedReduce *)(d + L.items),
                                     (const int *)(d + L.prefix), L.nitems, L.nblocks,
                                     b->stream);
        }
        if (e != hipSuccess) return hip_fail(e, "ragged launch");
    }
    HIPCHK(hipEventRecord(b->arena_ev[slot], b->stream));
    return CCFV_OK;
}

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

// the three stages of one Strang step of one problem
struct StepSpec {
    ccfv_ctx *c;
    const SolveTable *tx, *ty;
    const double *in;
    double *out;
    int nf;
    double t0, tDT, jdt, h;
    int field0;
    Epi epi;
    double *g_cache;
    const double *f_in;
    double *u_next;
    double *partial;
    int shape;  // 1: batch launch shapes (fine propagators), 0: one-field
};

void add_step_items(const StepSpec &s, std::vector<StageItem> st[3])
{
    ccfv_ctx *c = s.c;
    // fine propagators use the batch shapes; the coarse chain and the plain
    // sweep the one-field shapes, which also fix the epilogue's partial sums
    const int bt = s.shape;
    const long long fs = (long long)c->ncell;
    StageItem a;
    a.dir = DIR_X;
    a.bc = c->P.bc;
    a.post = false;
    a.epi = EPI_NONE;
    a.tile = c->tile[DIR_X][bt];
    a.item.a = make_args(c, DIR_X, 1, *s.tx, s.h, s.t0, s.tDT, s.field0, s.jdt);
    a.item.a.in = s.in;
    a.item.a.out = s.out;
    a.item.a.fstride = fs;
    a.item.nfields = s.nf;
    a.item.ntiles = c->P.ny / a.tile.w;
    st[0].push_back(a);
    StageItem b;
    b.dir = DIR_Y;
    b.bc = c->P.bc;
    b.post = false;
    b.epi = EPI_NONE;
    b.tile = c->tile[DIR_Y][bt];
    b.item.a = make_args(c, DIR_Y, 2, *s.ty, s.h, s.t0, s.tDT, s.field0, s.jdt);
    b.item.a.in = s.out;
    b.item.a.out = s.out;
    b.item.a.fstride = fs;
    b.item.nfields = s.nf;
    b.item.ntiles = c->P.nx / b.tile.w;
    st[1].push_back(b);
    StageItem x;
    x.dir = DIR_X;
    x.bc = c->P.bc;
    x.post = post_flag(c);
    x.epi = s.epi;
    x.tile = c->tile[DIR_X][bt];
    x.item.a = make_args(c, DIR_X, 3, *s.tx, s.h, s.t0, s.tDT, s.field0, s.jdt);
    x.item.a.in = s.out;
    x.item.a.out = s.out;
    x.item.a.fstride = fs;
    x.item.a.g_cache = s.g_cache;
    x.item.a.f_in = s.f_in;
    x.item.a.u_next = s.u_next;
    x.item.a.partial = s.partial;
    x.item.nfields = s.nf;
    x.item.ntiles = c->P.ny / x.tile.w;
    st[2].push_back(x);
}

void flush_round(Phase &ph, std::vector<StageItem> st[3])
{
    for (int i = 0; i < 3; i++) {
        if (!st[i].empty()) ph.add_sweep_stage(st[i]);
        st[i].clear();
    }
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
        batch_tiles(c);
        b->ctx.push_back(c);
    }
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->arena_ev[0], hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->arena_ev[1], hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess) {
        ccfv_batch_destroy(b);
        return hip_fail(e, "batch stream/events");
    }
    *out = b;
    return CCFV_OK;
}

void ccfv_batch_destroy(ccfv_batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (ccfv_ctx *c : b->ctx) ccfv_destroy(c);
    for (int s = 0; s < 2; s++) {
        if (b->hbuf[s]) (void)hipHostFree(b->hbuf[s]);
        if (b->dbuf[s]) (void)hipFree(b->dbuf[s]);
        if (b->arena_ev[s]) (void)hipEventDestroy(b->arena_ev[s]);
    }
    (void)hipFree(b->sums_dev);
    if (b->sums_host) (void)hipHostFree(b->sums_host);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
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
    if (!b || !t0 || !t1 || !nt) return fail(CCFV_EINVAL, "NULL argument");
    const int np = (int)b->ctx.size();
    int rc;
    if ((rc = join_contexts(b))) return rc;
    std::vector<double> dt(np);
    std::vector<const SolveTable *> tx(np), ty(np);
    int ntmax = 0;
    for (int p = 0; p < np; p++) {
        if (nt[p] < 1 || !(t1[p] > t0[p])) return fail(CCFV_EINVAL, "bad time interval");
        dt[p] = (t1[p] - t0[p]) / nt[p];
        if ((rc = get_table(b->ctx[p], DIR_X, dt[p], &tx[p])) ||
            (rc = get_table(b->ctx[p], DIR_Y, dt[p], &ty[p])))
            return rc;
        ntmax = std::max(ntmax, (int)nt[p]);
    }
    // phases of 64 rounds keep the descriptor upload small
    for (int r0 = 0; r0 < ntmax; r0 += 64) {
        Phase ph;
        for (int r = r0; r < std::min(ntmax, r0 + 64); r++) {
            std::vector<StageItem> st[3];
            for (int p = 0; p < np; p++) {
                if (r >= nt[p]) continue;
                ccfv_ctx *c = b->ctx[p];
                StepSpec s{c, tx[p], ty[p], c->state, c->state, 1, t0[p] + r * dt[p],
                           0.0, 0.0, dt[p], 0, EPI_NONE, nullptr, nullptr, nullptr,
                           nullptr, 0};
                add_step_items(s, st);
            }
            flush_round(ph, st);
        }
        if ((rc = run_phase(b, ph))) return rc;
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    for (ccfv_ctx *c : b->ctx)
        if ((rc = check_flag(c))) return rc;
    return CCFV_OK;
}

int ccfv_batch_parareal(ccfv_batch *b, const double *t0, const double *t1,
                        const int32_t *nt_fine, const int32_t *n_slices, int32_t kmax,
                        double tol, ccfv_parareal_report *reports)
{
    if (!b || !t0 || !t1 || !nt_fine || !n_slices)
        return fail(CCFV_EINVAL, "NULL argument");
    if (kmax < 0 || kmax > CCFV_MAX_ITERS) return fail(CCFV_EINVAL, "kmax must be in [0, 64]");
    const int np = (int)b->ctx.size();
    int rc;
    if ((rc = join_contexts(b))) return rc;
    struct P {
        ccfv_ctx *c;
        int ns, J, nb;
        double t0, DT, dt;
        const SolveTable *fx, *fy, *cx, *cy;
        bool running;
        size_t sums_off;
        std::vector<double> slot_u2;
        ccfv_parareal_report R;
    };
    std::vector<P> ps(np);
    size_t sums_total = 0;
    int nsmax = 0, Jmax = 0;
    for (int i = 0; i < np; i++) {
        P &p = ps[i];
        p.c = b->ctx[i];
        p.ns = n_slices[i];
        if (p.ns < 1 || nt_fine[i] < p.ns || nt_fine[i] % p.ns || !(t1[i] > t0[i]))
            return fail(CCFV_EINVAL, "problem " + std::to_string(i) +
                                         ": nt_fine must be a positive multiple of n_slices");
        p.J = nt_fine[i] / p.ns;
        p.t0 = t0[i];
        p.DT = (t1[i] - t0[i]) / p.ns;
        p.dt = p.DT / (double)p.J;
        if ((rc = ensure_para(p.c, p.ns))) return rc;
        p.nb = p.c->chain_blocks;
        if ((rc = get_table(p.c, DIR_X, p.dt, &p.fx)) || (rc = get_table(p.c, DIR_Y, p.dt, &p.fy)) ||
            (rc = get_table(p.c, DIR_X, p.DT, &p.cx)) || (rc = get_table(p.c, DIR_Y, p.DT, &p.cy)))
            return rc;
        p.running = kmax > 0;
        p.sums_off = sums_total;
        sums_total += 2 * (size_t)(p.ns + 1);
        p.slot_u2.assign(p.ns + 1, 0.0);
        memset(&p.R, 0, sizeof(p.R));
        p.R.n_slices = p.ns;
        p.R.j_fine = p.J;
        nsmax = std::max(nsmax, p.ns);
        Jmax = std::max(Jmax, p.J);
    }
    if (sums_total > b->sums_cap) {
        (void)hipFree(b->sums_dev);
        if (b->sums_host) (void)hipHostFree(b->sums_host);
        b->sums_dev = nullptr;
        b->sums_host = nullptr;
        HIPCHK(hipMalloc(&b->sums_dev, sizeof(double) * sums_total));
        HIPCHK(hipHostMalloc(&b->sums_host, sizeof(double) * sums_total));
        b->sums_cap = sums_total;
    }
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    for (P &p : ps) HIPCHK(launch_copy(p.c->state, p.c->U, p.c->ncell, b->stream));

    // k = 0: U[n+1] = G(U[n]), G cached (T2S4_Parareal.cu:1264-1280)
    {
        Phase ph;
        for (int n = 0; n < nsmax; n++) {
            std::vector<StageItem> st[3];
            for (P &p : ps) {
                if (n >= p.ns) continue;
                const size_t N = p.c->ncell;
                StepSpec s{p.c, p.cx, p.cy, p.c->U + n * N, p.c->Wk, 1, p.t0, p.DT,
                           0.0, p.DT, n, EPI_INIT, p.c->G + n * N, nullptr,
                           p.c->U + (n + 1) * N, nullptr, 0};
                add_step_items(s, st);
            }
            flush_round(ph, st);
        }
        if ((rc = run_phase(b, ph))) return rc;
        for (P &p : ps) p.R.coarse_steps_run = p.ns;
    }

    for (int k = 0; k < kmax; k++) {
        bool any = false;
        for (P &p : ps) any = any || (p.running && k < p.ns);
        if (!any) break;
        Phase ph;
        // fine propagators of slices k..ns-1 (T2S4_Parareal.cu:1319-1324)
        for (int j = 0; j < Jmax; j++) {
            std::vector<StageItem> st[3];
            for (P &p : ps) {
                if (!p.running || k >= p.ns || j >= p.J) continue;
                const size_t N = p.c->ncell;
                const double *src = j == 0 ? p.c->U + k * N : p.c->F + k * N;
                StepSpec s{p.c, p.fx, p.fy, src, p.c->F + k * N, p.ns - k, p.t0, p.DT,
                           j * p.dt, p.dt, k, EPI_NONE, nullptr, nullptr, nullptr,
                           nullptr, 1};
                add_step_items(s, st);
            }
            flush_round(ph, st);
        }
        // slot k+1 by the correction alone; coarse chain from slice k+1
        std::vector<RaggedCorr> corr;
        std::vector<int> cblocks;
        for (P &p : ps) {
            if (!p.running || k >= p.ns) continue;
            const size_t N = p.c->ncell;
            corr.push_back(RaggedCorr{p.c->G + k * N, p.c->G + k * N, p.c->F + k * N,
                                      p.c->U + (k + 1) * N, N,
                                      p.c->partial + 2 * (size_t)p.nb * (k + 1)});
            cblocks.push_back(p.nb);
        }
        ph.add_corrections(corr, cblocks);
        for (int n = k + 1; n < nsmax; n++) {
            std::vector<StageItem> st[3];
            for (P &p : ps) {
                if (!p.running || n >= p.ns) continue;
                const size_t N = p.c->ncell;
                StepSpec s{p.c, p.cx, p.cy, p.c->U + n * N, p.c->Wk, 1, p.t0, p.DT, 0.0,
                           p.DT, n, EPI_CORRECT, p.c->G + n * N, p.c->F + n * N,
                           p.c->U + (n + 1) * N, p.c->partial + 2 * (size_t)p.nb * (n + 1),
                           0};
                add_step_items(s, st);
            }
            flush_round(ph, st);
        }
        std::vector<RaggedReduce> red;
        for (P &p : ps) {
            if (!p.running || k >= p.ns) continue;
            red.push_back(RaggedReduce{p.c->partial, b->sums_dev + p.sums_off, p.nb, k + 1,
                                       p.ns - k});
        }
        ph.add_reductions(red);
        if ((rc = run_phase(b, ph))) return rc;
        HIPCHK(hipMemcpyAsync(b->sums_host, b->sums_dev, sizeof(double) * sums_total,
                              hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        for (P &p : ps) {
            if (!p.running || k >= p.ns) continue;
            const double *h = b->sums_host + p.sums_off;
            double num = 0.0, den = 0.0;
            for (int s = 1; s <= p.ns; s++) {
                if (s >= k + 1) {
                    num += h[2 * s];
                    p.slot_u2[s] = h[2 * s + 1];
                }
                den += p.slot_u2[s];
            }
            double inc = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
            p.R.increment[k] = inc;
            p.R.iterations = k + 1;
            p.R.fine_steps_run += (double)(p.ns - k) * p.J;
            p.R.coarse_steps_run += p.ns - k - 1;
            if (!std::isfinite(inc)) {
                p.running = false;
                continue;
            }
            if (tol > 0 && inc <= tol) {
                p.R.converged = 1;
                p.running = false;
            }
            if (k + 1 >= p.ns) {  // every slice exact
                p.R.converged = 1;
                p.running = false;
            }
        }
    }
    for (P &p : ps)
        HIPCHK(launch_copy(p.c->U + (size_t)p.ns * p.c->ncell, p.c->state, p.c->ncell,
                           b->stream));
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    int status = CCFV_OK;
    for (int i = 0; i < np; i++) {
        P &p = ps[i];
        p.R.wall_ms = ms;
        if (reports) reports[i] = p.R;
        if ((rc = check_flag(p.c))) status = rc;
        for (int k = 0; k < p.R.iterations; k++)
            if (!std::isfinite(p.R.increment[k]) && status == CCFV_OK)
                status = fail(CCFV_ENONFINITE, "problem " + std::to_string(i) +
                                                   ": Parareal iterate is not finite");
    }
    if (status) return status;
    if (tol > 0)
        for (int i = 0; i < np; i++)
            if (!ps[i].R.converged)
                return fail(CCFV_ENOTCONV, "problem " + std::to_string(i) +
                                               " did not reach the tolerance within kmax");
    return CCFV_OK;
}

}  // extern "C"
