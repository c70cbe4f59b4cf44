// ccfv_batch.h — batch of independent problems (ccfv_batch.hip) and the
// phase builder shared with the serving loop.
#pragma once

#include <map>
#include <tuple>
#include <vector>

#include "ccfv_ctx.h"

namespace ccfv {
namespace rt {

// one problem's contribution to one sweep stage
struct StageItem {
    Dir dir;
    int bc;
    bool post;
    Epi epi;
    Tile tile;
    RaggedItem item;
};

// one ragged launch of a phase; items / prefix are byte offsets into the
// phase's descriptor arena
struct Launch {
    enum Kind { SWEEP, CORR, REDUCE } kind;
    Dir dir = DIR_X;
    Tile tile{0, 0};
    int ch = 1, bc = 0;
    bool fast = false, post = false;
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
    void add_corrections(const std::vector<RaggedCorr> &c, const std::vector<int> &blocks);
    void add_reductions(const std::vector<RaggedReduce> &r);
};

}  // namespace rt
}  // namespace ccfv

struct ccfv_batch {
    std::vector<ccfv_ctx *> ctx;
    hipStream_t stream = nullptr;
    // two descriptor arenas (pinned host + device), alternating by phase
    char *hbuf[2] = {nullptr, nullptr};
    char *dbuf[2] = {nullptr, nullptr};
    size_t cap[2] = {0, 0};
    hipEvent_t arena_ev[2] = {nullptr, nullptr};
    int slot = 0;
    // per-slot increment sums of every problem
    double *sums_dev = nullptr, *sums_host = nullptr;
    size_t sums_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};
