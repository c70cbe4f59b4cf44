// Unit test of the serving engine's host bookkeeping (ccfv_sched.h): built
// with g++ and run by tests/test_sched.py.  Prints "ok" and exits 0 on
// success; otherwise names the failed check and exits 1.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ccfv_sched.h"

using namespace ccfv::sched;

#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            std::printf("FAILED line %d: %s\n", __LINE__, #c);     \
            std::exit(1);                                          \
        }                                                          \
    } while (0)

static void arena_first_fit_and_merge()
{
    ArenaAlloc a(4096, 256);
    CHECK(a.size() == 4096 && a.used() == 0 && a.free_blocks() == 1);
    size_t x = a.alloc(100);   // rounds to 256
    size_t y = a.alloc(512);
    size_t z = a.alloc(1000);  // rounds to 1024
    CHECK(x == 0 && y == 256 && z == 768);
    CHECK(a.used() == 256 + 512 + 1024);
    CHECK(a.largest_free() == 4096 - 1792);
    a.release(y, 512);  // hole [256, 768)
    CHECK(a.free_blocks() == 2);
    size_t w = a.alloc(300);  // first fit: the hole
    CHECK(w == 256);
    a.release(w, 300);
    a.release(x, 100);  // merges with the hole to [0, 768)
    CHECK(a.free_blocks() == 2 && a.largest_free() == 4096 - 1792);
    a.release(z, 1000);  // everything free again, one block
    CHECK(a.free_blocks() == 1 && a.largest_free() == 4096 && a.used() == 0);
    CHECK(a.alloc(5000) == NPOS);
    CHECK(!a.fits(4097) && a.fits(4096));
}

static void arena_fragmentation()
{
    ArenaAlloc a(1024, 256);
    size_t b[4];
    for (int i = 0; i < 4; i++) b[i] = a.alloc(256);
    a.release(b[0], 256);
    a.release(b[2], 256);
    CHECK(a.used() == 512 && a.largest_free() == 256);
    CHECK(a.alloc(512) == NPOS);  // 512 free, but not contiguous
    a.release(b[1], 256);
    CHECK(a.largest_free() == 768);
    CHECK(a.alloc(512) == 0);
}

static void footprint()
{
    // 64x64, 8 slices, 32 chain blocks: U 9 fields, G and F 8 each, work
    // field 1, partials 2*32*9 doubles
    const size_t f = 8 * 64 * 64;
    size_t want = 9 * f + 2 * 8 * f + f + round_up(8 * 2 * 32 * 9, 256);
    CHECK(request_bytes(64, 64, 8, 32) == want);
    CHECK(request_bytes(64, 128, 8, 64) > request_bytes(64, 64, 8, 32));
}

static void admission_order_and_slots()
{
    ArenaAlloc a(4096, 256);
    AdmissionQueue q(8);
    q.push(1, 1024);
    q.push(2, 1024);
    q.push(3, 1024);
    auto g = q.admit(a, 2);  // two slots
    CHECK(g.size() == 2 && g[0].id == 1 && g[1].id == 2 && q.size() == 1);
    CHECK(g[0].offset == 0 && g[1].offset == 1024);
    g = q.admit(a, 5);
    CHECK(g.size() == 1 && g[0].id == 3 && q.empty());
}

static void admission_passes_over_then_reserves()
{
    ArenaAlloc a(4096, 256);
    size_t held = a.alloc(2048);  // a resident request
    AdmissionQueue q(2);
    q.push(10, 4096);  // cannot fit until `held` is released
    q.push(11, 512);
    q.push(12, 512);
    q.push(13, 512);
    auto g = q.admit(a, 100);
    // 11 and 12 pass the blocked head; then the head has been passed over
    // twice and 13 must wait behind it
    CHECK(g.size() == 2 && g[0].id == 11 && g[1].id == 12 && q.size() == 2);
    g = q.admit(a, 100);
    CHECK(g.empty());
    a.release(held, 2048);
    a.release(2048, 512);  // 11 finishes
    a.release(2560, 512);  // 12 finishes
    g = q.admit(a, 100);
    CHECK(g.size() == 1 && g[0].id == 10 && g[0].offset == 0);
    CHECK(q.size() == 1);  // 13 waits: the arena is full
}

// xorshift64*: a seeded stream of pseudo-random numbers, the same on every
// platform
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return s * 0x2545F4914F6CDD1DULL;
    }
    size_t below(size_t n) { return (size_t)(next() % n); }
};

// The allocator against a unit-by-unit model of the arena: every alloc
// returns the lowest offset where enough free units start (first fit), the
// used bytes, the number of maximal free runs (released neighbours always
// merge) and the largest run agree after every operation.
static void arena_against_model(uint64_t seed)
{
    const size_t align = 256, units = 64;
    ArenaAlloc a(units * align, align);
    std::vector<int> owner(units, 0);  // 0 free, else the block id
    struct Live {
        size_t off, bytes;
        int id;
    };
    std::vector<Live> live;
    Rng r{seed};
    int next_id = 1;
    for (int step = 0; step < 4000; step++) {
        if (live.empty() || r.below(3) != 0) {
            const size_t bytes = 1 + r.below(12 * align);
            const size_t k = (bytes + align - 1) / align;
            size_t want = NPOS;
            for (size_t o = 0; o + k <= units && want == NPOS; o++) {
                bool ok = true;
                for (size_t u = o; u < o + k && ok; u++) ok = owner[u] == 0;
                if (ok) want = o * align;
            }
            const size_t got = a.alloc(bytes);
            CHECK(got == want);
            if (got != NPOS) {
                for (size_t u = got / align; u < got / align + k; u++) owner[u] = next_id;
                live.push_back(Live{got, bytes, next_id++});
            }
        } else {
            const size_t i = r.below(live.size());
            a.release(live[i].off, live[i].bytes);
            for (size_t u = 0; u < units; u++)
                if (owner[u] == live[i].id) owner[u] = 0;
            live.erase(live.begin() + (long)i);
        }
        size_t used = 0, runs = 0, run = 0, largest = 0;
        for (size_t u = 0; u < units; u++) {
            if (owner[u]) {
                used += align;
                run = 0;
            } else {
                if (run == 0) runs++;
                run++;
                largest = run > largest ? run : largest;
            }
        }
        CHECK(a.used() == used);
        CHECK(a.free_blocks() == runs);
        CHECK(a.largest_free() == largest * align);
    }
}

// Admission against its contract over random arrivals, sizes, slot counts
// and completions: grants come in arrival order among the requests granted
// in one call, every grant is a block the arena handed out, no more grants
// than free slots, and a waiting request is passed over by later arrivals at
// most max_bypass times while it is the oldest one waiting.
static void admission_against_contract(uint64_t seed)
{
    const size_t align = 256;
    const int max_bypass = 3;
    ArenaAlloc a(32 * align, align);
    AdmissionQueue q(max_bypass);
    Rng r{seed};
    struct Req {
        size_t bytes;
        size_t off = NPOS;
        int passed = 0;  // later arrivals granted while it was the oldest waiting
        bool waiting = true;
    };
    std::vector<Req> reqs;
    std::vector<int64_t> resident;
    std::vector<int64_t> waiting;  // arrival order
    for (int step = 0; step < 3000; step++) {
        const int arrivals = (int)r.below(3);
        for (int i = 0; i < arrivals; i++) {
            Req x;
            x.bytes = 1 + r.below(20 * align);
            reqs.push_back(x);
            q.push((int64_t)reqs.size() - 1, x.bytes);
            waiting.push_back((int64_t)reqs.size() - 1);
        }
        const int slots = (int)r.below(4);
        const int64_t oldest = waiting.empty() ? -1 : waiting.front();
        auto g = q.admit(a, slots);
        CHECK((int)g.size() <= slots);
        int64_t prev = -1;
        for (auto &gr : g) {
            CHECK(gr.id > prev);  // arrival order
            prev = gr.id;
            Req &x = reqs[(size_t)gr.id];
            CHECK(x.waiting);
            x.waiting = false;
            x.off = gr.offset;
            resident.push_back(gr.id);
            if (oldest >= 0 && gr.id != oldest && reqs[(size_t)oldest].waiting)
                reqs[(size_t)oldest].passed++;
        }
        if (oldest >= 0) CHECK(reqs[(size_t)oldest].passed <= max_bypass);
        waiting.erase(std::remove_if(waiting.begin(), waiting.end(),
                                     [&](int64_t id) { return !reqs[(size_t)id].waiting; }),
                      waiting.end());
        CHECK(q.size() == waiting.size());
        // granted blocks do not overlap
        std::vector<int> seen(32, 0);
        for (int64_t id : resident) {
            const Req &x = reqs[(size_t)id];
            for (size_t u = x.off / align; u < (x.off + x.bytes + align - 1) / align; u++) {
                CHECK(u < 32 && seen[u] == 0);
                seen[u] = 1;
            }
        }
        // some residents finish
        for (size_t i = 0; i < resident.size();) {
            if (r.below(3) == 0) {
                const Req &x = reqs[(size_t)resident[i]];
                a.release(x.off, x.bytes);
                resident.erase(resident.begin() + (long)i);
            } else {
                i++;
            }
        }
    }
}

int main()
{
    arena_first_fit_and_merge();
    arena_fragmentation();
    footprint();
    admission_order_and_slots();
    admission_passes_over_then_reserves();
    for (uint64_t seed = 1; seed <= 8; seed++) {
        arena_against_model(0x9E3779B97F4A7C15ULL * seed);
        admission_against_contract(0xD1B54A32D192ED03ULL * seed);
    }
    std::printf("ok\n");
    return 0;
}
