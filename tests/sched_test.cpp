// Unit test of the serving engine's host bookkeeping (ccfv_sched.h): built
// with g++ and run by tests/test_sched.py.  Prints "ok" and exits 0 on
// success; otherwise names the failed check and exits 1.
#include <cstdio>
#include <cstdlib>

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

int main()
{
    arena_first_fit_and_merge();
    arena_fragmentation();
    footprint();
    admission_order_and_slots();
    admission_passes_over_then_reserves();
    std::printf("ok\n");
    return 0;
}
