// CPU check of the solve server's admission bookkeeping (bola_amd/csrc/serve_cache.hpp),
// built and run by tests/test_serve_cache.py.  Device memory is stood in for by tagged
// host pointers; the release callback records what was freed.
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

#include "serve_cache.hpp"

using bola::resident_factor_bytes;
using bola::ServeCache;

static int failures = 0;
#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            std::printf("FAIL line %d: %s\n", __LINE__, #c);          \
            failures++;                                               \
        }                                                             \
    } while (0)

static std::set<void*> freed;
static void release(void* f, void* p) {
    if (f) freed.insert(f);
    if (p) freed.insert(p);
}
static void* tag(long k) { return (void*)(0x1000 + 16 * k); }

// One request's life: reserve, (factor), finish; returns false if it did not fit.
static bool run(ServeCache& c, uint64_t key, size_t n, size_t nrhs, bool& hit, long id, bool keep = true) {
    size_t reserved = 0;
    if (!c.reserve(key, 0, n, n * nrhs * 8, resident_factor_bytes(n), reserved, hit)) return false;
    if (hit) {
        bola::ResidentFactor* f = c.pinned(key);
        CHECK(f != nullptr);
        if (f) CHECK(freed.count(f->f) == 0);  // the factorization it reads is alive
        c.finish(reserved, false, key, 0, n, 0, nullptr, nullptr);
        c.unpin(key);
    } else {
        const bool stored =
            c.finish(reserved, keep && key != 0, key, 0, n, resident_factor_bytes(n), tag(2 * id), tag(2 * id + 1));
        if (!stored) release(tag(2 * id), tag(2 * id + 1));
    }
    return true;
}

int main() {
    const size_t n = 256;
    const size_t f = resident_factor_bytes(n);
    {
        // two resident factorizations fill the budget; a hit on the least recently used
        // one with a right-hand side that needs an eviction must evict the other one
        const size_t nrhs = 40;  // 80 KiB of right-hand side
        ServeCache c(2 * f + n * nrhs * 8 / 2, release);
        bool hit = false;
        CHECK(run(c, 1, n, 1, hit, 1) && !hit);
        CHECK(run(c, 2, n, 1, hit, 2) && !hit);
        CHECK(c.size() == 2 && c.in_use() == 2 * f);
        // key 1 is the least recently used; its hit needs room for 80 KiB
        CHECK(run(c, 1, n, nrhs, hit, 3) && hit);
        CHECK(c.evictions() == 1);
        CHECK(c.lookup(1, 0, n) != nullptr);  // its own factorization stayed
        CHECK(c.lookup(2, 0, n) == nullptr);  // the other one went
        CHECK(freed.count(tag(2 * 1)) == 0 && freed.count(tag(2 * 2)) == 1);
        CHECK(c.in_use() == f);
    }
    freed.clear();
    {
        // a hit that cannot fit even after evicting every other idle factorization is
        // refused and leaves its own factorization resident and unpinned
        ServeCache c(f + 4096, release);
        bool hit = false;
        CHECK(run(c, 7, n, 1, hit, 1) && !hit);
        size_t reserved = 0;
        CHECK(!c.reserve(7, 0, n, n * 64 * 8, f, reserved, hit));  // 128 KiB of right-hand side
        CHECK(!hit);
        CHECK(c.lookup(7, 0, n) != nullptr && c.lookup(7, 0, n)->users == 0);
        CHECK(freed.empty());
    }
    freed.clear();
    {
        // pinned factorizations are never evicted: two hits in flight on both entries,
        // and a miss that needs the room must wait
        ServeCache c(2 * f + 4096, release);
        bool hit = false;
        CHECK(run(c, 1, n, 1, hit, 1));
        CHECK(run(c, 2, n, 1, hit, 2));
        size_t r1 = 0, r2 = 0, r3 = 0;
        bool h1 = false, h2 = false, h3 = false;
        CHECK(c.reserve(1, 0, n, n * 8, f, r1, h1) && h1);
        CHECK(c.reserve(2, 0, n, n * 8, f, r2, h2) && h2);
        CHECK(!c.reserve(3, 0, n, n * 8, f, r3, h3));  // nothing idle to evict
        CHECK(freed.empty());
        c.finish(r1, false, 1, 0, n, 0, nullptr, nullptr);
        c.unpin(1);
        CHECK(c.reserve(3, 0, n, n * 8, f, r3, h3) && !h3);  // now key 1 can go
        CHECK(c.lookup(1, 0, n) == nullptr && c.lookup(2, 0, n) != nullptr);
        c.finish(r2, false, 2, 0, n, 0, nullptr, nullptr);
        c.unpin(2);
        c.finish(r3, false, 3, 0, n, f, tag(9), tag(10));
        CHECK(c.in_use() == f);
    }
    freed.clear();
    {
        // a factorization of another kind or order under the same key is not a hit,
        // and keeping it replaces the idle stale entry
        ServeCache c(8 * f, release);
        bool hit = false;
        CHECK(run(c, 5, n, 1, hit, 1) && !hit);
        size_t reserved = 0;
        CHECK(c.reserve(5, 1, n, n * 8, f, reserved, hit) && !hit);
        CHECK(c.finish(reserved, true, 5, 1, n, f, tag(20), tag(21)));
        CHECK(freed.count(tag(2)) == 1);
        CHECK(c.lookup(5, 1, n) != nullptr && c.lookup(5, 0, n) == nullptr);
        CHECK(c.in_use() == f);
        c.clear();
        CHECK(c.in_use() == 0 && freed.count(tag(20)) == 1);
    }
    freed.clear();
    {
        // random traffic: the budget holds, every hit reads a live factorization, and
        // everything is returned at the end
        ServeCache c(6 * f, release);
        unsigned s = 12345;
        auto rnd = [&] { s = s * 1103515245u + 12345u; return (s >> 8) & 0xffff; };
        long id = 100;
        for (int it = 0; it < 20000; it++) {
            const uint64_t key = rnd() % 4 == 0 ? 0 : 1 + rnd() % 12;
            const size_t nrhs = 1 + rnd() % 64;
            bool hit = false;
            run(c, key, n, nrhs, hit, id++);
            CHECK(c.in_use() <= c.budget());
            CHECK(c.peak() <= c.budget());
        }
        CHECK(c.in_use() == c.cached_bytes());
        c.clear();
        CHECK(c.in_use() == 0);
    }
    if (failures == 0) std::printf("serve_cache: all checks passed\n");
    return failures == 0 ? 0 : 1;
}
