// The serving loop's chunk pool (anime4kcpp_amd/csrc/chunk_pool.hpp) over a fake backend
// that hands out ranges of one host arena, so chunks can be made adjacent on purpose and
// allocation failures forced.  Sizes are in units of a 64-byte block, the first chunk 16
// blocks.  Exits non-zero and names the failed check on any mismatch.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "chunk_pool.hpp"

namespace {

int failures = 0;
#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            failures++;                                                          \
        }                                                                        \
    } while (0)

constexpr size_t B = 64;   // block

struct Arena {
    std::vector<char> mem = std::vector<char>(1 << 20);
    size_t top = 0;            // bump pointer: consecutive chunks are adjacent
    size_t limit = 1 << 20;    // alloc fails beyond this many bytes handed out at once
    size_t out = 0;
    int allocs = 0, releases = 0;
    std::map<void*, size_t> live;
};

struct FakeBackend {
    Arena* a;
    bool alloc(void** p, size_t n)
    {
        if (a->out + n > a->limit || a->top + n > a->mem.size()) return false;
        *p = a->mem.data() + a->top;
        a->top += n;
        a->out += n;
        a->live[*p] = n;
        a->allocs++;
        return true;
    }
    void release(void* p)
    {
        auto it = a->live.find(p);
        if (it == a->live.end()) {
            std::fprintf(stderr, "release of a pointer the backend did not hand out\n");
            failures++;
            return;
        }
        a->out -= it->second;
        a->live.erase(it);
        a->releases++;
    }
};
using Pool = achip::ChunkPool<FakeBackend>;

struct Span {
    char* p;
    size_t n;
};

bool disjoint(const std::vector<Span>& v)
{
    std::map<char*, size_t> m;
    for (const Span& s : v) m[s.p] = s.n;
    char* end = nullptr;
    for (const auto& kv : m) {
        if (end && kv.first < end) return false;
        end = kv.first + kv.second;
    }
    return m.size() == v.size();
}

// every handed-out block lies inside one chunk the backend gave out
bool inside_backend(const Arena& a, const std::vector<Span>& v)
{
    for (const Span& s : v) {
        auto it = a.live.upper_bound(static_cast<void*>(s.p));
        if (it == a.live.begin()) return false;
        --it;
        char* base = static_cast<char*>(it->first);
        if (s.p + s.n > base + it->second) return false;
    }
    return true;
}

void test_rounding_growth_and_reuse()
{
    Arena a;
    {
        Pool pool(FakeBackend{ &a }, B, 16 * B);
        pool.set_chunk_cap(64 * B);
        std::vector<Span> v;
        for (size_t n : { 1, 65, 128, 1000, 64 * 5 + 1 }) {
            void* p = nullptr;
            CHECK(pool.get(n, &p));
            v.push_back({ static_cast<char*>(p), pool.round(n) });
            CHECK((static_cast<char*>(p) - a.mem.data()) % B == 0);
        }
        CHECK(disjoint(v) && inside_backend(a, v));
        CHECK(pool.live_blocks() == 5);
        CHECK(a.allocs == 2);                     // 27 blocks: two chunks of 16
        // each further chunk is as large as everything held so far: few backend allocations
        for (int i = 0; i < 20; i++) {
            void* p = nullptr;
            CHECK(pool.get(3 * B, &p));
            v.push_back({ static_cast<char*>(p), 3 * B });
        }
        CHECK(disjoint(v) && inside_backend(a, v));
        CHECK(a.allocs <= 6);
        CHECK(pool.held() == a.out);
        // a freed block is reused for a request of its size, without the backend
        const int before = a.allocs;
        pool.put(v[3].p);
        void* p = nullptr;
        CHECK(pool.get(1000, &p) && p == v[3].p);
        CHECK(a.allocs == before);
        for (const Span& s : v) pool.put(s.p);
        CHECK(pool.live_blocks() == 0);
        CHECK(pool.free_ranges() == pool.chunks());   // everything coalesced back to whole chunks
        pool.trim();
        CHECK(pool.chunks() == 0 && pool.held() == 0 && a.out == 0);
    }
    CHECK(a.releases == a.allocs);
}

void test_coalescing_and_best_fit()
{
    Arena a;
    Pool pool(FakeBackend{ &a }, B, 16 * B);
    void *p1, *p2, *p3, *p4;
    CHECK(pool.get(4 * B, &p1) && pool.get(2 * B, &p2) && pool.get(4 * B, &p3) && pool.get(6 * B, &p4));
    CHECK(a.allocs == 1 && pool.free_ranges() == 0);   // 16 blocks exactly
    pool.put(p2);                                       // hole of 2
    pool.put(p4);                                       // hole of 6 at the end
    CHECK(pool.free_ranges() == 2);
    void* q = nullptr;
    CHECK(pool.get(2 * B, &q) && q == p2);              // best fit: the 2-block hole, not the 6
    pool.put(q);
    pool.put(p3);                                       // 2 + 4 + 6 merge into one range of 12
    CHECK(pool.free_ranges() == 1);
    CHECK(pool.get(12 * B, &q) && q == p2 && a.allocs == 1);
    pool.put(q);
    pool.put(p1);                                       // the whole chunk
    CHECK(pool.free_ranges() == 1);
    pool.trim();
    CHECK(pool.chunks() == 0 && a.out == 0);
}

void test_adjacent_chunks_stay_apart()
{
    Arena a;
    Pool pool(FakeBackend{ &a }, B, 16 * B);
    pool.set_chunk_cap(16 * B);                         // every chunk 16 blocks
    void *p1, *p2;
    CHECK(pool.get(16 * B, &p1) && pool.get(16 * B, &p2));
    CHECK(a.allocs == 2 && static_cast<char*>(p2) == static_cast<char*>(p1) + 16 * B);   // adjacent in memory
    pool.put(p1);
    pool.put(p2);
    CHECK(pool.free_ranges() == 2);                     // not merged across the chunk boundary
    void* q = nullptr;
    a.limit = 16 * B;                                   // the backend gives at most 16 blocks at once
    CHECK(!pool.get(32 * B, &q));                       // no 32-block range: trim both chunks, then fail
    CHECK(pool.chunks() == 0 && a.out == 0 && a.releases == 2);
}

void test_failure_trims_then_allocates_exactly()
{
    Arena a;
    Pool pool(FakeBackend{ &a }, B, 16 * B);
    pool.set_chunk_cap(1 << 20);
    void *p1, *p2;
    CHECK(pool.get(10 * B, &p1));                       // chunk of 16
    CHECK(pool.get(10 * B, &p2));                       // 6 left: a second chunk of 16
    CHECK(a.allocs == 2);
    pool.put(p2);                                       // second chunk wholly free
    a.limit = a.out + 4 * B;                            // room for 4 more blocks beyond what is out
    void* q = nullptr;
    // the geometric next chunk (32 blocks, as much as is held) would not fit; the pool trims
    // the free chunk (16 back) and asks for exactly 20 blocks, which now fits
    CHECK(pool.get(20 * B, &q));
    CHECK(a.releases == 1 && pool.chunks() == 2);
    pool.put(q);
    pool.put(p1);
    pool.put(p1);                                       // a second put of the same block is ignored
    pool.put(nullptr);
    pool.trim();
    CHECK(pool.chunks() == 0 && a.out == 0);
}

void test_random_churn()
{
    Arena a;
    Pool pool(FakeBackend{ &a }, B, 16 * B);
    pool.set_chunk_cap(256 * B);
    std::mt19937 rng(7);
    std::vector<Span> live;
    for (int step = 0; step < 4000; step++) {
        if (live.empty() || rng() % 3 != 0) {
            const size_t n = 1 + rng() % (8 * B);
            void* p = nullptr;
            if (!pool.get(n, &p)) continue;
            live.push_back({ static_cast<char*>(p), pool.round(n) });
        } else {
            const size_t i = rng() % live.size();
            pool.put(live[i].p);
            live.erase(live.begin() + static_cast<long>(i));
        }
        if (step % 97 == 0) CHECK(disjoint(live) && inside_backend(a, live));
    }
    CHECK(disjoint(live) && inside_backend(a, live));
    for (const Span& s : live) pool.put(s.p);
    CHECK(pool.free_ranges() == pool.chunks());
    pool.trim();
    CHECK(pool.chunks() == 0 && a.out == 0 && a.releases == a.allocs);
}

} // namespace

int main()
{
    test_rounding_growth_and_reuse();
    test_coalescing_and_best_fit();
    test_adjacent_chunks_stay_apart();
    test_failure_trims_then_allocates_exactly();
    test_random_churn();
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("pool_test: all checks passed\n");
    return 0;
}
