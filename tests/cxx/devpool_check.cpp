// Host check of the device allocator's pool logic (block-dda_amd/csrc/devmem_pool.hpp)
// over malloc with a capacity limit: best-fit reuse, splitting and merging,
// trimming on out-of-memory, content integrity under random alloc/free, and
// the same under four threads.  Prints "ok" and exits 0 when every check holds.
#include "../../block-dda_amd/csrc/devmem_pool.hpp"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <thread>
#include <vector>

namespace {

struct MallocRaw {
    static constexpr int kOom = 2;
    std::atomic<size_t>* used;
    size_t cap;
    std::map<void*, size_t>* sizes;
    std::mutex* m;
    int alloc(size_t bytes, void** p)
    {
        std::lock_guard<std::mutex> g(*m);
        if (used->load() + bytes > cap) return kOom;
        *p = std::aligned_alloc(4096, bytes);
        if (!*p) return 1;
        (*sizes)[*p] = bytes;
        *used += bytes;
        return 0;
    }
    void release(void* p)
    {
        std::lock_guard<std::mutex> g(*m);
        *used -= (*sizes)[p];
        sizes->erase(p);
        std::free(p);
    }
    size_t headroom = 0;
    bool low() { return used->load() + headroom > cap; }
};

int failures = 0;
#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            ++failures;                                                       \
        }                                                                     \
    } while (0)

struct Env {
    std::atomic<size_t> used{0};
    std::map<void*, size_t> sizes;
    std::mutex m;
    MallocRaw raw(size_t cap, size_t headroom = 0) { return MallocRaw{&used, cap, &sizes, &m, headroom}; }
};

constexpr size_t MiB = 1 << 20;

void basic()
{
    Env env;
    using Pool = bdda::CachingPool<MallocRaw>;
    Pool pool(env.raw(1024 * MiB));
    int err = 0;
    CHECK(pool.alloc(0, false, &err) == nullptr && err == 0);
    // medium requests share 128 MiB raw allocations: best fit, split, merge
    void* a = pool.alloc(5 * MiB + 3, false, &err);
    CHECK(a && err == 0);
    CHECK(pool.stats().held == 128 * MiB && pool.stats().raw_allocs == 1);
    CHECK(reinterpret_cast<uintptr_t>(a) % 512 == 0);
    CHECK(pool.free(a));
    auto idle = pool.idle_blocks();
    CHECK(idle.size() == 1 && idle[0].first == a && idle[0].second == 128 * MiB);
    void* b = pool.alloc(4 * MiB, false, &err);
    CHECK(b == a && pool.stats().raw_allocs == 1 && pool.stats().reuses == 1);
    void* c = pool.alloc(2 * MiB, false, &err);
    CHECK(c == static_cast<char*>(a) + 4 * MiB && pool.stats().raw_allocs == 1);
    CHECK(pool.free(b) && pool.free(c));
    idle = pool.idle_blocks();
    CHECK(idle.size() == 1 && idle[0].second == 128 * MiB);
    // small requests come from 2 MiB raw allocations of their own
    void* s1 = pool.alloc(1000, false, &err);
    void* s2 = pool.alloc(512, false, &err);
    CHECK(s1 && s2 && pool.stats().raw_allocs == 2);
    CHECK(static_cast<char*>(s2) == static_cast<char*>(s1) + 1024);
    CHECK(pool.stats().held == 130 * MiB);
    // kept large requests: one raw allocation each, never split, handed out
    // again for a request of at least 16/17 of the block
    void* l1 = pool.alloc(100 * MiB, true, &err);
    CHECK(l1 && pool.stats().raw_allocs == 3 && pool.stats().held == 230 * MiB);
    CHECK(pool.free(l1));
    void* l2 = pool.alloc(95 * MiB, true, &err);
    CHECK(l2 == l1 && pool.stats().raw_allocs == 3);
    CHECK(pool.stats().in_use == 100 * MiB + 1024 + 512);
    CHECK(pool.free(l2));
    void* l3 = pool.alloc(90 * MiB, true, &err);  // 100 MiB is more than 17/16 of it
    CHECK(l3 && l3 != l1 && pool.stats().raw_allocs == 4);
    CHECK(pool.free(l3));
    // a medium request never takes a large block, nor a small one a medium block
    void* m = pool.alloc(64 * MiB, false, &err);
    CHECK(m == a);
    CHECK(pool.free(m));
    // scratch large requests: best fit, split and merged
    void* t1 = pool.alloc(100 * MiB, false, &err);
    CHECK(t1 && pool.stats().raw_allocs == 5 && pool.stats().held == 420 * MiB);
    CHECK(pool.free(t1));
    void* t2 = pool.alloc(65 * MiB, false, &err);
    CHECK(t2 == t1 && pool.stats().raw_allocs == 5);
    void* t3 = pool.alloc(66 * MiB, false, &err);  // more than the 35 MiB left
    CHECK(t3 && pool.stats().raw_allocs == 6);
    void* t4 = pool.alloc(34 * MiB, false, &err);  // medium: never from a large block
    CHECK(t4 && (char*)t4 >= (char*)a && (char*)t4 < (char*)a + 128 * MiB);
    CHECK(pool.free(t2) && pool.free(t3) && pool.free(t4));
    // a kept request does not take an idle scratch block, nor the other way round
    void* k1 = pool.alloc(100 * MiB, true, &err);
    CHECK(k1 == l1);
    void* t5 = pool.alloc(99 * MiB, false, &err);
    CHECK(t5 == t1 && pool.stats().raw_allocs == 6);
    CHECK(pool.free(k1) && pool.free(t5));
    // a pointer that is not the pool's is refused, unchanged
    int x = 0;
    CHECK(!pool.free(&x));
    CHECK(pool.free(s1) && pool.free(s2));
    // trim returns every wholly idle raw allocation
    CHECK(pool.trim() == 130 * MiB + 190 * MiB + 166 * MiB && pool.stats().held == 0 && env.used.load() == 0);
    CHECK(pool.stats().raw_releases == 6);
}

void out_of_memory()
{
    Env env;
    bdda::CachingPool<MallocRaw> pool(env.raw(256 * MiB));
    int err = 0;
    void* a = pool.alloc(100 * MiB, true, &err);
    void* b = pool.alloc(100 * MiB, true, &err);
    CHECK(a && b);
    CHECK(pool.free(a) && pool.free(b));  // a is the least recently freed
    // 150 MiB fits neither idle block: the raw allocator is out of memory, the
    // pool gives back a (then 150 MiB fits) and keeps b
    void* c = pool.alloc(150 * MiB, true, &err);
    CHECK(c && err == 0 && pool.stats().raw_releases == 1 && pool.stats().held == 250 * MiB);
    auto idle = pool.idle_blocks();
    CHECK(idle.size() == 1 && idle[0].first == b);
    // beyond the capacity even after giving everything idle back
    void* d = pool.alloc(200 * MiB, false, &err);
    CHECK(d == nullptr && err == MallocRaw::kOom && pool.stats().held == 150 * MiB);
    // c idle: a 128 MiB medium allocation fits only once c went back
    CHECK(pool.free(c));
    void* e = pool.alloc(3 * MiB, false, &err);
    CHECK(e && pool.stats().held == 128 * MiB && pool.stats().raw_releases == 3);
    // a live medium block keeps its raw allocation from being trimmed
    CHECK(pool.trim() == 0 && pool.stats().held == 128 * MiB);
    CHECK(pool.free(e));
    CHECK(pool.trim() == 128 * MiB && env.used.load() == 0);
}

// after a raw allocation leaves less than the headroom free, idle raw
// allocations go back, least recently freed first
void headroom()
{
    Env env;
    bdda::CachingPool<MallocRaw> pool(env.raw(600 * MiB, 128 * MiB));
    int err = 0;
    void* a = pool.alloc(200 * MiB, true, &err);
    void* b = pool.alloc(100 * MiB, true, &err);
    CHECK(a && b && pool.free(a) && pool.free(b));
    void* c = pool.alloc(250 * MiB, true, &err);  // fits, but 550 MiB held is over 600 - 128
    CHECK(c && pool.stats().raw_releases == 1 && pool.stats().held == 350 * MiB);
    auto idle = pool.idle_blocks();
    CHECK(idle.size() == 1 && idle[0].first == b);
    CHECK(pool.free(c));
    CHECK(pool.trim() == 350 * MiB);
}

// a block's first and last 4 KiB and one byte in every 64 KiB between
void fill(char* p, size_t n, unsigned char tag)
{
    std::memset(p, tag, std::min<size_t>(n, 4096));
    std::memset(p + n - std::min<size_t>(n, 4096), tag, std::min<size_t>(n, 4096));
    for (size_t k = 0; k < n; k += 65536) p[k] = (char)tag;
}

bool intact(const char* p, size_t n, unsigned char tag)
{
    for (size_t k = 0; k < std::min<size_t>(n, 4096); ++k)
        if ((unsigned char)p[k] != tag || (unsigned char)p[n - 1 - k] != tag) return false;
    for (size_t k = 0; k < n; k += 65536)
        if ((unsigned char)p[k] != tag) return false;
    return true;
}

// every live block holds a fill pattern of its own; the bytes of all live
// blocks plus the idle ones add up to what the pool holds, and no two live
// blocks overlap
void random_sequence(unsigned seed, size_t cap, int steps)
{
    Env env;
    bdda::CachingPool<MallocRaw> pool(env.raw(cap));
    std::mt19937_64 rng(seed);
    struct Live {
        char* p;
        size_t n;
        unsigned char tag;
    };
    std::vector<Live> live;
    for (int it = 0; it < steps; ++it) {
        const bool do_alloc = live.empty() || (rng() % 100) < 55;
        if (do_alloc) {
            size_t n;
            switch (rng() % 5) {
            case 0: n = 1 + rng() % 4096; break;
            case 1: n = 1 + rng() % MiB; break;
            case 2: n = MiB + rng() % (8 * MiB); break;
            case 3: n = 8 * MiB + rng() % (56 * MiB); break;
            default: n = 64 * MiB + rng() % (96 * MiB); break;
            }
            int err = 0;
            char* p = static_cast<char*>(pool.alloc(n, rng() % 2 == 0, &err));
            if (!p) {
                CHECK(err == MallocRaw::kOom);
                continue;
            }
            const unsigned char tag = (unsigned char)(rng() % 251 + 1);
            fill(p, n, tag);
            for (const Live& l : live) CHECK(p + n <= l.p || l.p + l.n <= p);
            live.push_back({p, n, tag});
        } else {
            const size_t i = rng() % live.size();
            const Live l = live[i];
            CHECK(intact(l.p, l.n, l.tag));
            CHECK(pool.free(l.p));
            live[i] = live.back();
            live.pop_back();
        }
        if (it % 64 == 0) {
            size_t idle = 0;
            for (auto& b : pool.idle_blocks()) idle += b.second;
            const auto st = pool.stats();
            CHECK(st.in_use + idle == st.held);
            CHECK(st.held == env.used.load());
        }
        if (it % 500 == 499) pool.trim();
    }
    for (const Live& l : live) CHECK(pool.free(l.p));
    pool.trim();
    CHECK(pool.stats().held == 0 && env.used.load() == 0);
}

void threads()
{
    Env env;
    bdda::CachingPool<MallocRaw> pool(env.raw(2048 * MiB));
    std::atomic<int> bad{0};
    auto work = [&](unsigned seed) {
        std::mt19937_64 rng(seed);
        std::vector<std::pair<char*, size_t>> mine;
        for (int it = 0; it < 3000; ++it) {
            if (mine.empty() || rng() % 2) {
                const size_t r = rng() % 3;
                const size_t n = r == 0 ? 1 + rng() % 4096 : r == 1 ? MiB + rng() % (6 * MiB) : 64 * MiB + rng() % (32 * MiB);
                int err = 0;
                char* p = static_cast<char*>(pool.alloc(n, rng() % 2 == 0, &err));
                if (!p) continue;
                fill(p, n, (unsigned char)((seed & 0x7f) + 1));
                mine.emplace_back(p, n);
            } else {
                const size_t i = rng() % mine.size();
                auto [p, n] = mine[i];
                if (!intact(p, n, (unsigned char)((seed & 0x7f) + 1))) ++bad;
                if (!pool.free(p)) ++bad;
                mine[i] = mine.back();
                mine.pop_back();
            }
        }
        for (auto& m : mine)
            if (!pool.free(m.first)) ++bad;
    };
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < 4; ++t) ts.emplace_back(work, 100 + t);
    for (auto& t : ts) t.join();
    CHECK(bad.load() == 0);
    CHECK(pool.stats().in_use == 0);
    pool.trim();
    CHECK(pool.stats().held == 0 && env.used.load() == 0);
}

}  // namespace

int main()
{
    basic();
    out_of_memory();
    headroom();
    random_sequence(1, 768 * MiB, 6000);
    random_sequence(2, 320 * MiB, 6000);   // tight: out-of-memory and trimming on the way
    random_sequence(3, 4096 * MiB, 4000);
    threads();
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
