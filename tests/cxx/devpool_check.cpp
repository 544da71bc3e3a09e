// Host check of the device allocator's pool logic (block-dda_amd/csrc/devmem_pool.hpp)
// over malloc with a capacity limit: best-fit reuse, splitting and merging,
// trimming on out-of-memory, content integrity under random alloc/free, and
// the same under four threads.  Prints "ok" and exits 0 when every check holds.
#include "../../block-dda_amd/csrc/devmem_pool.hpp"

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
    MallocRaw raw(size_t cap) { return MallocRaw{&used, cap, &sizes, &m}; }
};

constexpr size_t MiB = 1 << 20;

void basic()
{
    Env env;
    bdda::CachingPool<MallocRaw> pool(env.raw(64 * MiB));
    int err = 0;
    CHECK(pool.alloc(0, &err) == nullptr && err == 0);
    // a large request gets its own raw allocation, rounded to 2 MiB
    void* a = pool.alloc(5 * MiB + 3, &err);
    CHECK(a && err == 0);
    CHECK(pool.stats().held == 6 * MiB && pool.stats().raw_allocs == 1);
    CHECK(reinterpret_cast<uintptr_t>(a) % 512 == 0);
    // freed and asked again: no new raw allocation
    CHECK(pool.free(a));
    void* b = pool.alloc(4 * MiB, &err);
    CHECK(b == a && pool.stats().raw_allocs == 1 && pool.stats().reuses == 1);
    // the rest (2 MiB >= the split minimum) was split off and is handed out next
    void* c = pool.alloc(2 * MiB, &err);
    CHECK(c == static_cast<char*>(a) + 4 * MiB && pool.stats().raw_allocs == 1);
    CHECK(pool.idle_blocks().empty());
    // merging: free both, one idle block of the whole raw allocation
    CHECK(pool.free(b) && pool.free(c));
    auto idle = pool.idle_blocks();
    CHECK(idle.size() == 1 && idle[0].first == a && idle[0].second == 6 * MiB);
    // small requests come from 2 MiB raw allocations of their own
    void* s1 = pool.alloc(1000, &err);
    void* s2 = pool.alloc(512, &err);
    CHECK(s1 && s2 && pool.stats().raw_allocs == 2);
    CHECK(static_cast<char*>(s2) == static_cast<char*>(s1) + 1024);
    CHECK(pool.stats().held == 8 * MiB);
    // a pointer that is not the pool's is refused, unchanged
    int x = 0;
    CHECK(!pool.free(&x));
    CHECK(pool.free(s1) && pool.free(s2));
    // trim returns every wholly idle raw allocation
    CHECK(pool.trim() == 8 * MiB && pool.stats().held == 0 && env.used.load() == 0);
    CHECK(pool.stats().raw_releases == 2);
}

void out_of_memory()
{
    Env env;
    bdda::CachingPool<MallocRaw> pool(env.raw(16 * MiB));
    int err = 0;
    void* a = pool.alloc(10 * MiB, &err);
    CHECK(a);
    CHECK(pool.free(a));
    // 12 MiB does not fit beside the idle 10 MiB: the pool trims and retries
    void* b = pool.alloc(12 * MiB, &err);
    CHECK(b && err == 0 && pool.stats().raw_releases == 1 && pool.stats().held == 12 * MiB);
    // beyond the capacity even after trimming: the raw code comes back
    void* c = pool.alloc(8 * MiB, &err);
    CHECK(c == nullptr && err == MallocRaw::kOom);
    // a live split block keeps its raw allocation: not trimmed
    CHECK(pool.free(b));
    void* d = pool.alloc(4 * MiB, &err);
    CHECK(d == b);
    CHECK(pool.trim() == 0 && pool.stats().held == 12 * MiB);
    CHECK(pool.free(d));
    CHECK(pool.trim() == 12 * MiB);
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
            switch (rng() % 4) {
            case 0: n = 1 + rng() % 4096; break;
            case 1: n = 1 + rng() % MiB; break;
            case 2: n = MiB + rng() % (8 * MiB); break;
            default: n = 8 * MiB + rng() % (24 * MiB); break;
            }
            int err = 0;
            char* p = static_cast<char*>(pool.alloc(n, &err));
            if (!p) {
                CHECK(err == MallocRaw::kOom);
                continue;
            }
            const unsigned char tag = (unsigned char)(rng() % 251 + 1);
            std::memset(p, tag, n);
            for (const Live& l : live) CHECK(p + n <= l.p || l.p + l.n <= p);
            live.push_back({p, n, tag});
        } else {
            const size_t i = rng() % live.size();
            const Live l = live[i];
            for (size_t k = 0; k < l.n; k += 97) CHECK((unsigned char)l.p[k] == l.tag);
            CHECK((unsigned char)l.p[l.n - 1] == l.tag);
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
    bdda::CachingPool<MallocRaw> pool(env.raw(512 * MiB));
    std::atomic<int> bad{0};
    auto work = [&](unsigned seed) {
        std::mt19937_64 rng(seed);
        std::vector<std::pair<char*, size_t>> mine;
        for (int it = 0; it < 3000; ++it) {
            if (mine.empty() || rng() % 2) {
                const size_t n = (rng() % 3 == 0) ? 1 + rng() % 4096 : MiB + rng() % (6 * MiB);
                int err = 0;
                char* p = static_cast<char*>(pool.alloc(n, &err));
                if (!p) continue;
                std::memset(p, (int)(seed & 0x7f) + 1, n);
                mine.emplace_back(p, n);
            } else {
                const size_t i = rng() % mine.size();
                auto [p, n] = mine[i];
                if ((unsigned char)p[0] != (seed & 0x7f) + 1 || (unsigned char)p[n - 1] != (seed & 0x7f) + 1) ++bad;
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
    random_sequence(1, 96 * MiB, 6000);
    random_sequence(2, 40 * MiB, 6000);   // tight: out-of-memory and trimming on the way
    random_sequence(3, 1024 * MiB, 4000);
    threads();
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
