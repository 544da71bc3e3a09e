// devmem_pool.hpp -- caching sub-allocator over a raw device allocator.
//
// Blocks freed by the library stay with the pool and are handed out again,
// so that the raw allocator is asked for memory only when no idle block
// fits.  Size classes and lifetimes keep a long-lived block from pinning a
// large raw allocation (which could then never go back to the raw allocator):
//
//   small   (up to 1 MiB)     2 MiB raw allocations, shared: best fit, split,
//   medium  (up to 64 MiB)    128 MiB raw allocations,   idle neighbours merged
//   large, scratch            raw allocations of the request (2 MiB multiple),
//     (freed within the call  best fit, split and merged like the above: only
//      that made it)          blocks that live for one call share them
//   large, kept               one raw allocation per block, never split; handed
//     (operator data)         out again for a request of at least 16/17 of its
//                             size (a live block is at most 1/16 larger than
//                             asked for)
//
// When the raw allocator reports out of memory, wholly idle raw allocations
// go back to it one at a time, least recently freed first, and the request is
// tried again after each.  After every raw allocation, idle ones also go back
// while the raw allocator reports memory low: the runtime allocates memory of
// its own too (scratch for kernels that spill registers, at their first
// dispatch on a queue), and a pool that held every free byte idle would make
// that fail.
//
// Header-only and free of HIP types: the library instantiates it over
// hipMalloc/hipFree (devmem.hip), and a host test over malloc/free.
//
// RAW provides  int alloc(size_t bytes, void** p)  (0 on success, RAW::kOom
// when out of memory, another nonzero code on any other failure),
// void release(void* p) and bool low() (free memory below the headroom the
// runtime needs).
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <mutex>
#include <set>
#include <unordered_map>
#include <utility>
#include <vector>

namespace bdda {

template <class RAW>
class CachingPool {
public:
    static constexpr size_t kAlign = 512;  // every block starts on a 512-byte boundary
    static constexpr size_t kSmallMax = size_t(1) << 20;
    static constexpr size_t kSmallChunk = size_t(2) << 20;
    static constexpr size_t kMediumMax = size_t(64) << 20;
    static constexpr size_t kMediumChunk = size_t(128) << 20;
    static constexpr size_t kGranule = size_t(2) << 20;  // large raw allocations are multiples of 2 MiB

    enum Kind : int { SMALL = 0, MEDIUM = 1, LARGE_SCRATCH = 2, LARGE_KEEP = 3 };

    struct Stats {
        size_t held = 0;    // bytes of all raw allocations
        size_t in_use = 0;  // bytes of live blocks (a block may exceed its request)
        int64_t raw_allocs = 0, raw_releases = 0, reuses = 0;
    };

    explicit CachingPool(RAW raw) : raw_(std::move(raw)) {}
    CachingPool(const CachingPool&) = delete;
    CachingPool& operator=(const CachingPool&) = delete;

    static Kind kind_of(size_t sz, bool keep)
    {
        return sz <= kSmallMax ? SMALL : sz <= kMediumMax ? MEDIUM : keep ? LARGE_KEEP : LARGE_SCRATCH;
    }

    // nullptr for bytes == 0 (*err = 0) and on failure (*err = the raw code);
    // keep: the block outlives the call that asks for it
    void* alloc(size_t bytes, bool keep, int* err)
    {
        *err = 0;
        if (bytes == 0) return nullptr;
        const size_t sz = round_up(bytes, kAlign);
        const Kind k = kind_of(sz, keep);
        {
            std::lock_guard<std::mutex> g(m_);
            if (Block* b = take_idle(sz, k)) return b->ptr;
        }
        const size_t raw_bytes = k == SMALL ? kSmallChunk : k == MEDIUM ? kMediumChunk : round_up(sz, kGranule);
        void* p = nullptr;
        int e = raw_.alloc(raw_bytes, &p);
        // out of memory: idle raw allocations back one at a time, least
        // recently freed first, until the request fits or none is left
        while (e == RAW::kOom && release_idle(1) > 0) {
            {
                // another thread may have freed a fitting block meanwhile
                std::lock_guard<std::mutex> g(m_);
                if (Block* b = take_idle(sz, k)) return b->ptr;
            }
            e = raw_.alloc(raw_bytes, &p);
        }
        if (e != 0) {
            *err = e;
            return nullptr;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            Block* b = new Block{static_cast<char*>(p), raw_bytes, false, k, nullptr, nullptr, 0};
            st_.held += raw_bytes;
            ++st_.raw_allocs;
            use(b, sz);
        }
        while (raw_.low() && release_idle(1) > 0) {
        }
        return p;
    }

    // false when p did not come from this pool (nothing is changed)
    bool free(void* p)
    {
        if (!p) return true;
        std::lock_guard<std::mutex> g(m_);
        auto it = live_.find(p);
        if (it == live_.end()) return false;
        Block* b = it->second;
        live_.erase(it);
        b->used = false;
        st_.in_use -= b->size;
        if (b->prev && !b->prev->used) {
            Block* a = b->prev;
            idle_.erase(a);
            a->size += b->size;
            a->next = b->next;
            if (b->next) b->next->prev = a;
            delete b;
            b = a;
        }
        if (b->next && !b->next->used) {
            Block* c = b->next;
            idle_.erase(c);
            b->size += c->size;
            b->next = c->next;
            if (c->next) c->next->prev = b;
            delete c;
        }
        b->stamp = ++tick_;
        idle_.insert(b);
        return true;
    }

    // every wholly idle raw allocation back to the raw allocator; returns the bytes released
    size_t trim() { return release_idle(SIZE_MAX); }

    bool owns(void* p)
    {
        std::lock_guard<std::mutex> g(m_);
        return live_.count(p) != 0;
    }

    Stats stats()
    {
        std::lock_guard<std::mutex> g(m_);
        return st_;
    }

    // idle blocks as (address, bytes), ascending address (for tests)
    std::vector<std::pair<const void*, size_t>> idle_blocks()
    {
        std::lock_guard<std::mutex> g(m_);
        std::vector<std::pair<const void*, size_t>> v;
        for (Block* b : idle_) v.emplace_back(b->ptr, b->size);
        std::sort(v.begin(), v.end());
        return v;
    }

private:
    struct Block {
        char* ptr;
        size_t size;
        bool used;
        Kind kind;
        Block* prev;  // neighbours within one raw allocation, by address
        Block* next;
        uint64_t stamp;  // when it last became idle
    };
    struct BySize {
        bool operator()(const Block* a, const Block* b) const
        {
            if (a->kind != b->kind) return a->kind < b->kind;
            if (a->size != b->size) return a->size < b->size;
            return std::less<const char*>()(a->ptr, b->ptr);
        }
    };

    static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

    // smallest idle block of class k that holds sz (lock held); a kept large
    // block only for a request of at least 16/17 of its size
    Block* take_idle(size_t sz, Kind k)
    {
        Block key{nullptr, sz, false, k, nullptr, nullptr, 0};
        auto it = idle_.lower_bound(&key);
        if (it == idle_.end() || (*it)->kind != k) return nullptr;
        Block* b = *it;
        if (k == LARGE_KEEP && b->size - sz > sz / 16) return nullptr;
        idle_.erase(it);
        ++st_.reuses;
        use(b, sz);
        return b;
    }

    // b (idle, not in the idle set) becomes live with sz bytes; the rest of
    // a block of a shared class is split off (lock held)
    void use(Block* b, size_t sz)
    {
        const size_t min_rest = b->kind == LARGE_KEEP ? SIZE_MAX : b->kind == LARGE_SCRATCH ? kGranule : kAlign;
        if (b->size - sz >= min_rest) {
            Block* r = new Block{b->ptr + sz, b->size - sz, false, b->kind, b, b->next, b->stamp};
            if (b->next) b->next->prev = r;
            b->next = r;
            b->size = sz;
            idle_.insert(r);
        }
        b->used = true;
        live_[b->ptr] = b;
        st_.in_use += b->size;
    }

    // wholly idle raw allocations back, least recently freed first, until at
    // least `want` bytes are released (SIZE_MAX: all); returns the bytes released
    size_t release_idle(size_t want)
    {
        std::vector<Block*> out;
        {
            std::lock_guard<std::mutex> g(m_);
            std::vector<Block*> whole;
            for (Block* b : idle_)
                if (!b->prev && !b->next) whole.push_back(b);
            std::sort(whole.begin(), whole.end(), [](const Block* a, const Block* b) { return a->stamp < b->stamp; });
            size_t n = 0;
            for (Block* b : whole) {
                if (n >= want) break;
                idle_.erase(b);
                st_.held -= b->size;
                ++st_.raw_releases;
                n += b->size;
                out.push_back(b);
            }
        }
        size_t n = 0;
        for (Block* b : out) {
            raw_.release(b->ptr);
            n += b->size;
            delete b;
        }
        return n;
    }

    RAW raw_;
    std::mutex m_;
    std::set<Block*, BySize> idle_;
    std::unordered_map<void*, Block*> live_;
    Stats st_;
    uint64_t tick_ = 0;
};

}  // namespace bdda
