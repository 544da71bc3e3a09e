// devmem_pool.hpp -- caching sub-allocator over a raw device allocator.
//
// Blocks freed by the library stay with the pool and are handed out again:
// best fit among the idle blocks, the remainder split off as an idle block of
// its own, and idle neighbours within one raw allocation merged again on
// free.  The raw allocator is asked for memory only when no idle block fits;
// when it reports out of memory, every raw allocation that is wholly idle
// goes back to it (trim) and the request is tried once more.
//
// Requests up to 1 MiB come from 2 MiB raw allocations of their own, so that
// a small live block never pins a large raw allocation that could otherwise
// be trimmed.  Larger requests get raw allocations rounded up to 2 MiB.
//
// Header-only and free of HIP types: the library instantiates it over
// hipMalloc/hipFree (devmem.hip), and a host test over malloc/free.
//
// RAW provides  int alloc(size_t bytes, void** p)  (0 on success, RAW::kOom
// when out of memory, another nonzero code on any other failure) and
// void release(void* p).
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
    static constexpr size_t kAlign = 512;         // every block starts on a 512-byte boundary
    static constexpr size_t kSmallMax = 1 << 20;  // requests up to 1 MiB share 2 MiB raw allocations
    static constexpr size_t kSmallChunk = 2 << 20;
    static constexpr size_t kGranule = 2 << 20;   // larger raw allocations are multiples of 2 MiB
    static constexpr size_t kMinSplit = 1 << 20;  // a large block is split when at least this much is left

    struct Stats {
        size_t held = 0;       // bytes of all raw allocations
        size_t in_use = 0;     // bytes of live blocks (request sizes rounded to kAlign)
        int64_t raw_allocs = 0, raw_releases = 0, reuses = 0;
    };

    explicit CachingPool(RAW raw) : raw_(std::move(raw)) {}
    CachingPool(const CachingPool&) = delete;
    CachingPool& operator=(const CachingPool&) = delete;

    // nullptr for bytes == 0 (*err = 0) and on failure (*err = the raw code)
    void* alloc(size_t bytes, int* err)
    {
        *err = 0;
        if (bytes == 0) return nullptr;
        const size_t sz = round_up(bytes, kAlign);
        const bool small = sz <= kSmallMax;
        {
            std::lock_guard<std::mutex> g(m_);
            if (Block* b = take_idle(sz, small)) return b->ptr;
        }
        const size_t raw_bytes = small ? kSmallChunk : round_up(sz, kGranule);
        void* p = nullptr;
        int e = raw_.alloc(raw_bytes, &p);
        if (e == RAW::kOom) {
            trim();
            {
                // another thread may have freed a fitting block meanwhile
                std::lock_guard<std::mutex> g(m_);
                if (Block* b = take_idle(sz, small)) return b->ptr;
            }
            e = raw_.alloc(raw_bytes, &p);
        }
        if (e != 0) {
            *err = e;
            return nullptr;
        }
        std::lock_guard<std::mutex> g(m_);
        Block* b = new Block{static_cast<char*>(p), raw_bytes, false, small, nullptr, nullptr};
        st_.held += raw_bytes;
        ++st_.raw_allocs;
        use(b, sz);
        return b->ptr;
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
        idle_.insert(b);
        return true;
    }

    // every wholly idle raw allocation back to the raw allocator; returns the bytes released
    size_t trim()
    {
        std::vector<std::pair<void*, size_t>> out;
        {
            std::lock_guard<std::mutex> g(m_);
            for (auto it = idle_.begin(); it != idle_.end();) {
                Block* b = *it;
                if (!b->prev && !b->next) {
                    out.emplace_back(b->ptr, b->size);
                    st_.held -= b->size;
                    ++st_.raw_releases;
                    it = idle_.erase(it);
                    delete b;
                } else {
                    ++it;
                }
            }
        }
        size_t n = 0;
        for (auto& o : out) {
            raw_.release(o.first);
            n += o.second;
        }
        return n;
    }

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
        bool small;
        Block* prev;  // neighbours within one raw allocation, by address
        Block* next;
    };
    struct BySize {
        bool operator()(const Block* a, const Block* b) const
        {
            if (a->small != b->small) return a->small;
            if (a->size != b->size) return a->size < b->size;
            return std::less<const char*>()(a->ptr, b->ptr);
        }
    };

    static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

    // smallest idle block of the size class that holds sz (lock held)
    Block* take_idle(size_t sz, bool small)
    {
        Block key{nullptr, sz, false, small, nullptr, nullptr};
        auto it = idle_.lower_bound(&key);
        if (it == idle_.end() || (*it)->small != small) return nullptr;
        Block* b = *it;
        idle_.erase(it);
        ++st_.reuses;
        use(b, sz);
        return b;
    }

    // b (idle, not in the idle set) becomes live with sz bytes; the rest is split off (lock held)
    void use(Block* b, size_t sz)
    {
        const size_t rest = b->size - sz;
        if (rest >= (b->small ? kAlign : kMinSplit)) {
            Block* r = new Block{b->ptr + sz, rest, false, b->small, b, b->next};
            if (b->next) b->next->prev = r;
            b->next = r;
            b->size = sz;
            idle_.insert(r);
        }
        b->used = true;
        live_[b->ptr] = b;
        st_.in_use += b->size;
    }

    RAW raw_;
    std::mutex m_;
    std::set<Block*, BySize> idle_;
    std::unordered_map<void*, Block*> live_;
    Stats st_;
};

}  // namespace bdda
