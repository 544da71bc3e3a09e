// devmem.hip -- the caching device allocator over hipMalloc / hipFree.
#include "devmem.hpp"
#include "devmem_pool.hpp"
#include <cstdlib>
#include <cstring>

namespace bdda {
namespace {

struct HipRaw {
    static constexpr int kOom = (int)hipErrorOutOfMemory;
    int alloc(size_t bytes, void** p)
    {
        hipError_t e = hipMalloc(p, bytes);
        // a failed hipMalloc is also the thread's last error; clear it so that
        // the next kernel launch check does not report it
        if (e != hipSuccess) (void)hipGetLastError();
        return (int)e;
    }
    void release(void* p) { (void)hipFree(p); }
    // under 4 GiB free on the device (the runtime's own allocations need it)
    bool low()
    {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        return fr < (size_t(4) << 30);
    }
};

constexpr int kMaxDev = 64;

struct Pools {
    CachingPool<HipRaw>* dev[kMaxDev] = {};
    std::mutex m;
    bool on = true;
    Pools()
    {
        const char* v = std::getenv("BDDA_DEVICE_CACHE");
        on = !(v && std::strcmp(v, "0") == 0);
    }
};

// never destroyed: blocks may be freed from destructors that run at exit,
// and the runtime may already be gone by then
Pools& pools()
{
    static Pools* p = new Pools();
    return *p;
}

CachingPool<HipRaw>* pool_of(int dev)
{
    if (dev < 0 || dev >= kMaxDev) return nullptr;
    Pools& P = pools();
    std::lock_guard<std::mutex> g(P.m);
    if (!P.dev[dev]) P.dev[dev] = new CachingPool<HipRaw>(HipRaw{});
    return P.dev[dev];
}

CachingPool<HipRaw>* current_pool()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    return pool_of(dev);
}

}  // namespace

void* dev_malloc(size_t bytes, bool keep)
{
    if (bytes == 0) return nullptr;
    CachingPool<HipRaw>* pl = pools().on ? current_pool() : nullptr;
    if (!pl) {
        void* p = nullptr;
        BDDA_HIP(hipMalloc(&p, bytes));
        return p;
    }
    int err = 0;
    void* p = pl->alloc(bytes, keep, &err);
    if (!p) throw HipError{(hipError_t)err, "hipMalloc (device cache)"};
    return p;
}

void dev_free(void* p, hipStream_t st) noexcept
{
    if (!p) return;
    if (!pools().on) {
        (void)hipFree(p);
        return;
    }
    (void)(st ? hipStreamSynchronize(st) : hipDeviceSynchronize());
    if (CachingPool<HipRaw>* pl = current_pool())
        if (pl->free(p)) return;
    for (int d = 0; d < kMaxDev; ++d) {
        CachingPool<HipRaw>* pl;
        {
            std::lock_guard<std::mutex> g(pools().m);
            pl = pools().dev[d];
        }
        if (pl && pl->free(p)) return;
    }
    (void)hipFree(p);
}

size_t dev_trim() noexcept
{
    CachingPool<HipRaw>* pl = pools().on ? current_pool() : nullptr;
    return pl ? pl->trim() : 0;
}

DevStats dev_stats() noexcept
{
    DevStats s;
    CachingPool<HipRaw>* pl = pools().on ? current_pool() : nullptr;
    if (!pl) return s;
    const auto t = pl->stats();
    s.held = t.held;
    s.in_use = t.in_use;
    s.runtime_allocs = t.raw_allocs;
    s.runtime_frees = t.raw_releases;
    s.reuses = t.reuses;
    return s;
}

}  // namespace bdda
