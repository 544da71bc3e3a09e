// devmem.hpp -- device memory of the library: a caching allocator per GPU.
//
// Memory new to the process is slow to get from the runtime: in the serving
// stream (scenes of up to 34 GB of operator build transient, workspaces of up
// to 32 GiB) single hipMalloc calls took 0.3-3 s, 18 of them 19 s of a 74 s
// run, and every hipFree waits for the whole device, the other workers'
// streams included.  Every device buffer of the library therefore comes from
// a CachingPool (devmem_pool.hpp) per device, which hands freed blocks out
// again and asks the runtime only when no idle block fits.
//
// dev_free(p, st) first waits for stream st (the stream the last work on p
// was queued on; nullptr waits for the whole device, as hipFree does), so the
// block is idle for any stream or thread once it is back in the pool.  The
// environment variable BDDA_DEVICE_CACHE=0 turns the cache off (plain
// hipMalloc / hipFree).
#pragma once
#include "bdda_common.hpp"

namespace bdda {

// throws HipError (hipErrorOutOfMemory when nothing fits, after the pool's
// idle runtime allocations went back to the runtime); nullptr for 0 bytes.
// keep: the block outlives the call that asks for it (operator data, the
// caller's buffers); otherwise it is freed before that call returns, and
// large ones may share runtime allocations with other such blocks.
void* dev_malloc(size_t bytes, bool keep = false);
// p: nullptr, or a pointer from dev_malloc (others go to hipFree)
void dev_free(void* p, hipStream_t st) noexcept;
// the current device's wholly idle runtime allocations back to the runtime;
// returns the bytes released
size_t dev_trim() noexcept;

struct DevStats {
    size_t held = 0;    // bytes the pool holds from the runtime (current device)
    size_t in_use = 0;  // bytes of live blocks
    int64_t runtime_allocs = 0, runtime_frees = 0, reuses = 0;
};
DevStats dev_stats() noexcept;

template <typename T>
T* dev_alloc_n(size_t n, bool keep = false)
{
    return static_cast<T*>(dev_malloc(sizeof(T) * n, keep));
}

// an owned device array of n T for the duration of one call, freed after its
// stream's work
template <typename T>
struct DevPtr {
    T* p = nullptr;
    hipStream_t st = nullptr;
    DevPtr(size_t n, hipStream_t stream) : st(stream)
    {
        if (n) p = dev_alloc_n<T>(n);
    }
    ~DevPtr() { dev_free(p, st); }
    DevPtr(const DevPtr&) = delete;
    DevPtr& operator=(const DevPtr&) = delete;
};

}  // namespace bdda
