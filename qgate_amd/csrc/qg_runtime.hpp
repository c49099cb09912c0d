// Device context of the qgate HIP runtime: one device, one stream, a caching allocator and
// small scratch buffers. Every kernel of the library is ordered on that one stream.
#pragma once

#include <mutex>

#include "qg_common.hpp"

namespace qg {

struct Runtime {
    bool initialized = false;
    int device = 0;
    int n_cus = 0;
    size_t total_mem = 0;
    hipStream_t stream = nullptr;
    double *d_partials = nullptr;   // per-block partial sums of reductions
    double *h_pinned = nullptr;     // pinned host words for scalar readbacks
    std::recursive_mutex mtx;
};

Runtime &runtime();

// Initialise the runtime on `device` (no-op if already initialised on it).
void runtime_init(int device);
// Release cached memory, scratch and the stream.
void runtime_clear();
// Lazily initialise on the default device and return the stream.
hipStream_t stream();
void sync();

// Device allocations: power-of-two size classes up to 1 GiB are cached after free and
// reused without a device synchronisation (all users share the one stream, so reuse is
// stream-ordered). Larger blocks go straight back to the driver.
void *dev_alloc(size_t bytes);
void dev_free(void *ptr, size_t bytes);
size_t dev_cached_bytes();

// Grid size that fills the device for grid-stride reductions.
inline int reduce_grid() { return 4096; }

}  // namespace qg
