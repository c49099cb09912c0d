// Device context of the qgate HIP runtime: one device, one stream, a caching allocator and
// small scratch buffers. Every kernel of the library is ordered on that one stream.
#pragma once

#include <mutex>
#include <string>

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

// Message of the last failed C-ABI call on this thread.
std::string &last_error();

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

// A launch covers at most 2^32 - 1 work items, so kernels over index ranges that can exceed
// that (one work item per amplitude or pair at 33+ qubits) are launched in pieces of
// launch_piece_items() items (default 2^30; a power of two >= 2^12), each piece given the
// index of its first item.
u64 launch_piece_items();
void set_launch_piece_items(u64 items);

template <class F>
void for_each_piece(u64 n_items, F &&f) {
    const u64 piece = launch_piece_items();
    for (u64 first = 0; first < n_items; first += piece)
        f(first, n_items - first < piece ? n_items - first : piece);
}

inline unsigned grid_for(u64 count, int block) { return (unsigned)((count + block - 1) / block); }

// Full-size tile passes keep the tile in registers (default) or work on it in LDS.
bool use_register_tiles();
void set_register_tiles(bool on);

// Device-to-host readout through pinned staging. Chunks of up to kStageBytes go by DMA into
// one of two pinned host buffers (hipHostMalloc, allocated on first use) while host threads
// copy the previous chunk out to the caller's (pageable) memory.
constexpr size_t kStageBytes = size_t(64) << 20;

class HostStager {
  public:
    HostStager();
    ~HostStager();
    // Queue a copy of `bytes` (<= kStageBytes) from device memory `src`, ordered on the
    // library stream after the work already queued, to host memory `dst`.
    void push(void *dst, const void *src, size_t bytes);
    // Complete every queued copy.
    void finish();

  private:
    void drain(int b);
    struct Slot {
        char *dst = nullptr;
        size_t bytes = 0;
        bool active = false;
    } slot_[2];
    int next_ = 0;
};

// Copy `bytes` of device memory to pageable host memory through the stager; synchronises.
void copy_to_host(void *dst, const void *src, size_t bytes);

}  // namespace qg
