// Device context, stream and caching allocator.
// Counterpart of the reference device setup and memory store
// (qgate/simulator/src/CUDADevice.cpp:60-153, MultiDeviceMemoryStore.cpp:164-269), for one GPU:
// no chunking across devices, and no device-wide synchronisation on allocate/free.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <thread>
#include <vector>

#include "qg_runtime.hpp"

namespace qg {

void raise_error(const char *fmt, ...) {
    char buf[1024];
    va_list va;
    va_start(va, fmt);
    vsnprintf(buf, sizeof(buf), fmt, va);
    va_end(va);
    throw Error(buf);
}

Runtime &runtime() {
    static Runtime rt;
    return rt;
}

std::string &last_error() {
    thread_local std::string msg;
    return msg;
}

namespace {

constexpr size_t kCacheMaxBlock = size_t(1) << 30;   // cache blocks up to 1 GiB
constexpr size_t kCacheMaxTotal = size_t(8) << 30;   // keep at most 8 GiB cached

std::map<size_t, std::vector<void *>> &free_lists() {
    static std::map<size_t, std::vector<void *>> lists;
    return lists;
}
size_t g_cached = 0;

size_t size_class(size_t bytes) {
    size_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
}

void release_cache() {
    for (auto &kv : free_lists())
        for (void *p : kv.second) hipFree(p);
    free_lists().clear();
    g_cached = 0;
}

}  // namespace

namespace {
u64 g_piece_items = u64(1) << 30;
}

u64 launch_piece_items() { return g_piece_items; }

namespace {
bool g_register_tiles = true;
}

bool use_register_tiles() { return g_register_tiles; }
void set_register_tiles(bool on) { g_register_tiles = on; }

void set_launch_piece_items(u64 items) {
    QG_REQUIRE(items >= (u64(1) << 12) && items <= (u64(1) << 30) && (items & (items - 1)) == 0,
               "launch piece must be a power of two in [2^12, 2^30]");
    g_piece_items = items;
}

void runtime_init(int device) {
    Runtime &rt = runtime();
    std::lock_guard<std::recursive_mutex> lk(rt.mtx);
    if (rt.initialized) {
        QG_REQUIRE(rt.device == device, "runtime already initialised on device %d", rt.device);
        return;
    }
    int count = 0;
    QG_HIP(hipGetDeviceCount(&count));
    QG_REQUIRE(count > 0, "no HIP device is visible");
    QG_REQUIRE(0 <= device && device < count, "device %d out of range (%d devices)", device, count);
    QG_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    QG_HIP(hipGetDeviceProperties(&prop, device));
    rt.device = device;
    rt.n_cus = prop.multiProcessorCount;
    rt.total_mem = prop.totalGlobalMem;
    // default priority stream (non-blocking w.r.t. the legacy null stream)
    QG_HIP(hipStreamCreateWithFlags(&rt.stream, hipStreamNonBlocking));
    QG_HIP(hipMalloc(&rt.d_partials, sizeof(double) * (reduce_grid() + 64)));
    QG_HIP(hipHostMalloc(&rt.h_pinned, sizeof(double) * 64, hipHostMallocDefault));
    rt.initialized = true;
}

namespace {

struct StageBuffers {
    char *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
};

StageBuffers &stage_buffers() {
    static StageBuffers sb;
    return sb;
}

StageBuffers &staging() {
    stream();
    StageBuffers &sb = stage_buffers();
    if (sb.buf[0] == nullptr) {
        for (int b = 0; b < 2; ++b) {
            QG_HIP(hipHostMalloc(reinterpret_cast<void **>(&sb.buf[b]), kStageBytes, hipHostMallocDefault));
            QG_HIP(hipEventCreateWithFlags(&sb.ev[b], hipEventDisableTiming));
        }
    }
    return sb;
}

void release_staging() {
    StageBuffers &sb = stage_buffers();
    for (int b = 0; b < 2; ++b) {
        if (sb.buf[b] != nullptr) hipHostFree(sb.buf[b]);
        if (sb.ev[b] != nullptr) hipEventDestroy(sb.ev[b]);
        sb.buf[b] = nullptr;
        sb.ev[b] = nullptr;
    }
}

// memcpy split over up to 8 host threads (first touch of fresh pages included)
void parallel_memcpy(char *dst, const char *src, size_t bytes) {
    constexpr size_t kMinPerThread = size_t(4) << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t n = std::min<size_t>({8, hw, std::max<size_t>(1, bytes / kMinPerThread)});
    if (n <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = (bytes / n + 4095) & ~size_t(4095);
    std::vector<std::thread> th;
    for (size_t i = 1; i < n; ++i) {
        const size_t off = i * part;
        if (off >= bytes) break;
        const size_t len = std::min(part, bytes - off);
        th.emplace_back([=] { std::memcpy(dst + off, src + off, len); });
    }
    std::memcpy(dst, src, std::min(part, bytes));
    for (std::thread &t : th) t.join();
}

}  // namespace

HostStager::HostStager() { staging(); }

HostStager::~HostStager() {
    // a stager abandoned by an exception still waits for its copies into the staging buffers
    for (int b = 0; b < 2; ++b)
        if (slot_[b].active) hipEventSynchronize(stage_buffers().ev[b]);
}

void HostStager::drain(int b) {
    StageBuffers &sb = stage_buffers();
    QG_HIP(hipEventSynchronize(sb.ev[b]));
    slot_[b].active = false;
    parallel_memcpy(slot_[b].dst, sb.buf[b], slot_[b].bytes);
}

void HostStager::push(void *dst, const void *src, size_t bytes) {
    QG_REQUIRE(bytes <= kStageBytes, "staged chunk of %zu bytes exceeds the staging buffer", bytes);
    StageBuffers &sb = staging();
    const int b = next_;
    if (slot_[b].active) drain(b);
    QG_HIP(hipMemcpyAsync(sb.buf[b], src, bytes, hipMemcpyDeviceToHost, stream()));
    QG_HIP(hipEventRecord(sb.ev[b], stream()));
    slot_[b].dst = static_cast<char *>(dst);
    slot_[b].bytes = bytes;
    slot_[b].active = true;
    next_ ^= 1;
}

void HostStager::finish() {
    // the older pending chunk sits in buffer next_, the newer one in next_ ^ 1
    for (int k = 0; k < 2; ++k) {
        const int b = next_ ^ k;
        if (slot_[b].active) drain(b);
    }
}

void copy_to_host(void *dst, const void *src, size_t bytes) {
    HostStager st;
    const char *s = static_cast<const char *>(src);
    char *d = static_cast<char *>(dst);
    for (size_t done = 0; done < bytes; done += kStageBytes)
        st.push(d + done, s + done, std::min(kStageBytes, bytes - done));
    st.finish();
}

void runtime_clear() {
    Runtime &rt = runtime();
    std::lock_guard<std::recursive_mutex> lk(rt.mtx);
    if (!rt.initialized) return;
    hipStreamSynchronize(rt.stream);
    release_staging();
    release_cache();
    hipFree(rt.d_partials);
    hipHostFree(rt.h_pinned);
    hipStreamDestroy(rt.stream);
    rt.d_partials = nullptr;
    rt.h_pinned = nullptr;
    rt.stream = nullptr;
    rt.initialized = false;
}

hipStream_t stream() {
    Runtime &rt = runtime();
    if (!rt.initialized) runtime_init(0);
    return rt.stream;
}

void sync() { QG_HIP(hipStreamSynchronize(stream())); }

void *dev_alloc(size_t bytes) {
    stream();
    if (bytes == 0) bytes = 1;
    Runtime &rt = runtime();
    std::lock_guard<std::recursive_mutex> lk(rt.mtx);
    if (bytes <= kCacheMaxBlock) {
        size_t c = size_class(bytes);
        auto it = free_lists().find(c);
        if (it != free_lists().end() && !it->second.empty()) {
            void *p = it->second.back();
            it->second.pop_back();
            g_cached -= c;
            return p;
        }
        bytes = c;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        // give cached blocks back and retry once
        (void)hipGetLastError();
        sync();
        release_cache();
        e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        raise_error("Out of device memory (requested %zu bytes).", bytes);
    }
    return p;
}

void dev_free(void *ptr, size_t bytes) {
    if (ptr == nullptr) return;
    Runtime &rt = runtime();
    std::lock_guard<std::recursive_mutex> lk(rt.mtx);
    if (bytes == 0) bytes = 1;
    if (bytes <= kCacheMaxBlock) {
        size_t c = size_class(bytes);
        if (g_cached + c <= kCacheMaxTotal) {
            free_lists()[c].push_back(ptr);
            g_cached += c;
            return;
        }
    }
    // the block may still be read by queued kernels: drain the stream first
    QG_HIP(hipStreamSynchronize(rt.stream));
    QG_HIP(hipFree(ptr));
}

size_t dev_cached_bytes() { return g_cached; }

}  // namespace qg
