// runtime.hip — context, caching allocator, pinned staging, columns and
// Arrow C Data Interface import/export.  Replaces the UPMEM host movement
// layer (pimdal/support/transfer_helper.h:87-283): columns are copied once
// into device-resident buffers instead of being scattered across DPUs.
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <new>
#include <vector>

#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

static thread_local std::string g_last_error;

static thread_local uint64_t g_error_count = 0;

void set_last_error(const std::string& m) {
    g_last_error = m;
    g_error_count++;
}

uint64_t error_count() { return g_error_count; }

[[noreturn]] void fail(int code, const std::string& msg) { throw Error(code, msg); }

void hip_check(hipError_t e, const char* what, const char* file, int line) {
    if (e != hipSuccess) {
        (void)hipGetLastError();
        int code = (e == hipErrorOutOfMemory) ? PIMDAL_E_NOMEM : PIMDAL_E_HIP;
        fail(code, std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) + ")");
    }
}

uint32_t type_width(pimdal_type t) {
    switch (t) {
        case PIMDAL_U8: return 1;
        case PIMDAL_I32: case PIMDAL_U32: return 4;
        case PIMDAL_I64: case PIMDAL_U64: return 8;
        case PIMDAL_I128: return 16;
    }
    fail(PIMDAL_E_INVALID, "unknown column type");
}

// ---------------- caching device allocator ----------------

uint64_t DevicePool::round_size(uint64_t b) {
    if (b == 0) b = 1;
    if (b <= (1ull << 20)) return (b + 255) & ~255ull;
    const uint64_t g = 2ull << 20;
    return (b + g - 1) / g * g;
}

void DevicePool::init(int device, uint64_t budget) {
    device_ = device;
    budget_ = budget;
}

// Events that mark when a freed block's last use completes are recycled:
// creating and destroying one per free costs more host time than the
// small operators' kernels.
hipEvent_t DevicePool::take_event() {
    if (!idle_events_.empty()) {
        hipEvent_t e = idle_events_.back();
        idle_events_.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    PD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
}

void DevicePool::give_event(hipEvent_t e) {
    if (e) idle_events_.push_back(e);
}

// Frees cached blocks, largest first, until reserved + need fits the budget
// (or nothing is cached).  Caller holds mu_.
void DevicePool::trim_locked(uint64_t need) {
    if (reserved_ + need <= budget_ || cached_.empty()) return;
    PD_HIP(hipDeviceSynchronize());
    while (!cached_.empty() && reserved_ + need > budget_) {
        auto it = std::prev(cached_.end());
        (void)hipFree(it->second.p);
        give_event(it->second.ev);
        reserved_ -= it->second.size;
        cached_.erase(it);
    }
}

void DevicePool::set_budget(uint64_t b) {
    std::lock_guard<std::mutex> lk(mu_);
    budget_ = b;
    trim_locked(0);   // a lower budget gives cached memory back at once
}

void* DevicePool::alloc(uint64_t bytes, hipStream_t stream) {
    std::lock_guard<std::mutex> lk(mu_);
    const uint64_t sz = round_size(bytes);
    // reuse the smallest cached block of sz .. sz + sz/8 bytes whose free is
    // on the same stream or has completed (requests of many sizes, as a
    // serving stream makes, would otherwise each leave a block of their own)
    for (auto it = cached_.lower_bound(sz); it != cached_.end() && it->first <= sz + sz / 8; ++it) {
        Block& b = it->second;
        bool ok = (b.stream == stream) || (hipEventQuery(b.ev) == hipSuccess);
        if (ok) {
            void* p = b.p;
            const uint64_t bsz = b.size;
            give_event(b.ev);
            cached_.erase(it);
            live_ += bsz;
            live_blocks_[p] = bsz;
            peak_ = std::max(peak_, live_);
            return p;
        }
    }
    if (reserved_ + sz > budget_) {
        // give cached memory back, largest blocks first, until the request fits
        trim_locked(sz);
        if (reserved_ + sz > budget_)
            fail(PIMDAL_E_NOMEM, "HBM budget exceeded: need " + std::to_string(sz) + " B, reserved " +
                                     std::to_string(reserved_) + " B of " + std::to_string(budget_) + " B");
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(PIMDAL_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    reserved_ += sz;
    live_ += sz;
    live_blocks_[p] = sz;
    peak_ = std::max(peak_, live_);
    return p;
}

void DevicePool::free(void* p, hipStream_t stream) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = live_blocks_.find(p);
    if (it == live_blocks_.end()) fail(PIMDAL_E_INVALID, "free of a pointer the pool does not own");
    const uint64_t sz = it->second;
    live_blocks_.erase(it);
    live_ -= sz;
    hipEvent_t ev = take_event();
    PD_HIP(hipEventRecord(ev, stream));
    cached_.emplace(sz, Block{p, sz, stream, ev});
}

void DevicePool::release_cached() {
    std::lock_guard<std::mutex> lk(mu_);
    (void)hipDeviceSynchronize();
    for (auto& kv : cached_) {
        (void)hipFree(kv.second.p);
        give_event(kv.second.ev);
        reserved_ -= kv.second.size;
    }
    cached_.clear();
}

void DevicePool::destroy() {
    release_cached();
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& kv : live_blocks_) {
        (void)hipFree(kv.first);
        reserved_ -= kv.second;
    }
    live_blocks_.clear();
    live_ = 0;
    for (hipEvent_t e : idle_events_) (void)hipEventDestroy(e);
    idle_events_.clear();
}

// ---------------- profiling ----------------

hipEvent_t take_event(pimdal_ctx* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    PD_HIP(hipEventCreate(&e));
    return e;
}

void record_launch(pimdal_ctx* c, const char* name, hipEvent_t e0, hipStream_t s) {
    hipEvent_t e1 = take_event(c);
    PD_HIP(hipEventRecord(e1, s));
    c->pending.push_back({name, e0, e1});
}

hipStream_t side_fork(pimdal_ctx* c, hipStream_t s) {
    if (!c->side) {
        PD_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        for (hipEvent_t& e : c->side_ev) PD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    PD_HIP(hipEventRecord(c->side_ev[0], s));
    PD_HIP(hipStreamWaitEvent(c->side, c->side_ev[0], 0));
    return c->side;
}

void side_join(pimdal_ctx* c, hipStream_t s) {
    PD_HIP(hipEventRecord(c->side_ev[1], c->side));
    PD_HIP(hipStreamWaitEvent(s, c->side_ev[1], 0));
}

static void resolve_pending(pimdal_ctx* c) {
    for (auto& p : c->pending) {
        PD_HIP(hipEventSynchronize(p.e1));
        float ms = 0.f;
        PD_HIP(hipEventElapsedTime(&ms, p.e0, p.e1));
        KernelStat& st = c->stats[p.name];
        st.launches++;
        st.total_ms += ms;
        c->event_pool.push_back(p.e0);
        c->event_pool.push_back(p.e1);
    }
    c->pending.clear();
}

// ---------------- small utility kernels ----------------

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void fill_u32(pimdal_ctx* c, void* p, uint32_t value, uint64_t count, hipStream_t s) {
    if (!count) return;
    uint32_t g = (uint32_t)std::min<uint64_t>(ceil_div(count, 256), 4096);
    launch(c, "fill_u32", k_fill_u32, dim3(g), dim3(256), 0, s, (uint32_t*)p, value, count);
}

void fill_u64(pimdal_ctx* c, void* p, uint64_t value, uint64_t count, hipStream_t s) {
    if (!count) return;
    uint32_t g = (uint32_t)std::min<uint64_t>(ceil_div(count, 256), 4096);
    launch(c, "fill_u64", k_fill_u64, dim3(g), dim3(256), 0, s, (uint64_t*)p, value, count);
}

void copy_d2d(pimdal_ctx* c, void* dst, const void* src, uint64_t bytes, hipStream_t s) {
    (void)c;
    if (bytes) PD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
}

uint64_t read_u64(pimdal_ctx* c, const void* dev, hipStream_t s) {
    uint64_t v = 0;
    read_small(c, &v, dev, 8, s);
    return v;
}

// Small device-to-host reads (counts, flags) through a pinned word buffer
// of the context: a copy into pageable memory goes through the runtime's
// staging path, several times slower for a few bytes.
void read_small(pimdal_ctx* c, void* dst, const void* dev, uint64_t bytes, hipStream_t s) {
    PD_REQUIRE(!c->capture, PIMDAL_E_UNSUPPORTED,
               "a result read back to the host cannot be captured into a plan; use an asynchronous device form");
    if (bytes > pimdal_ctx::READBACK_BYTES) {
        PD_HIP(hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToHost, s));
        PD_HIP(hipStreamSynchronize(s));
        return;
    }
    if (!c->readback) PD_HIP(hipHostMalloc(&c->readback, pimdal_ctx::READBACK_BYTES, hipHostMallocDefault));
    PD_HIP(hipMemcpyAsync(c->readback, dev, bytes, hipMemcpyDeviceToHost, s));
    PD_HIP(hipStreamSynchronize(s));
    memcpy(dst, c->readback, bytes);
}

// ---------------- pinned staging ----------------

static constexpr uint64_t kStageBytes = 32ull << 20;

static void ensure_pinned(pimdal_ctx* c) {
    if (c->pinned[0]) return;
    for (int i = 0; i < 2; i++) {
        PD_HIP(hipHostMalloc(&c->pinned[i], kStageBytes, hipHostMallocDefault));
        PD_HIP(hipEventCreateWithFlags(&c->pinned_ev[i], hipEventDisableTiming));
        PD_HIP(hipEventRecord(c->pinned_ev[i], c->stream));
    }
    c->pinned_bytes = kStageBytes;
}

// host memcpy into one pinned buffer while the other drains to the device
static void stage_h2d(pimdal_ctx* c, void* dst, const void* src, uint64_t bytes) {
    ensure_pinned(c);
    uint64_t done = 0;
    int k = 0;
    while (done < bytes) {
        const uint64_t n = std::min(kStageBytes, bytes - done);
        PD_HIP(hipEventSynchronize(c->pinned_ev[k]));
        memcpy(c->pinned[k], (const char*)src + done, n);
        PD_HIP(hipMemcpyAsync((char*)dst + done, c->pinned[k], n, hipMemcpyHostToDevice, c->stream));
        PD_HIP(hipEventRecord(c->pinned_ev[k], c->stream));
        done += n;
        k ^= 1;
    }
    PD_HIP(hipStreamSynchronize(c->stream));
}

static void stage_d2h(pimdal_ctx* c, void* dst, const void* src, uint64_t bytes) {
    ensure_pinned(c);
    PD_HIP(hipStreamSynchronize(c->stream));
    uint64_t done = 0;
    uint64_t pend_off[2] = {0, 0}, pend_n[2] = {0, 0};
    int k = 0;
    while (done < bytes || pend_n[0] || pend_n[1]) {
        if (pend_n[k]) {
            PD_HIP(hipEventSynchronize(c->pinned_ev[k]));
            memcpy((char*)dst + pend_off[k], c->pinned[k], pend_n[k]);
            pend_n[k] = 0;
        }
        if (done < bytes) {
            const uint64_t n = std::min(kStageBytes, bytes - done);
            PD_HIP(hipMemcpyAsync(c->pinned[k], (const char*)src + done, n, hipMemcpyDeviceToHost, c->stream));
            PD_HIP(hipEventRecord(c->pinned_ev[k], c->stream));
            pend_off[k] = done;
            pend_n[k] = n;
            done += n;
        }
        k ^= 1;
    }
}

}  // namespace pd

using namespace pd;

// ---------------- C ABI: context ----------------

extern "C" const char* pimdal_last_error(void) { return pd::g_last_error.c_str(); }

extern "C" int pimdal_version(void) { return 10000; }

extern "C" int pimdal_ctx_create(int device, uint64_t hbm_budget_bytes, pimdal_ctx** out) {
    PD_API_BEGIN
    PD_REQUIRE(out, PIMDAL_E_INVALID, "out is NULL");
    int n = 0;
    PD_HIP(hipGetDeviceCount(&n));
    PD_REQUIRE(device >= 0 && device < n, PIMDAL_E_INVALID, "no such device");
    PD_HIP(hipSetDevice(device));
    // device address of the no-allocation select's scratch array (once per device)
    pd::select_noalloc_init(device);
    auto* c = new pimdal_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    PD_HIP(hipGetDeviceProperties(&prop, device));
    c->cu_count = prop.multiProcessorCount;
    size_t free_b = 0, total_b = 0;
    PD_HIP(hipMemGetInfo(&free_b, &total_b));
    if (hbm_budget_bytes == 0) hbm_budget_bytes = (uint64_t)((double)free_b * 0.9);
    c->pool.init(device, hbm_budget_bytes);
    PD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    *out = c;
    PD_API_END
}

extern "C" int pimdal_ctx_destroy(pimdal_ctx* c) {
    PD_API_BEGIN
    if (!c) return PIMDAL_OK;
    (void)hipSetDevice(c->device);
    if (c->capture) {   // a capture left open: end it (the plan itself stays with the caller, empty)
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        c->capture->capturing = false;
        c->capture = nullptr;
    }
    (void)hipStreamSynchronize(c->stream);
    c->pool.destroy();
    for (int i = 0; i < 2; i++) {
        if (c->pinned[i]) (void)hipHostFree(c->pinned[i]);
        if (c->pinned_ev[i]) (void)hipEventDestroy(c->pinned_ev[i]);
    }
    if (c->readback) (void)hipHostFree(c->readback);
    for (auto& p : c->pending) { (void)hipEventDestroy(p.e0); (void)hipEventDestroy(p.e1); }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    if (c->side) {
        (void)hipStreamSynchronize(c->side);
        (void)hipStreamDestroy(c->side);
        for (hipEvent_t e : c->side_ev) (void)hipEventDestroy(e);
    }
    (void)hipStreamDestroy(c->stream);
    delete c;
    PD_API_END
}

extern "C" int pimdal_ctx_sync(pimdal_ctx* c) {
    PD_API_BEGIN
    PD_REQUIRE(c, PIMDAL_E_INVALID, "ctx is NULL");
    PD_NO_CAPTURE(c);
    PD_HIP(hipStreamSynchronize(c->stream));
    PD_API_END
}

extern "C" int pimdal_ctx_stream(pimdal_ctx* c, void** stream) {
    PD_API_BEGIN
    PD_REQUIRE(c && stream, PIMDAL_E_INVALID, "NULL argument");
    *stream = (void*)c->stream;
    PD_API_END
}

extern "C" int pimdal_ctx_mem_stats(pimdal_ctx* c, uint64_t* live, uint64_t* reserved, uint64_t* peak, uint64_t* budget) {
    PD_API_BEGIN
    PD_REQUIRE(c, PIMDAL_E_INVALID, "ctx is NULL");
    if (live) *live = c->pool.live();
    if (reserved) *reserved = c->pool.reserved();
    if (peak) *peak = c->pool.peak();
    if (budget) *budget = c->pool.budget();
    PD_API_END
}

extern "C" int pimdal_ctx_set_budget(pimdal_ctx* c, uint64_t b) {
    PD_API_BEGIN
    PD_REQUIRE(c && b, PIMDAL_E_INVALID, "bad argument");
    c->pool.set_budget(b);
    PD_API_END
}

extern "C" int pimdal_ctx_set_profiling(pimdal_ctx* c, int on) {
    PD_API_BEGIN
    PD_REQUIRE(c, PIMDAL_E_INVALID, "ctx is NULL");
    c->profiling = on != 0;
    PD_API_END
}

extern "C" int pimdal_ctx_kernel_stat(pimdal_ctx* c, const char* kernel, uint64_t* launches, double* total_ms) {
    PD_API_BEGIN
    PD_REQUIRE(c && kernel, PIMDAL_E_INVALID, "NULL argument");
    resolve_pending(c);
    auto it = c->stats.find(kernel);
    if (launches) *launches = it == c->stats.end() ? 0 : it->second.launches;
    if (total_ms) *total_ms = it == c->stats.end() ? 0.0 : it->second.total_ms;
    PD_API_END
}

extern "C" int pimdal_ctx_reset_stats(pimdal_ctx* c) {
    PD_API_BEGIN
    PD_REQUIRE(c, PIMDAL_E_INVALID, "ctx is NULL");
    resolve_pending(c);
    c->stats.clear();
    PD_API_END
}

extern "C" int pimdal_ctx_set_tune(pimdal_ctx* c, const char* key, int64_t value) {
    PD_API_BEGIN
    PD_REQUIRE(c && key, PIMDAL_E_INVALID, "NULL argument");
    c->tune[key] = value;
    PD_API_END
}

extern "C" int pimdal_ctx_get_tune(pimdal_ctx* c, const char* key, int64_t* value) {
    PD_API_BEGIN
    PD_REQUIRE(c && key && value, PIMDAL_E_INVALID, "NULL argument");
    auto it = c->tune.find(key);
    PD_REQUIRE(it != c->tune.end(), PIMDAL_E_INVALID, std::string("tuning key not set: ") + key);
    *value = it->second;
    PD_API_END
}

extern "C" int pimdal_ctx_clear_tune(pimdal_ctx* c, const char* key) {
    PD_API_BEGIN
    PD_REQUIRE(c && key, PIMDAL_E_INVALID, "NULL argument");
    c->tune.erase(key);
    PD_API_END
}

// ---------------- C ABI: memory ----------------

extern "C" int pimdal_dev_alloc(pimdal_ctx* c, uint64_t bytes, void** dptr) {
    PD_API_BEGIN
    PD_REQUIRE(c && dptr, PIMDAL_E_INVALID, "NULL argument");
    *dptr = c->pool.alloc(bytes, c->stream);
    PD_API_END
}

extern "C" int pimdal_dev_free(pimdal_ctx* c, void* dptr) {
    PD_API_BEGIN
    PD_REQUIRE(c, PIMDAL_E_INVALID, "ctx is NULL");
    c->pool.free(dptr, c->stream);
    PD_API_END
}

extern "C" int pimdal_dev_copy(pimdal_ctx* c, void* dst, const void* src, uint64_t bytes) {
    PD_API_BEGIN
    PD_REQUIRE(c && (bytes == 0 || (dst && src)), PIMDAL_E_INVALID, "NULL argument");
    copy_d2d(c, dst, src, bytes, c->stream);
    PD_API_END
}

extern "C" int pimdal_memcpy_h2d(pimdal_ctx* c, void* dst, const void* src, uint64_t bytes) {
    PD_API_BEGIN
    PD_REQUIRE(c && (bytes == 0 || (dst && src)), PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    stage_h2d(c, dst, src, bytes);
    PD_API_END
}

extern "C" int pimdal_memcpy_d2h(pimdal_ctx* c, void* dst, const void* src, uint64_t bytes) {
    PD_API_BEGIN
    PD_REQUIRE(c && (bytes == 0 || (dst && src)), PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    stage_d2h(c, dst, src, bytes);
    PD_API_END
}

// ---------------- C ABI: columns ----------------

extern "C" int pimdal_col_alloc(pimdal_ctx* c, pimdal_type type, uint64_t length, pimdal_col** out) {
    PD_API_BEGIN
    PD_REQUIRE(c && out, PIMDAL_E_INVALID, "NULL argument");
    const uint32_t w = type_width(type);
    auto* col = new pimdal_col{type, length, length, nullptr, true};
    try {
        col->data = c->pool.alloc(std::max<uint64_t>(length, 1) * w, c->stream);
    } catch (...) {
        delete col;
        throw;
    }
    *out = col;
    PD_API_END
}

extern "C" int pimdal_col_from_host(pimdal_ctx* c, pimdal_type type, const void* data, uint64_t length, pimdal_col** out) {
    int rc = pimdal_col_alloc(c, type, length, out);
    if (rc) return rc;
    PD_API_BEGIN
    PD_REQUIRE(data || length == 0, PIMDAL_E_INVALID, "data is NULL");
    stage_h2d(c, (*out)->data, data, length * type_width(type));
    PD_API_END
}

extern "C" int pimdal_col_wrap(pimdal_ctx* c, pimdal_type type, void* dptr, uint64_t length, pimdal_col** out) {
    PD_API_BEGIN
    PD_REQUIRE(c && out && (dptr || length == 0), PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    (void)type_width(type);
    *out = new pimdal_col{type, length, length, dptr, false};
    PD_API_END
}

extern "C" int pimdal_col_to_host(pimdal_ctx* c, const pimdal_col* col, void* dst, uint64_t dst_bytes) {
    PD_API_BEGIN
    PD_REQUIRE(c && col && (dst || col->length == 0), PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    const uint64_t b = col->length * type_width(col->type);
    PD_REQUIRE(dst_bytes >= b, PIMDAL_E_CAPACITY, "destination too small");
    stage_d2h(c, dst, col->data, b);
    PD_API_END
}

extern "C" int pimdal_col_info(const pimdal_col* col, pimdal_type* type, uint64_t* length, void** dptr) {
    PD_API_BEGIN
    PD_REQUIRE(col, PIMDAL_E_INVALID, "col is NULL");
    if (type) *type = col->type;
    if (length) *length = col->length;
    if (dptr) *dptr = col->data;
    PD_API_END
}

extern "C" int pimdal_col_free(pimdal_ctx* c, pimdal_col* col) {
    PD_API_BEGIN
    PD_REQUIRE(c, PIMDAL_E_INVALID, "ctx is NULL");
    if (!col) return PIMDAL_OK;
    if (col->owned && col->data) c->pool.free(col->data, c->stream);
    if (col->validity) c->pool.free(col->validity, c->stream);
    delete col;
    PD_API_END
}

// ---------------- Arrow C Data Interface ----------------

static bool arrow_type(const char* fmt, pimdal_type* t) {
    if (!fmt) return false;
    if (!strcmp(fmt, "C")) { *t = PIMDAL_U8; return true; }
    if (!strcmp(fmt, "i")) { *t = PIMDAL_I32; return true; }
    if (!strcmp(fmt, "tdD")) { *t = PIMDAL_I32; return true; }   // date32: int32 days since 1970-01-01
    if (!strcmp(fmt, "I")) { *t = PIMDAL_U32; return true; }
    if (!strcmp(fmt, "l")) { *t = PIMDAL_I64; return true; }
    if (!strcmp(fmt, "L")) { *t = PIMDAL_U64; return true; }
    if (!strcmp(fmt, "d:38,0") || !strcmp(fmt, "d:38,0,128")) { *t = PIMDAL_I128; return true; }
    return false;
}

static const char* arrow_format(pimdal_type t) {
    switch (t) {
        case PIMDAL_U8: return "C";
        case PIMDAL_I32: return "i";
        case PIMDAL_U32: return "I";
        case PIMDAL_I64: return "l";
        case PIMDAL_U64: return "L";
        case PIMDAL_I128: return "d:38,0";
    }
    return "";
}

// Type of an importable Arrow schema (primitive, fixed width, no children).
static pimdal_type import_type(const struct ArrowSchema* s) {
    PD_REQUIRE(s->release, PIMDAL_E_INVALID, "released Arrow schema");
    pimdal_type t;
    PD_REQUIRE(arrow_type(s->format, &t), PIMDAL_E_UNSUPPORTED,
               std::string("unsupported Arrow format '") + (s->format ? s->format : "") + "'");
    PD_REQUIRE(s->n_children == 0 && !s->dictionary, PIMDAL_E_UNSUPPORTED,
               "nested or dictionary arrays are not supported");
    return t;
}

// Host address of an importable array's first value (its offset applied).
static const char* import_values(const struct ArrowArray* a, uint32_t w) {
    PD_REQUIRE(a, PIMDAL_E_INVALID, "NULL Arrow array");
    PD_REQUIRE(a->release, PIMDAL_E_INVALID, "released Arrow array");
    PD_REQUIRE(a->n_children == 0 && !a->dictionary, PIMDAL_E_UNSUPPORTED,
               "nested or dictionary arrays are not supported");
    PD_REQUIRE(a->n_buffers == 2, PIMDAL_E_UNSUPPORTED, "expected a primitive array with 2 buffers");
    PD_REQUIRE(a->null_count == 0 || a->buffers[0] != nullptr || a->null_count == -1, PIMDAL_E_INVALID,
               "null_count > 0 without a validity buffer");
    PD_REQUIRE(a->length >= 0 && a->offset >= 0, PIMDAL_E_INVALID, "negative length or offset");
    const char* base = (const char*)a->buffers[1];
    PD_REQUIRE(base || a->length == 0, PIMDAL_E_INVALID, "missing data buffer");
    return base ? base + (uint64_t)a->offset * w : nullptr;
}

extern "C" int pimdal_col_import(pimdal_ctx* c, const struct ArrowArray* a, const struct ArrowSchema* s, pimdal_col** out) {
    PD_API_BEGIN
    PD_REQUIRE(c && a && s && out, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    return pimdal_col_import_chunks(c, &a, 1, s, out);
    PD_API_END
}

extern "C" int pimdal_col_import_chunks(pimdal_ctx* c, const struct ArrowArray* const* chunks, uint32_t nchunks,
                                        const struct ArrowSchema* s, pimdal_col** out) {
    PD_API_BEGIN
    PD_REQUIRE(c && s && out && (chunks || nchunks == 0), PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    const pimdal_type t = import_type(s);
    const uint32_t w = type_width(t);
    // every chunk is checked before anything is allocated or copied
    uint64_t total = 0;
    for (uint32_t i = 0; i < nchunks; i++) {
        import_values(chunks[i], w);
        total += (uint64_t)chunks[i]->length;
    }
    pimdal_col* col = nullptr;
    int rc = pimdal_col_alloc(c, t, total, &col);
    if (rc) fail(rc, pimdal_last_error());
    try {
        uint64_t at = 0;
        for (uint32_t i = 0; i < nchunks; i++) {
            const uint64_t len = (uint64_t)chunks[i]->length;
            if (len) stage_h2d(c, (char*)col->data + at * w, import_values(chunks[i], w), len * w);
            at += len;
        }
        // validity: the chunks' bitmaps (their offsets applied) concatenated
        // at bit level into one bitmap from bit 0; chunks without one are all
        // valid; no bitmap at all when nothing is null
        bool any = false;
        for (uint32_t i = 0; i < nchunks; i++) any |= chunks[i]->buffers[0] != nullptr && chunks[i]->null_count != 0;
        if (any) {
            std::vector<uint32_t> bits(valid_words(total), 0u);
            uint64_t nulls = 0;
            at = 0;
            for (uint32_t i = 0; i < nchunks; i++) {
                const struct ArrowArray* a = chunks[i];
                const uint8_t* vb = (const uint8_t*)a->buffers[0];
                const uint64_t len = (uint64_t)a->length, off = (uint64_t)a->offset;
                for (uint64_t r = 0; r < len; r++) {
                    const bool ok = !vb || a->null_count == 0 || ((vb[(off + r) >> 3] >> ((off + r) & 7)) & 1);
                    if (ok) bits[(at + r) >> 5] |= 1u << ((at + r) & 31);
                    else nulls++;
                }
                at += len;
            }
            if (nulls) set_validity_words(c, col, bits.data(), nulls);
        }
    } catch (...) {
        pimdal_col_free(c, col);
        throw;
    }
    *out = col;
    PD_API_END
}

// 32-bit words of a validity bitmap of n rows.
uint64_t pd::valid_words(uint64_t n) { return (n + 31) / 32; }

// Gives `col` the validity of host bitmap words `bits` (valid_words(length)
// words, bits past the length ignored) and its null count.
void pd::set_validity_words(pimdal_ctx* c, pimdal_col* col, const uint32_t* bits, uint64_t nulls) {
    const uint64_t nw = valid_words(col->length);
    if (!col->validity) col->validity = c->pool.alloc(std::max<uint64_t>(nw, 1) * 4, c->stream);
    stage_h2d(c, col->validity, bits, nw * 4);
    col->null_count = nulls;
}

extern "C" int pimdal_col_validity(const pimdal_col* col, void** dptr, uint64_t* null_count) {
    PD_API_BEGIN
    PD_REQUIRE(col, PIMDAL_E_INVALID, "col is NULL");
    if (dptr) *dptr = col->validity;
    if (null_count) *null_count = col->null_count;
    PD_API_END
}

extern "C" int pimdal_col_set_validity(pimdal_ctx* c, pimdal_col* col, const uint8_t* bits) {
    PD_API_BEGIN
    PD_REQUIRE(c && col, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    if (!bits) {
        if (col->validity) c->pool.free(col->validity, c->stream);
        col->validity = nullptr;
        col->null_count = 0;
        return PIMDAL_OK;
    }
    std::vector<uint32_t> w(std::max<uint64_t>(valid_words(col->length), 1), 0u);
    memcpy(w.data(), bits, (col->length + 7) / 8);
    uint64_t nulls = 0;
    for (uint64_t r = 0; r < col->length; r++) nulls += !((w[r >> 5] >> (r & 31)) & 1);
    if (col->length % 32) w[col->length / 32] &= (1u << (col->length % 32)) - 1u;   // bits past the length: 0
    if (!nulls) {
        if (col->validity) c->pool.free(col->validity, c->stream);
        col->validity = nullptr;
        col->null_count = 0;
        return PIMDAL_OK;
    }
    set_validity_words(c, col, w.data(), nulls);
    PD_API_END
}

namespace {
struct ExportPriv {
    void* data;
    void* valid;
    const void* buffers[2];
};
void release_array(struct ArrowArray* a) {
    if (!a || !a->release) return;
    auto* p = (ExportPriv*)a->private_data;
    free(p->data);
    free(p->valid);
    delete p;
    a->release = nullptr;
}
void release_schema(struct ArrowSchema* s) {
    if (!s || !s->release) return;
    s->release = nullptr;
}
}  // namespace

extern "C" int pimdal_col_export(pimdal_ctx* c, const pimdal_col* col, struct ArrowArray* out, struct ArrowSchema* os) {
    PD_API_BEGIN
    PD_REQUIRE(c && col && out && os, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    const uint64_t bytes = col->length * type_width(col->type);
    void* host = aligned_alloc(64, std::max<uint64_t>((bytes + 63) & ~63ull, 64));
    if (!host) throw std::bad_alloc();
    try {
        stage_d2h(c, host, col->data, bytes);
    } catch (...) {
        free(host);
        throw;
    }
    void* vhost = nullptr;
    if (col->validity) {
        const uint64_t vb = valid_words(col->length) * 4;
        vhost = aligned_alloc(64, std::max<uint64_t>((vb + 63) & ~63ull, 64));
        if (!vhost) {
            free(host);
            throw std::bad_alloc();
        }
        try {
            stage_d2h(c, vhost, col->validity, vb);
        } catch (...) {
            free(host);
            free(vhost);
            throw;
        }
    }
    auto* p = new ExportPriv;
    p->data = host;
    p->valid = vhost;
    p->buffers[0] = vhost;
    p->buffers[1] = host;
    out->length = (int64_t)col->length;
    out->null_count = (int64_t)col->null_count;
    out->offset = 0;
    out->n_buffers = 2;
    out->n_children = 0;
    out->buffers = p->buffers;
    out->children = nullptr;
    out->dictionary = nullptr;
    out->release = release_array;
    out->private_data = p;
    os->format = arrow_format(col->type);
    os->name = "";
    os->metadata = nullptr;
    os->flags = 0;
    os->n_children = 0;
    os->children = nullptr;
    os->dictionary = nullptr;
    os->release = release_schema;
    os->private_data = nullptr;
    PD_API_END
}

// ---------------- C ABI: captured plans ----------------

extern "C" int pimdal_plan_begin(pimdal_ctx* c, pimdal_plan** out) {
    PD_API_BEGIN
    PD_REQUIRE(c && out, PIMDAL_E_INVALID, "NULL argument");
    PD_REQUIRE(!c->capture, PIMDAL_E_INVALID, "a plan is already being captured on this context");
    auto* p = new pimdal_plan();
    p->ctx = c;
    // relaxed: the pool may take new device memory for the plan while capturing
    const hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
        delete p;
        PD_HIP(e);
    }
    p->capturing = true;
    p->errors_at_begin = pd::error_count();
    c->capture = p;
    *out = p;
    PD_API_END
}

static void plan_release(pimdal_plan* p) {
    pimdal_ctx* c = p->ctx;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
    for (void* b : p->blocks) c->pool.free(b, c->stream);
    p->blocks.clear();
}

extern "C" int pimdal_plan_end(pimdal_plan* p) {
    PD_API_BEGIN
    PD_REQUIRE(p && p->capturing, PIMDAL_E_INVALID, "plan is not being captured");
    pimdal_ctx* c = p->ctx;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    p->capturing = false;
    c->capture = nullptr;
    const bool failed = pd::error_count() != p->errors_at_begin;
    if (e != hipSuccess || failed || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        plan_release(p);
        PD_REQUIRE(!failed, PIMDAL_E_INVALID, "a call inside the capture failed; the plan holds nothing to launch");
        PD_HIP(e);
        pd::fail(PIMDAL_E_INTERNAL, "capture produced no graph");
    }
    const hipError_t ei = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        p->exec = nullptr;
        plan_release(p);
        PD_HIP(ei);
    }
    PD_API_END
}

extern "C" int pimdal_plan_launch(pimdal_plan* p) {
    PD_API_BEGIN
    PD_REQUIRE(p && p->exec && !p->capturing, PIMDAL_E_INVALID, "plan has no instantiated graph");
    PD_REQUIRE(!p->ctx->capture, PIMDAL_E_INVALID, "cannot launch a plan while capturing another");
    PD_HIP(hipGraphLaunch(p->exec, p->ctx->stream));
    PD_API_END
}

extern "C" int pimdal_plan_destroy(pimdal_plan* p) {
    PD_API_BEGIN
    if (!p) return PIMDAL_OK;
    pimdal_ctx* c = p->ctx;
    if (p->capturing) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        c->capture = nullptr;
        p->capturing = false;
    }
    (void)hipStreamSynchronize(c->stream);   // no replay still running on its memory
    plan_release(p);
    delete p;
    PD_API_END
}
