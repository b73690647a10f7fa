// pd_runtime.hpp — host runtime shared by the operators: the context (device,
// stream, caching device allocator under an HBM budget, pinned staging,
// per-kernel hipEvent timing) and error propagation to the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <map>
#include <mutex>
#include <string>
#include <vector>
#include <stdexcept>
#include <tuple>

#include "../../include/pimdal_gpu.h"

namespace pd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what, const char* file, int line);
#define PD_HIP(x) ::pd::hip_check((x), #x, __FILE__, __LINE__)
#define PD_REQUIRE(cond, code, msg) do { if (!(cond)) ::pd::fail((code), (msg)); } while (0)

void set_last_error(const std::string& m);
// Errors reported at the C ABI by this thread so far (a plan's capture
// checks that no call inside it failed).
uint64_t error_count();

// Caching device allocator.  Blocks are rounded to 256 B (small) or 2 MiB
// (large) and kept per rounded size after free; a cached block is reused on
// the stream it was freed on, or on any stream once its free-event has
// completed.  Reserved bytes (live + cached) never exceed the budget: cached
// blocks are released first, then the request fails with PIMDAL_E_NOMEM.
class DevicePool {
public:
    void init(int device, uint64_t budget);
    void* alloc(uint64_t bytes, hipStream_t stream);
    void free(void* p, hipStream_t stream);
    void release_cached();
    void destroy();
    uint64_t live() const { return live_; }
    uint64_t reserved() const { return reserved_; }
    uint64_t peak() const { return peak_; }
    uint64_t budget() const { return budget_; }
    // Takes the pool lock; when reserved bytes exceed the new budget,
    // cached blocks are freed (largest first) until they fit, as far as
    // cached memory allows (live blocks stay).
    void set_budget(uint64_t b);
private:
    void trim_locked(uint64_t need);
    struct Block { void* p; uint64_t size; hipStream_t stream; hipEvent_t ev; };
    static uint64_t round_size(uint64_t b);
    hipEvent_t take_event();            // an idle event (created when none is idle)
    void give_event(hipEvent_t e);      // back to the idle list
    std::vector<hipEvent_t> idle_events_;
    int device_ = 0;
    uint64_t budget_ = 0, live_ = 0, reserved_ = 0, peak_ = 0;
    std::multimap<uint64_t, Block> cached_;
    std::map<void*, uint64_t> live_blocks_;
    std::mutex mu_;
};

struct KernelStat { uint64_t launches = 0; double total_ms = 0.0; };

}  // namespace pd

// A captured operator sequence (hipGraph) and the device memory its nodes
// use, held until the plan is destroyed.
struct pimdal_plan {
    pimdal_ctx* ctx = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void*> blocks;       // pool blocks the captured kernels use
    uint64_t errors_at_begin = 0;
    bool capturing = false;
};

struct pimdal_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    pd::DevicePool pool;
    // pinned staging for host <-> device column copies (double buffered)
    void* pinned[2] = {nullptr, nullptr};
    hipEvent_t pinned_ev[2] = {nullptr, nullptr};
    uint64_t pinned_bytes = 0;
    // pinned words for small device-to-host reads (read_small)
    void* readback = nullptr;
    static constexpr uint64_t READBACK_BYTES = 16384;   // pinned bytes for read_small
    // profiling: hipEvent pairs per kernel name on the stream it ran on
    bool profiling = false;
    struct Pending { std::string name; hipEvent_t e0, e1; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, pd::KernelStat> stats;
    // a second stream for independent work of one call (created on first
    // use, default priority) and its fork / join events
    hipStream_t side = nullptr;
    hipEvent_t side_ev[2] = {nullptr, nullptr};
    // the plan being captured on `stream` (nullptr: eager).  While set,
    // scratch memory belongs to the plan, launches carry no profiling events,
    // host reads of device results are refused, and the radix passes use
    // ordinary launches instead of cooperative ones.
    pimdal_plan* capture = nullptr;
    // device properties
    int cu_count = 256;
    int xcd_count = 8;
    // tuning parameters (autotuner writes these)
    std::map<std::string, int64_t> tune;
    int64_t tune_get(const std::string& k, int64_t dflt) const {
        auto it = tune.find(k);
        return it == tune.end() ? dflt : it->second;
    }
};

struct pimdal_col {
    pimdal_type type;
    uint64_t length;
    uint64_t capacity;
    void* data;      // device memory
    bool owned;      // allocated from the context pool
    // validity bitmap (device, Arrow layout: bit i % 8 of byte i / 8, 1 =
    // valid, ceil(length / 8) bytes rounded up to whole 32-bit words), pool
    // memory owned by the column; nullptr = no nulls
    void* validity = nullptr;
    uint64_t null_count = 0;
};

namespace pd {

uint32_t type_width(pimdal_type t);
// Validity bitmaps: 32-bit words for n rows; set a column's from host words.
uint64_t valid_words(uint64_t n);
void set_validity_words(pimdal_ctx* c, pimdal_col* col, const uint32_t* bits, uint64_t nulls);

// Temporary device buffer from the context pool, freed on scope exit
// (stream ordered on the context stream).
// While a plan is being captured the buffer belongs to the plan (freed
// when the plan is destroyed): the captured kernels use it at every replay.
class Scratch {
public:
    Scratch(pimdal_ctx* c, uint64_t bytes, hipStream_t s = nullptr)
        : c_(c), s_(s ? s : c->stream), p_(bytes ? c->pool.alloc(bytes, s ? s : c->stream) : nullptr) {
        if (p_ && c->capture) {
            c->capture->blocks.push_back(p_);
            plan_owned_ = true;
        }
    }
    ~Scratch() { if (p_ && !plan_owned_) c_->pool.free(p_, s_); }
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
    template <typename T> T* as() const { return reinterpret_cast<T*>(p_); }
    void* get() const { return p_; }
private:
    pimdal_ctx* c_;
    hipStream_t s_;
    void* p_;
    bool plan_owned_ = false;
};

// Kernel launch with optional hipEvent timing on the launch stream.
hipEvent_t take_event(pimdal_ctx* c);
void record_launch(pimdal_ctx* c, const char* name, hipEvent_t e0, hipStream_t s);
// The context's side stream, started after the work enqueued so far on s
// (fork); join makes s wait for everything enqueued on the side stream.
hipStream_t side_fork(pimdal_ctx* c, hipStream_t s);
void side_join(pimdal_ctx* c, hipStream_t s);
// Joins the side stream into s at join() or, on an error path, when it goes
// out of scope: buffers declared before it are freed on s only after the
// side stream's work.
class SideJoin {
public:
    SideJoin(pimdal_ctx* c, hipStream_t s, bool on) : c_(c), s_(s), on_(on) {}
    ~SideJoin() {
        if (on_) {
            (void)hipEventRecord(c_->side_ev[1], c_->side);
            (void)hipStreamWaitEvent(s_, c_->side_ev[1], 0);
        }
    }
    void join() {
        if (on_) side_join(c_, s_);
        on_ = false;
    }
    SideJoin(const SideJoin&) = delete;
    SideJoin& operator=(const SideJoin&) = delete;
private:
    pimdal_ctx* c_;
    hipStream_t s_;
    bool on_;
};

template <typename... KArgs, typename... Args>
void launch(pimdal_ctx* c, const char* name, void (*kernel)(KArgs...), dim3 grid, dim3 block,
            uint32_t shmem, hipStream_t s, Args... args) {
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
    const bool prof = c->profiling && !c->capture;   // no timing events inside a captured plan
    hipEvent_t e0 = nullptr;
    if (prof) {
        e0 = take_event(c);
        PD_HIP(hipEventRecord(e0, s));
    }
    hipLaunchKernelGGL(kernel, grid, block, shmem, s, static_cast<KArgs>(args)...);
    PD_HIP(hipGetLastError());
    if (prof) record_launch(c, name, e0, s);
}

// Cooperative launch: the runtime starts the grid only if every block is
// resident at once (or fails), so blocks may wait on one another's
// progress.  Persistent kernels that walk tiles in a fixed order (block b
// takes tiles b, b + grid, ...) use it instead of a tile counter: one
// device-wide atomic per tile serializes at about 11 ns each.
// Returns false, with nothing launched, when the runtime refuses the grid:
// too large to be resident at once (fewer CUs available than the device
// reports, for example), or the device's cooperative dispatch not available
// in its present state (seen while another process on the same GPU has run
// cooperative launches).  Other errors throw.
template <typename... KArgs, typename... Args>
bool launch_coop(pimdal_ctx* c, const char* name, void (*kernel)(KArgs...), dim3 grid, dim3 block,
                 uint32_t shmem, hipStream_t s, Args... args) {
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return true;
    if (c->capture) return false;   // plans use ordinary launches: taken as a refusal
    std::tuple<KArgs...> vals(static_cast<KArgs>(args)...);
    void* ptrs[sizeof...(KArgs) > 0 ? sizeof...(KArgs) : 1];
    std::apply([&](auto&... v) { int i = 0; ((ptrs[i++] = (void*)&v), ...); }, vals);
    hipEvent_t e0 = nullptr;
    if (c->profiling) {
        e0 = take_event(c);
        PD_HIP(hipEventRecord(e0, s));
    }
    const hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kernel), grid, block, ptrs, shmem, s);
    if (e == hipErrorCooperativeLaunchTooLarge || e == hipErrorIllegalState) {
        (void)hipGetLastError();   // clear the sticky launch error
        if (e0) c->event_pool.push_back(e0);
        return false;
    }
    PD_HIP(e);
    if (c->profiling) record_launch(c, name, e0, s);
    return true;
}

// Blocks of `kernel` (at `block` threads, no dynamic LDS) that one CU holds
// at once; at least 1.
template <typename K>
uint32_t blocks_per_cu(K kernel, int block) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) != hipSuccess || nb < 1) nb = 1;
    return (uint32_t)nb;
}

// Fill device memory with a small kernel (graph-capture safe, unlike
// hipMemsetAsync).
void fill_u32(pimdal_ctx* c, void* p, uint32_t value, uint64_t count, hipStream_t s);
void fill_u64(pimdal_ctx* c, void* p, uint64_t value, uint64_t count, hipStream_t s);
void copy_d2d(pimdal_ctx* c, void* dst, const void* src, uint64_t bytes, hipStream_t s);

uint64_t read_u64(pimdal_ctx* c, const void* dev, hipStream_t s);
// Copies `bytes` (<= 64 through pinned memory) from the device and waits.
// Refused (PIMDAL_E_UNSUPPORTED) while a plan is being captured.
void read_small(pimdal_ctx* c, void* dst, const void* dev, uint64_t bytes, hipStream_t s);

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

}  // namespace pd

// Entry points that wait for the device or return host results cannot be
// recorded into a plan.
#define PD_NO_CAPTURE(c)                                                                            \
    PD_REQUIRE(!(c)->capture, PIMDAL_E_UNSUPPORTED,                                                 \
               "this call waits for results on the host and cannot be captured into a plan; use "   \
               "an asynchronous device form")

// C-ABI boundary: run `body`, translate exceptions into return codes.
#define PD_API_BEGIN try {
#define PD_API_END                                                       \
    return PIMDAL_OK;                                                    \
    } catch (const pd::Error& e) {                                       \
        pd::set_last_error(e.what());                                    \
        return e.code;                                                   \
    } catch (const std::bad_alloc&) {                                    \
        pd::set_last_error("host allocation failed");                    \
        return PIMDAL_E_NOMEM;                                           \
    } catch (const std::exception& e) {                                  \
        pd::set_last_error(e.what());                                    \
        return PIMDAL_E_INTERNAL;                                        \
    }
