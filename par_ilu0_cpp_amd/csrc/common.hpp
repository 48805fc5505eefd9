// Shared host/device helpers for the gfx950 ILU(k)-GMRES library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace pilu {

// ---------------------------------------------------------------------
// Error handling: every HIP call on the product path is checked and turns
// into std::runtime_error with the call site.
// ---------------------------------------------------------------------
[[noreturn]] void hip_fail(hipError_t e, const char* what, const char* file, int line);

#define PILU_HIP(call)                                                   \
    do {                                                                 \
        hipError_t pilu_e_ = (call);                                     \
        if (pilu_e_ != hipSuccess) ::pilu::hip_fail(pilu_e_, #call, __FILE__, __LINE__); \
    } while (0)

// Device allocation behind DevBuf.  Inside a SetupScope (per host thread)
// freed blocks are kept and handed out again (best fit, at most twice the
// request) instead of going back to the driver; the outermost scope's end
// synchronises its stream and frees them.  All work on such blocks inside a
// scope must be ordered on that one stream.
void* dev_alloc(size_t bytes);
void dev_free(void* p, size_t bytes);

class SetupScope {
public:
    explicit SetupScope(hipStream_t s);
    ~SetupScope();
    SetupScope(const SetupScope&) = delete;
    SetupScope& operator=(const SetupScope&) = delete;
};

// Owning device allocation (dev_alloc / dev_free).
template <typename T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t count) { resize(count); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    // Grows (never shrinks) to at least `count` elements; contents are not kept.
    void reserve(size_t count) { if (count > n_) resize(count); }
    void resize(size_t count) {
        release();
        if (count) p_ = static_cast<T*>(dev_alloc(count * sizeof(T)));
        n_ = count;
    }
    void release() {
        if (p_) dev_free(p_, n_ * sizeof(T));
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// ---------------------------------------------------------------------
// "Value as flag": a published FP64 value is stored with one 64-bit
// relaxed agent-scope atomic store; consumers spin with relaxed agent-scope
// atomic loads until the slot no longer holds kPending.  kPending is a
// signalling NaN; arithmetic never produces one and publish() turns any NaN
// into the canonical quiet NaN, so a published value can never look pending.
// On gfx950 these compile to plain global_load/store with sc1 (coherent at
// device scope across XCDs) and need no L2 writeback or invalidate.
// ---------------------------------------------------------------------
constexpr uint64_t kPending = 0x7FF4DEADBEEF0001ull;
constexpr uint64_t kQuietNan = 0x7FF8000000000000ull;

__device__ __forceinline__ uint64_t peek_bits(const double* p) {
    return __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void publish(double* p, double v) {
    uint64_t b = __double_as_longlong(v);
    if (v != v) b = kQuietNan;
    __hip_atomic_store(reinterpret_cast<uint64_t*>(p), b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void publish_raw(uint64_t* p, uint64_t b) {
    __hip_atomic_store(p, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t peek_raw(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int peek_int(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void publish_int(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------
// Work tickets.  Sync-free kernels hand out work in dependency order with
// one atomic ticket per workgroup: a workgroup can only wait on work held
// by a workgroup that took its ticket earlier and is therefore resident,
// so progress never depends on the hardware's dispatch order.  The last
// workgroup to finish puts the counters back to zero, so the kernel needs
// no host-side reset and can be replayed inside a graph.
// ctr[0] = next ticket, ctr[1] = finished workgroups.
// ---------------------------------------------------------------------
__device__ __forceinline__ int take_ticket(int* ctr) {
    __shared__ int s_ticket;
    if (threadIdx.x == 0)
        s_ticket = __hip_atomic_fetch_add(&ctr[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return s_ticket;
}

// nblocks: workgroups that take tickets from ctr (the whole grid unless a
// launch splits its workgroups over several ticket counters).
__device__ __forceinline__ void retire_ticket(int* ctr, int nblocks = -1) {
    __syncthreads();
    if (threadIdx.x == 0) {
        int done = __hip_atomic_fetch_add(&ctr[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (done == (nblocks < 0 ? static_cast<int>(gridDim.x) : nblocks) - 1) {
            __hip_atomic_store(&ctr[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctr[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// v of the lane K places up within the lane's row of 16 (DPP row shift; 0
// past the row's end).
template <int K>
__device__ __forceinline__ double dpp_row_shl(double v) {
    static_assert(K >= 1 && K <= 15, "row shift");
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b & 0xffffffffu), 0x100 | K, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), 0x100 | K, 0xf, 0xf, true);
    return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                                                       static_cast<uint32_t>(lo)));
}

// Sum over G adjacent lanes (G a power of two <= 64), valid in the group's
// first lane (the others hold partial sums).  Lane 0 adds lane G/2's value,
// then G/4's, ..., then lane 1's: the addition tree of the xor butterfly,
// so the result is deterministic for a given G.  Offsets below 16 are DPP
// row shifts (no LDS traffic); 16 and 32 cross rows through ds_bpermute.
// Every lane of the group must be active.
template <int G>
__device__ __forceinline__ double group_sum(double v) {
    if constexpr (G >= 64) v += __shfl_down(v, 32, 64);
    if constexpr (G >= 32) v += __shfl_down(v, 16, 64);
    if constexpr (G >= 16) v += dpp_row_shl<8>(v);
    if constexpr (G >= 8) v += dpp_row_shl<4>(v);
    if constexpr (G >= 4) v += dpp_row_shl<2>(v);
    if constexpr (G >= 2) v += dpp_row_shl<1>(v);
    return v;
}

template <int G>
__device__ __forceinline__ int group_max(int v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        int o = __shfl_xor(v, off, G);
        v = o > v ? o : v;
    }
    return v;
}

// True for every lane whose G-lane group has `pred` set on all lanes.
template <int G>
__device__ __forceinline__ bool group_all(bool pred) {
    const uint64_t m = __ballot(pred);
    if constexpr (G == 64) {
        return m == ~0ull;
    } else {
        const int lane = threadIdx.x & 63;
        const int shift = lane & ~(G - 1);
        const uint64_t gm = (1ull << G) - 1ull;
        return ((m >> shift) & gm) == gm;
    }
}

__device__ __forceinline__ void spin_pause() { __builtin_amdgcn_s_sleep(1); }

// Watchdog for sync-free waits: a wait that outlives `limit` ticks of the
// constant-rate device wall clock gives up, raises kErrTimeout in the
// kernel's error word and lets the kernel drain (the waiting row publishes
// NaN), so a scheduling fault turns into an error instead of a hang.
constexpr int kErrTimeout = 1 << 8;

struct Watchdog {
    uint64_t start;
    uint64_t limit;
    __device__ explicit Watchdog(uint64_t lim) : start(wall_clock64()), limit(lim) {}
    __device__ bool expired() const { return wall_clock64() - start > limit; }
};

// Ticks of the device wall clock in `seconds` (host side).
uint64_t watchdog_ticks(double seconds);

// Host-side phase times of the set-up path, printed to stderr when the
// environment holds PAR_ILU_TIMING (each mark first waits for `stream`).
bool phase_timing_enabled();
struct PhaseTimer {
    const char* scope;
    hipStream_t stream;
    bool on;
    double t0;
    PhaseTimer(const char* sc, hipStream_t s);
    void mark(const char* what);
};

inline int ceil_div(long a, long b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace pilu
