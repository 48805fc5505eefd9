// lljbash::IluSolver on the GPU: pattern setup and level analysis,
// numeric ILU and the triangular sweeps.
// Interface: include/ilu_solver.hpp (reference include/ilu_solver.hpp:8-30,
// behaviour src/ilu_solver.cpp:225-831).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <stdexcept>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "ilu_internal.hpp"
#include "ilu_kernels.hpp"
#include "ilu_sweep_ws.hpp"
#include "iluk_device.hpp"

namespace pilu {

[[noreturn]] void hip_fail(hipError_t e, const char* what, const char* file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", static_cast<int>(e),
                  hipGetErrorString(e), what, file, line);
    throw std::runtime_error(buf);
}

namespace {
// Blocks freed inside the current thread's SetupScope.
struct ScopeCache {
    int depth = 0;
    hipStream_t stream = nullptr;
    std::vector<std::pair<void*, size_t>> blocks;
};
thread_local ScopeCache g_scope;
}  // namespace

void* dev_alloc(size_t bytes) {
    if (g_scope.depth > 0) {
        size_t best = 0;
        int at = -1;
        for (size_t i = 0; i < g_scope.blocks.size(); ++i) {
            const size_t b = g_scope.blocks[i].second;
            if (b >= bytes && b / 2 <= bytes && (at < 0 || b < best)) {
                best = b;
                at = static_cast<int>(i);
            }
        }
        if (at >= 0) {
            void* p = g_scope.blocks[at].first;
            g_scope.blocks.erase(g_scope.blocks.begin() + at);
            return p;
        }
    }
    void* p = nullptr;
    PILU_HIP(hipMalloc(&p, bytes));
    return p;
}

void dev_free(void* p, size_t bytes) {
    if (!p) return;
    if (g_scope.depth > 0) {
        g_scope.blocks.emplace_back(p, bytes);
        return;
    }
    (void)hipFree(p);
}

SetupScope::SetupScope(hipStream_t s) {
    if (g_scope.depth++ == 0) g_scope.stream = s;
}

SetupScope::~SetupScope() {
    if (--g_scope.depth > 0 || g_scope.blocks.empty()) return;
    (void)hipStreamSynchronize(g_scope.stream);  // no kernel may still use a block
    for (auto& b : g_scope.blocks) (void)hipFree(b.first);
    g_scope.blocks.clear();
}

int lanes_for_avg(double avg_len) {
    int g = 4;
    while (g < 64 && g < avg_len) g <<= 1;
    return g;
}

hipStream_t ilu_stream(const lljbash::IluSolver::Ext* ext) { return ext->stream; }

uint64_t watchdog_ticks(double seconds) {
    static int rate_khz = 0;
    if (rate_khz <= 0) {
        int dev = 0;
        PILU_HIP(hipGetDevice(&dev));
        PILU_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
        if (rate_khz <= 0) rate_khz = 100000;
    }
    return static_cast<uint64_t>(seconds * rate_khz * 1000.0);
}

// Error-word slots of IluSolver::Ext::err.
enum ErrSlot { kErrFactor = 0, kErrSweep = 1, kErrSorted = 2, kErrDiag = 3, kErrMaxLev = 4, kErrRowPtr = 5,
               kErrLevels = 6, kErrPivotRow = 7, kErrExtent = 8 /* 3 words */, kErrMaxLevLU = 12 /* 2 words */ };

void check_sweep_errors(lljbash::IluSolver::Ext* e) {
    int v = 0;
    PILU_HIP(hipMemcpyAsync(&v, e->err.get() + kErrSweep, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    PILU_HIP(hipStreamSynchronize(e->stream));
    if (v & kErrTimeout) {
        PILU_HIP(hipMemsetAsync(e->err.get() + kErrSweep, 0, sizeof(int), e->stream));
        e->y_armed = false;
        throw std::runtime_error("triangular sweep: dependency wait timed out (watchdog)");
    }
    if (v & kErrStream) {
        PILU_HIP(hipMemsetAsync(e->err.get() + kErrSweep, 0, sizeof(int), e->stream));
        e->y_armed = false;
        throw std::runtime_error("wave-specialised sweep: record stream row or face id out of range");
    }
}

double* device_scratch();

bool phase_timing_enabled() {
    static const bool on = std::getenv("PAR_ILU_TIMING") != nullptr;
    return on;
}

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

PhaseTimer::PhaseTimer(const char* sc, hipStream_t s)
    : scope(sc), stream(s), on(phase_timing_enabled()), t0(on ? now_ms() : 0.0) {}

void PhaseTimer::mark(const char* what) {
    if (!on) return;
    (void)hipStreamSynchronize(stream);
    const double t = now_ms();
    std::fprintf(stderr, "[par-ilu timing] %s: %s %.2f ms\n", scope, what, t - t0);
    t0 = t;
}

// Raised when a sweep structure is missing during a no-allocation apply
// (building it would allocate device memory).
void refuse_if_no_alloc(const lljbash::IluSolver::Ext* e, const char* what) {
    if (e->no_alloc)
        throw std::runtime_error(std::string("no-allocation apply: ") + what +
                                 " not built for the current sweep configuration (run IluSolverSubstituteDeviceEx "
                                 "once first)");
}

namespace {

// f(begin, end) over [0, n) in contiguous chunks on up to 16 host threads.
template <typename F>
void parallel_for(long n, F f) {
    const long hw = static_cast<long>(std::thread::hardware_concurrency());
    const long t = std::max(1L, std::min({hw > 0 ? hw : 1L, 16L, n / 65536 + 1}));
    if (t == 1) {
        f(0L, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(static_cast<size_t>(t));
    for (long k = 0; k < t; ++k) th.emplace_back(f, n * k / t, n * (k + 1) / t);
    for (auto& x : th) x.join();
}

// Host copy of the pattern, the diagonal positions and the partition,
// copied on first use after the pattern or the partition changed (only the
// wave-specialised sweep's stream builder needs them).
void host_pattern(lljbash::IluSolver::Ext* e) {
    const size_t n = static_cast<size_t>(e->n);
    if (e->h_pattern_version != e->pattern_version) {
        e->h_rp.resize(n + 1);
        e->h_dg.resize(n);
        e->h_ci.resize(static_cast<size_t>(e->nnz));
        PILU_HIP(hipMemcpy(e->h_rp.data(), e->rp.get(), sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
        PILU_HIP(hipMemcpy(e->h_dg.data(), e->dg.get(), sizeof(int) * n, hipMemcpyDeviceToHost));
        PILU_HIP(hipMemcpy(e->h_ci.data(), e->ci.get(), sizeof(int) * e->h_ci.size(), hipMemcpyDeviceToHost));
        e->h_pattern_version = e->pattern_version;
    }
    if (e->nblk > 0 && e->h_part_version != e->part_version) {
        e->h_block_of.resize(n);
        PILU_HIP(hipMemcpy(e->h_block_of.data(), e->block_of.get(), sizeof(int) * n, hipMemcpyDeviceToHost));
        e->h_part_version = e->part_version;
    }
}

using Ext = lljbash::IluSolver::Ext;

inline int grid_for(long n, int per_block) { return std::max(1, ceil_div(n, per_block)); }

int read_int(const int* d, hipStream_t s) {
    int h = 0;
    PILU_HIP(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, s));
    PILU_HIP(hipStreamSynchronize(s));
    return h;
}

template <bool UPPER, int G>
void levels_g(Ext* e, const int* block_of, int* level, int* maxlev, uint64_t wd) {
    constexpr int rows = (kBlock / 64) * (64 / G);
    k_levels<UPPER, G><<<grid_for(e->n, rows), kBlock, 0, e->stream>>>(e->n, e->rp.get(), e->ci.get(), e->dg.get(),
                                                                        level, maxlev, e->ctr.get(),
                                                                        e->err.get() + kErrLevels, wd, block_of);
}

// Level analysis of one direction: lanes per row from the average number
// of dependencies per row.
template <bool UPPER>
void launch_levels(Ext* e, const int* block_of, int* level, int* maxlev, uint64_t wd) {
    const double avg = e->n > 0 ? static_cast<double>(e->nnz - e->n) / (2.0 * e->n) : 0.0;
    if (avg <= 8) levels_g<UPPER, 1>(e, block_of, level, maxlev, wd);
    else if (avg <= 32) levels_g<UPPER, 8>(e, block_of, level, maxlev, wd);
    else levels_g<UPPER, 64>(e, block_of, level, maxlev, wd);
}

// Rows sorted by (level, row) and each level's first position, from the
// levels in `level` (nlev levels).
void sort_by_level(Ext* e, const int* level, int nlev, DevBuf<int>& perm, DevBuf<int>& lev_off) {
    const int n = e->n;
    hipStream_t s = e->stream;
    DevBuf<int> keys_out(n), rows(n);
    k_iota<<<grid_for(n, 256), 256, 0, s>>>(n, rows.get());
    perm.reserve(n);
    unsigned end_bit = 1;
    while (end_bit < 31 && (1u << end_bit) <= static_cast<unsigned>(nlev)) ++end_bit;
    size_t bytes = 0;
    PILU_HIP(rocprim::radix_sort_pairs(nullptr, bytes, level, keys_out.get(), rows.get(), perm.get(),
                                       static_cast<unsigned>(n), 0u, end_bit, s));
    e->sort_tmp.reserve(bytes);
    PILU_HIP(rocprim::radix_sort_pairs(e->sort_tmp.get(), bytes, level, keys_out.get(), rows.get(), perm.get(),
                                       static_cast<unsigned>(n), 0u, end_bit, s));
    lev_off.reserve(static_cast<size_t>(nlev) + 1);
    // every level 0..nlev-1 holds a row (a row's level is one more than its
    // deepest dependency's); the fill only keeps the array defined if not
    k_fill_int<<<grid_for(static_cast<long>(nlev) + 1, 256), 256, 0, s>>>(static_cast<long>(nlev) + 1, lev_off.get(), n);
    k_level_starts<<<grid_for(n, 256), 256, 0, s>>>(n, nlev, keys_out.get(), lev_off.get());
    PILU_HIP(hipGetLastError());
    PILU_HIP(hipStreamSynchronize(s));
}

// Levels of both dependency graphs in one launch (the two chains run at the
// same time) and the rows of each sorted by (level, row).
void build_schedules(Ext* e) {
    const int n = e->n;
    hipStream_t s = e->stream;
    DevBuf<int> level_u(n);
    int* level_l = e->level.get();
    int* maxlev = e->err.get() + kErrMaxLevLU;  // [0] L, [1] U
    const uint64_t wd = watchdog_ticks(e->watchdog_s);
    k_fill_int<<<grid_for(n, 256 * 8), 256, 0, s>>>(n, level_l, -1);
    k_fill_int<<<grid_for(n, 256 * 8), 256, 0, s>>>(n, level_u.get(), -1);
    PILU_HIP(hipMemsetAsync(maxlev, 0, 2 * sizeof(int), s));
    const double avg = n > 0 ? static_cast<double>(e->nnz - n) / (2.0 * n) : 0.0;
    auto run = [&](auto g) {
        constexpr int G = decltype(g)::value;
        constexpr int rows = (kBlock / 64) * (64 / G);
        const int nb = grid_for(n, rows);
        k_levels_lu<G><<<2 * nb, kBlock, 0, s>>>(n, e->rp.get(), e->ci.get(), e->dg.get(), level_l, level_u.get(),
                                                 maxlev, e->ctr.get(), e->err.get() + kErrLevels, wd, nb);
    };
    if (avg <= 8) run(std::integral_constant<int, 1>{});
    else if (avg <= 32) run(std::integral_constant<int, 8>{});
    else run(std::integral_constant<int, 64>{});
    PILU_HIP(hipGetLastError());
    int ml[2] = {0, 0};
    PILU_HIP(hipMemcpyAsync(ml, maxlev, sizeof(ml), hipMemcpyDeviceToHost, s));
    PILU_HIP(hipStreamSynchronize(s));
    if (read_int(e->err.get() + kErrLevels, s) & kErrTimeout)
        throw std::runtime_error("level analysis: dependency wait timed out (watchdog)");
    e->nlev_l = ml[0] + 1;
    e->nlev_u = ml[1] + 1;
    sort_by_level(e, level_l, e->nlev_l, e->perm_l, e->lvoff_l);
    sort_by_level(e, level_u.get(), e->nlev_u, e->perm_u, e->lvoff_u);
}

__global__ void k_block_keys(int n, const int* __restrict__ block_of, const int* __restrict__ pos,
                             const int* __restrict__ level, int lbits, unsigned* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = (static_cast<unsigned>(pos[block_of[i]]) << lbits) | static_cast<unsigned>(level[i]);
}

// Schedule of the block-local sweeps in one direction: levels restricted to
// each block's own rows, rows sorted by (processing position of their
// block, local level, row), and the group offsets of every (block, level).
// head[i]: the sorted key i starts a new (block, level) group.
__global__ void k_group_heads(int n, const unsigned* __restrict__ keys, unsigned char* __restrict__ head) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) head[i] = i == 0 || keys[i] != keys[i - 1];
}

// bl[p] = first group of block position p (the block part of the key is
// above lbits); untouched for positions without rows.
__global__ void k_block_first(int ngroups, const int* __restrict__ off, const unsigned* __restrict__ keys, int lbits,
                              int* __restrict__ bl) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    const unsigned p = keys[off[g]] >> lbits;
    if (g == 0 || p != (keys[off[g - 1]] >> lbits)) bl[p] = g;
}

// block_of[row]: block of every row (nblk blocks); pos[b]: processing
// position of block b.
void build_block_schedule(Ext* e, bool upper, const int* block_of, int nblk, const std::vector<int>& pos,
                          DevBuf<int>& perm, DevBuf<int>& blev, DevBuf<int>& loff) {
    const int n = e->n;
    hipStream_t s = e->stream;
    int* level = e->level.get();
    int* maxlev = e->err.get() + kErrMaxLev;
    k_fill_int<<<grid_for(n, 256 * 8), 256, 0, s>>>(n, level, -1);
    PILU_HIP(hipMemsetAsync(maxlev, 0, sizeof(int), s));
    const uint64_t wd = watchdog_ticks(e->watchdog_s);
    const int* bo = block_of;
    if (upper) launch_levels<true>(e, bo, level, maxlev, wd);
    else launch_levels<false>(e, bo, level, maxlev, wd);
    PILU_HIP(hipGetLastError());
    const int nlev = read_int(maxlev, s) + 1;
    if (read_int(e->err.get() + kErrLevels, s) & kErrTimeout)
        throw std::runtime_error("block level analysis: dependency wait timed out (watchdog)");
    int lbits = 1, pbits = 1;
    while ((1 << lbits) < nlev) ++lbits;
    while ((1 << pbits) < nblk) ++pbits;
    if (lbits + pbits > 31) throw std::runtime_error("block schedule: too many blocks x levels");
    DevBuf<int> dpos(nblk);
    PILU_HIP(hipMemcpy(dpos.get(), pos.data(), sizeof(int) * nblk, hipMemcpyHostToDevice));
    DevBuf<unsigned> keys(n), keys_out(n);
    DevBuf<int> rows(n);
    k_block_keys<<<grid_for(n, 256), 256, 0, s>>>(n, bo, dpos.get(), level, lbits, keys.get());
    k_iota<<<grid_for(n, 256), 256, 0, s>>>(n, rows.get());
    perm.reserve(n);
    size_t bytes = 0;
    const unsigned end_bit = static_cast<unsigned>(lbits + pbits);
    PILU_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys.get(), keys_out.get(), rows.get(), perm.get(),
                                       static_cast<unsigned>(n), 0u, end_bit, s));
    e->sort_tmp.reserve(bytes);
    PILU_HIP(rocprim::radix_sort_pairs(e->sort_tmp.get(), bytes, keys.get(), keys_out.get(), rows.get(), perm.get(),
                                       static_cast<unsigned>(n), 0u, end_bit, s));
    // group starts (a new (block, level) key) selected on the device into
    // loff, then the first group of every block
    DevBuf<unsigned char> head(n);
    DevBuf<unsigned> nsel(1);
    loff.reserve(static_cast<size_t>(n) + 1);
    k_group_heads<<<grid_for(n, 256), 256, 0, s>>>(n, keys_out.get(), head.get());
    PILU_HIP(hipGetLastError());
    rocprim::counting_iterator<int> idx(0);
    bytes = 0;
    PILU_HIP(rocprim::select(nullptr, bytes, idx, head.get(), loff.get(), nsel.get(), static_cast<size_t>(n),
                             s));
    e->sort_tmp.reserve(bytes);
    PILU_HIP(rocprim::select(e->sort_tmp.get(), bytes, idx, head.get(), loff.get(), nsel.get(),
                             static_cast<size_t>(n), s));
    unsigned ng = 0;
    PILU_HIP(hipMemcpyAsync(&ng, nsel.get(), sizeof(unsigned), hipMemcpyDeviceToHost, s));
    PILU_HIP(hipStreamSynchronize(s));
    const int ngroups = static_cast<int>(ng);
    k_fill_int<<<1, 64, 0, s>>>(1, loff.get() + ngroups, n);
    blev.reserve(static_cast<size_t>(nblk) + 1);
    k_fill_int<<<grid_for(nblk + 1, 256), 256, 0, s>>>(nblk + 1, blev.get(), -1);
    k_block_first<<<grid_for(ngroups, 256), 256, 0, s>>>(ngroups, loff.get(), keys_out.get(), lbits, blev.get());
    PILU_HIP(hipGetLastError());
    std::vector<int> bl(static_cast<size_t>(nblk) + 1);
    PILU_HIP(hipMemcpyAsync(bl.data(), blev.get(), sizeof(int) * bl.size(), hipMemcpyDeviceToHost, s));
    PILU_HIP(hipStreamSynchronize(s));
    bl[nblk] = ngroups;
    for (int p = nblk - 1; p >= 0; --p)
        if (bl[p] < 0) bl[p] = bl[p + 1];  // a block without rows
    PILU_HIP(hipMemcpy(blev.get(), bl.data(), sizeof(int) * bl.size(), hipMemcpyHostToDevice));
}

// Processing order of the blocks of a partition in one direction: a row
// depends on the rows of its L (forward) or U (backward) part, so block bj
// must run before block bi when a row of bi depends on a row of bj.  The
// blocks are ordered by their depth in that graph (longest dependency
// chain), ties by block id: a topological order that starts the blocks of
// one wavefront together.  Throws if the block graph has a cycle (the
// partition cannot be swept block by block).  Returns pos[block].
// Cross-block dependencies of one direction: per row, its entries whose
// column lies in another block (count pass, then keys (block of the row,
// block of the column) in 64 bits).
__global__ void k_cross_count(int n, const int* __restrict__ rp, const int* __restrict__ ci, const int* __restrict__ dg,
                              const int* __restrict__ bo, int upper, int* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = bo[i], p0 = upper ? dg[i] + 1 : rp[i], p1 = upper ? rp[i + 1] : dg[i];
    int c = 0;
    for (int p = p0; p < p1; ++p) c += bo[ci[p]] != b;
    cnt[i] = c;
}

__global__ void k_cross_write(int n, const int* __restrict__ rp, const int* __restrict__ ci, const int* __restrict__ dg,
                              const int* __restrict__ bo, int upper, const int* __restrict__ off,
                              unsigned long long* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = bo[i], p0 = upper ? dg[i] + 1 : rp[i], p1 = upper ? rp[i + 1] : dg[i];
    int k = off[i];
    for (int p = p0; p < p1; ++p) {
        const int c = bo[ci[p]];
        if (c != b) keys[k++] = (static_cast<unsigned long long>(static_cast<unsigned>(b)) << 32) | static_cast<unsigned>(c);
    }
}

// Processing order of the blocks of a partition in one direction: a row
// depends on the rows of its L (forward) or U (backward) part, so block bj
// must run before block bi when a row of bi depends on a row of bj.  The
// blocks are ordered by their depth in that graph (longest dependency
// chain), ties by block id: a topological order that starts the blocks of
// one wavefront together.  The distinct block edges are found on the
// device (cross-block entries, sorted and deduplicated); Kahn's algorithm
// runs on the host over them.  Throws if the block graph has a cycle (the
// partition cannot be swept block by block).  Returns pos[block].
std::vector<int> block_order(Ext* e, bool upper) {
    const int n = e->n, nblk = e->nblk;
    hipStream_t s = e->stream;
    DevBuf<int> cnt(static_cast<size_t>(n) + 1), off(static_cast<size_t>(n) + 1);
    k_fill_int<<<1, 1, 0, s>>>(1, cnt.get() + n, 0);
    k_cross_count<<<grid_for(n, 256), 256, 0, s>>>(n, e->rp.get(), e->ci.get(), e->dg.get(), e->block_of.get(),
                                                  upper ? 1 : 0, cnt.get());
    PILU_HIP(hipGetLastError());
    size_t bytes = 0;
    PILU_HIP(rocprim::exclusive_scan(nullptr, bytes, cnt.get(), off.get(), 0, static_cast<size_t>(n) + 1,
                                     rocprim::plus<int>(), s));
    e->sort_tmp.reserve(bytes);
    PILU_HIP(rocprim::exclusive_scan(e->sort_tmp.get(), bytes, cnt.get(), off.get(), 0, static_cast<size_t>(n) + 1,
                                     rocprim::plus<int>(), s));
    const int total = read_int(off.get() + n, s);
    std::vector<int> pred_ptr(static_cast<size_t>(nblk) + 1, 0), pred;
    if (total > 0) {
        DevBuf<unsigned long long> keys(total), sorted(total), uniq(total);
        DevBuf<unsigned> nuniq(1);
        k_cross_write<<<grid_for(n, 256), 256, 0, s>>>(n, e->rp.get(), e->ci.get(), e->dg.get(), e->block_of.get(),
                                                      upper ? 1 : 0, off.get(), keys.get());
        PILU_HIP(hipGetLastError());
        int pbits = 1;
        while ((1 << pbits) < nblk) ++pbits;
        bytes = 0;
        PILU_HIP(rocprim::radix_sort_keys(nullptr, bytes, keys.get(), sorted.get(), static_cast<unsigned>(total), 0u,
                                          static_cast<unsigned>(32 + pbits), s));
        e->sort_tmp.reserve(bytes);
        PILU_HIP(rocprim::radix_sort_keys(e->sort_tmp.get(), bytes, keys.get(), sorted.get(),
                                          static_cast<unsigned>(total), 0u, static_cast<unsigned>(32 + pbits), s));
        bytes = 0;
        PILU_HIP(rocprim::unique(nullptr, bytes, sorted.get(), uniq.get(), nuniq.get(), static_cast<size_t>(total),
                                 rocprim::equal_to<unsigned long long>(), s));
        e->sort_tmp.reserve(bytes);
        PILU_HIP(rocprim::unique(e->sort_tmp.get(), bytes, sorted.get(), uniq.get(), nuniq.get(),
                                 static_cast<size_t>(total), rocprim::equal_to<unsigned long long>(), s));
        unsigned nu = 0;
        PILU_HIP(hipMemcpyAsync(&nu, nuniq.get(), sizeof(unsigned), hipMemcpyDeviceToHost, s));
        PILU_HIP(hipStreamSynchronize(s));
        std::vector<unsigned long long> edges(nu);
        PILU_HIP(hipMemcpy(edges.data(), uniq.get(), sizeof(unsigned long long) * nu, hipMemcpyDeviceToHost));
        pred.resize(nu);
        for (unsigned k = 0; k < nu; ++k) {
            ++pred_ptr[static_cast<int>(edges[k] >> 32) + 1];
            pred[k] = static_cast<int>(edges[k] & 0xffffffffu);
        }
        for (int b = 0; b < nblk; ++b) pred_ptr[b + 1] += pred_ptr[b];
    }
    // successor lists and Kahn's algorithm, computing depths on the way
    std::vector<int> indeg(nblk), succ_ptr(static_cast<size_t>(nblk) + 1, 0), succ(pred.size());
    for (int b = 0; b < nblk; ++b) {
        indeg[b] = pred_ptr[b + 1] - pred_ptr[b];
        for (int q = pred_ptr[b]; q < pred_ptr[b + 1]; ++q) ++succ_ptr[pred[q] + 1];
    }
    for (int b = 0; b < nblk; ++b) succ_ptr[b + 1] += succ_ptr[b];
    std::vector<int> sf(succ_ptr.begin(), succ_ptr.end() - 1);
    for (int b = 0; b < nblk; ++b)
        for (int q = pred_ptr[b]; q < pred_ptr[b + 1]; ++q) succ[sf[pred[q]]++] = b;
    std::vector<int> depth(nblk, 0), queue;
    queue.reserve(nblk);
    for (int b = 0; b < nblk; ++b)
        if (indeg[b] == 0) queue.push_back(b);
    for (size_t h = 0; h < queue.size(); ++h) {
        const int b = queue[h];
        for (int q = succ_ptr[b]; q < succ_ptr[b + 1]; ++q) {
            const int c = succ[q];
            depth[c] = std::max(depth[c], depth[b] + 1);
            if (--indeg[c] == 0) queue.push_back(c);
        }
    }
    if (static_cast<int>(queue.size()) != nblk)
        throw std::runtime_error("sweep partition: the block dependency graph has a cycle");
    std::vector<int> order(nblk);
    for (int b = 0; b < nblk; ++b) order[b] = b;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return depth[a] < depth[c]; });
    std::vector<int> pos(nblk);
    for (int k = 0; k < nblk; ++k) pos[order[k]] = k;
    return pos;
}

// Rows of every block.  Consecutive rows mostly share a block, and a wave
// of them spans at most two blocks of a tiled or contiguous partition: the
// lanes in the block of the wave's first row add with one atomic, those in
// the block of its last row with another, any other lane on its own.
__global__ void k_block_counts(int n, const int* __restrict__ block_of, int* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool act = i < n;
    const int b = act ? block_of[i] : -1;
    const int last = 63 - __clzll(__ballot(act));  // highest active lane (act lanes are the low ones)
    const int b0 = __shfl(b, 0, 64), b1 = __shfl(b, last < 0 ? 0 : last, 64);
    const unsigned long long m0 = __ballot(act && b == b0);
    const unsigned long long m1 = __ballot(act && b == b1 && b != b0);
    if (lane == 0 && m0) atomicAdd(cnt + b0, __popcll(m0));
    if (m1 && lane == __ffsll(static_cast<long long>(m1)) - 1) atomicAdd(cnt + b1, __popcll(m1));
    if (act && b != b0 && b != b1) atomicAdd(cnt + b, 1);
}

__global__ void k_block_contiguous(int n, int blk, int* __restrict__ block_of) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) block_of[i] = i / blk;
}

// Installs the partition in e->block_of (device, blocks 0..nblk-1): orders
// the blocks in both directions and builds both block schedules.
// Contiguous row blocks keep ascending (forward) and descending (backward)
// order, which k_sweep_blk relies on.
void install_partition(Ext* e, int nblk, bool contiguous) {
    const int n = e->n;
    hipStream_t s = e->stream;
    PhaseTimer pt("partition", s);
    e->nblk = nblk;
    {
        DevBuf<int> cnt(std::max(nblk, 1));
        PILU_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(int) * nblk, s));
        k_block_counts<<<grid_for(n, 256), 256, 0, s>>>(n, e->block_of.get(), cnt.get());
        PILU_HIP(hipGetLastError());
        std::vector<int> hc(nblk);
        PILU_HIP(hipMemcpyAsync(hc.data(), cnt.get(), sizeof(int) * nblk, hipMemcpyDeviceToHost, s));
        PILU_HIP(hipStreamSynchronize(s));
        e->max_block_rows = *std::max_element(hc.begin(), hc.end());
    }
    if (e->max_block_rows > kMaxBlockRows)
        throw std::runtime_error("sweep partition: a block holds more than 16384 rows");
    std::vector<int> pos_l(nblk), pos_u(nblk);
    if (contiguous) {
        for (int b = 0; b < nblk; ++b) {
            pos_l[b] = b;
            pos_u[b] = nblk - 1 - b;
        }
    } else {
        pt.mark("block sizes");
        pos_l = block_order(e, false);
        pt.mark("L block order");
        pos_u = block_order(e, true);
        pt.mark("U block order");
    }
    build_block_schedule(e, false, e->block_of.get(), e->nblk, pos_l, e->bperm_l, e->blev_l, e->loff_l);
    pt.mark("L block schedule");
    build_block_schedule(e, true, e->block_of.get(), e->nblk, pos_u, e->bperm_u, e->blev_u, e->loff_u);
    pt.mark("U block schedule");
    ++e->part_version;  // packs are rebuilt on first use
}

// Contiguous blocks of blk rows.
void use_blocks(Ext* e, int blk) {
    if (blk < 64 || blk > kMaxBlockRows) throw std::runtime_error("block rows must be in [64, 16384]");
    if (blk == e->blk_rows && e->tile[0] == 0) return;
    e->block_of.reserve(e->n);
    k_block_contiguous<<<grid_for(e->n, 256), 256, 0, e->stream>>>(e->n, blk, e->block_of.get());
    PILU_HIP(hipGetLastError());
    e->blk_rows = blk;
    e->tile[0] = e->tile[1] = e->tile[2] = 0;
    install_partition(e, ceil_div(e->n, blk), true);
}

// Skew factors of the tiled partition from the L entries' grid offsets
// (dx, dy, dz) (a row at x + nx (y + ny z)): pass 0 gives b from the
// same-plane entries with dy < 0 (ceil(dx / -dy)) and c2 from the entries
// with dz < 0 (ceil(dy / -dz)); pass 1 gives c (ceil((dx + b dy) / -dz)).
// Maxima over all entries, at least 0, in out[0] (b, c) and out[1] (c2).
__device__ __forceinline__ long skew_ceil_div(long a, long b) { return a >= 0 ? (a + b - 1) / b : -((-a) / b); }

__global__ void k_skew(int n, long nx, long ny, const int* __restrict__ rp, const int* __restrict__ ci,
                       const int* __restrict__ dg, int pass, long b, unsigned long long* __restrict__ out) {
    __shared__ long red[2][256];
    long l0 = 0, l1 = 0;
    const long nxy = nx * ny;
    for (long r = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; r < n;
         r += static_cast<long>(gridDim.x) * blockDim.x) {
        const long x = r % nx, y = (r / nx) % ny, z = r / nxy;
        for (int p = rp[r]; p < dg[r]; ++p) {
            const long q = ci[p];
            const long dx = q % nx - x, dy = (q / nx) % ny - y, dz = q / nxy - z;
            if (pass == 0) {
                if (dz == 0 && dy < 0) l0 = max(l0, skew_ceil_div(dx, -dy));
                if (dz < 0) l1 = max(l1, skew_ceil_div(dy, -dz));
            } else if (dz < 0) {
                l0 = max(l0, skew_ceil_div(dx + b * dy, -dz));
            }
        }
    }
    red[0][threadIdx.x] = l0;
    red[1][threadIdx.x] = l1;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) {
            red[0][threadIdx.x] = max(red[0][threadIdx.x], red[0][threadIdx.x + h]);
            red[1][threadIdx.x] = max(red[1][threadIdx.x], red[1][threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMax(&out[0], static_cast<unsigned long long>(red[0][0]));
        atomicMax(&out[1], static_cast<unsigned long long>(red[1][0]));
    }
}

// Tiles of the grid declared by SetGridShape (row x + nx (y + ny z)).
// Tile of every row (row r at x + nx (y + ny z)) and, by atomicMin, the
// first row of every tile.
__global__ void k_tile_of(int n, long nx, long ny, long b, long c, long c2, int tx, int ty, int tz, long ntu,
                          long ntv, int* __restrict__ tile_of, int* __restrict__ first) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const long x = r % nx, yz = r / nx, y = yz % ny, z = yz / ny;
    const long u = x + b * y + c * z, v = y + c2 * z;
    const int t = static_cast<int>(((z / tz) * ntv + v / ty) * ntu + u / tx);
    tile_of[r] = t;
    // a wave spans at most two tiles mostly: one atomic for the tile of its
    // first row (lane 0's row, the smallest), one for the tile of its last
    // row (its first lane there), any other lane on its own
    const int lane = threadIdx.x & 63;
    const int last = 63 - __clzll(__ballot(true));
    const int t0 = __shfl(t, 0, 64), t1 = __shfl(t, last, 64);
    const unsigned long long m1 = __ballot(t == t1 && t != t0);
    if (lane == 0) atomicMin(first + t0, r);
    if (m1 && lane == __ffsll(static_cast<long long>(m1)) - 1) atomicMin(first + t1, r);
    if (t != t0 && t != t1) atomicMin(first + t, r);
}

__global__ void k_apply_tile_id(int n, const int* __restrict__ tile_of, const int* __restrict__ id,
                                int* __restrict__ block_of) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) block_of[r] = id[tile_of[r]];
}

// Tiles are boxes of tx x ty x tz in skewed coordinates
//   u = x + b y + c z,  v = y + c2 z,  w = z,
// with the smallest non-negative integers b, c, c2 for which no forward
// dependency (an entry of L, offset (dx, dy, dz) from its row) points to a
// larger u, v or w: then the tile index never increases along a dependency
// and the tiles form an acyclic graph.  A 7-point ILU(0) pattern gives
// b = c = c2 = 0 (plain boxes); ILU(1), whose fill couples (x+1, y-1),
// (x+1, z-1) and (y+1, z-1), gives b = c = c2 = 1.  The skew is unimodular,
// so an interior tile holds exactly tx ty tz rows.
void use_tiles(Ext* e, int tx, int ty, int tz) {
    const int nx = e->grid[0], ny = e->grid[1], nz = e->grid[2];
    if (nx <= 0) throw std::runtime_error("tiled sweeps need the grid shape (SetGridShape)");
    if (tx < 1 || ty < 1 || tz < 1 || static_cast<long>(tx) * ty * tz > kMaxBlockRows)
        throw std::runtime_error("tile must hold 1..16384 rows");
    if (e->tile[0] == tx && e->tile[1] == ty && e->tile[2] == tz) return;
    const int n = e->n;
    PhaseTimer pt("tiles", e->stream);
    hipStream_t st = e->stream;
    long b = 0, c = 0, c2 = 0;
    {
        DevBuf<unsigned long long> mx(2);
        const int grid = std::min(grid_for(n, 256), 4096);
        for (int pass = 0; pass < 2; ++pass) {  // b first, then c (which depends on b)
            PILU_HIP(hipMemsetAsync(mx.get(), 0, 2 * sizeof(unsigned long long), st));
            k_skew<<<grid, 256, 0, st>>>(n, nx, ny, e->rp.get(), e->ci.get(), e->dg.get(), pass, b, mx.get());
            PILU_HIP(hipGetLastError());
            unsigned long long r[2] = {0, 0};
            PILU_HIP(hipMemcpyAsync(r, mx.get(), sizeof(r), hipMemcpyDeviceToHost, st));
            PILU_HIP(hipStreamSynchronize(st));
            if (pass == 0) {
                b = static_cast<long>(r[0]);
                c2 = static_cast<long>(r[1]);
            } else {
                c = static_cast<long>(r[0]);
            }
        }
    }
    pt.mark("skew");
    const long umax = (nx - 1) + b * (ny - 1) + c * (nz - 1), vmax = (ny - 1) + c2 * (nz - 1);
    const long ntu = umax / tx + 1, ntv = vmax / ty + 1, ntw = (nz - 1) / tz + 1;
    if (ntu * ntv * ntw > (1L << 30)) throw std::runtime_error("tiled sweeps: too many tiles");
    const int ntiles = static_cast<int>(ntu * ntv * ntw);
    // tile of every row and the first row of every tile on the device; the
    // non-empty tiles get dense ids in order of their first row
    DevBuf<int> tile_of(std::max(n, 1)), first(ntiles), id(ntiles);
    k_fill_int<<<grid_for(ntiles, 256 * 8), 256, 0, st>>>(ntiles, first.get(), INT_MAX);
    k_tile_of<<<grid_for(n, 256), 256, 0, st>>>(n, nx, ny, b, c, c2, tx, ty, tz, ntu, ntv, tile_of.get(), first.get());
    PILU_HIP(hipGetLastError());
    std::vector<int> hf(ntiles), hid(ntiles, -1);
    PILU_HIP(hipMemcpyAsync(hf.data(), first.get(), sizeof(int) * ntiles, hipMemcpyDeviceToHost, st));
    PILU_HIP(hipStreamSynchronize(st));
    std::vector<std::pair<int, int>> order;
    for (int t = 0; t < ntiles; ++t)
        if (hf[t] != INT_MAX) order.emplace_back(hf[t], t);
    std::sort(order.begin(), order.end());
    const int nblk = static_cast<int>(order.size());
    for (int k = 0; k < nblk; ++k) hid[order[k].second] = k;
    PILU_HIP(hipMemcpyAsync(id.get(), hid.data(), sizeof(int) * ntiles, hipMemcpyHostToDevice, st));
    e->block_of.reserve(n);
    k_apply_tile_id<<<grid_for(n, 256), 256, 0, st>>>(n, tile_of.get(), id.get(), e->block_of.get());
    PILU_HIP(hipGetLastError());
    PILU_HIP(hipStreamSynchronize(st));  // hid goes out of scope
    e->blk_rows = 0;
    e->tile[0] = tx;
    e->tile[1] = ty;
    e->tile[2] = tz;
    e->skew[0] = static_cast<int>(b);
    e->skew[1] = static_cast<int>(c);
    e->skew[2] = static_cast<int>(c2);
    pt.mark("tile of every row");
    install_partition(e, nblk, false);
}

// Longest row, L part and U part: ext[0..2] (atomic maxima).  A grid of at
// most 1024 workgroups strides over the rows; each reduces in registers and
// LDS and adds one atomic per extent.
__global__ __launch_bounds__(256) void k_row_extents(int n, const int* __restrict__ rp, const int* __restrict__ dg,
                                                     int* ext) {
    __shared__ int s_m[3][4];
    int r = 0, lo = 0, up = 0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += 256L * gridDim.x) {
        r = max(r, rp[i + 1] - rp[i]);
        lo = max(lo, dg[i] - rp[i]);
        up = max(up, rp[i + 1] - dg[i] - 1);
    }
    for (int o = 32; o > 0; o >>= 1) {
        r = max(r, __shfl_xor(r, o, 64));
        lo = max(lo, __shfl_xor(lo, o, 64));
        up = max(up, __shfl_xor(up, o, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_m[0][w] = r;
        s_m[1][w] = lo;
        s_m[2][w] = up;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int* m = s_m[threadIdx.x];
        atomicMax(ext + threadIdx.x, max(max(m[0], m[1]), max(m[2], m[3])));
    }
}

// Device pattern is in e->rp / e->ci: validate, locate diagonals, schedule.
void analyse(Ext* e) {
    const int n = e->n;
    PhaseTimer pt("analyse", e->stream);
    e->analysed = false;  // until every schedule below is built
    e->factored = false;
    hipStream_t s = e->stream;
    e->dg.reserve(n);
    e->level.reserve(n);
    e->ctr.reserve(16);
    e->err.reserve(16);
    PILU_HIP(hipMemsetAsync(e->ctr.get(), 0, 16 * sizeof(int), s));
    PILU_HIP(hipMemsetAsync(e->err.get(), 0, 16 * sizeof(int), s));
    k_check_rowptr<<<grid_for(static_cast<long>(n) + 1, 256), 256, 0, s>>>(n, e->nnz, e->rp.get(),
                                                                        e->err.get() + kErrRowPtr);
    PILU_HIP(hipGetLastError());
    if (read_int(e->err.get() + kErrRowPtr, s))
        throw std::runtime_error("row_ptr must start at 0, be non-decreasing and end at nnz");
    k_check_sorted<<<grid_for(n, 256), 256, 0, s>>>(n, e->rp.get(), e->ci.get(), e->err.get() + kErrSorted);
    k_diag_ptr<<<grid_for(n, 256), 256, 0, s>>>(n, e->rp.get(), e->ci.get(), e->dg.get(), e->err.get() + kErrDiag);
    PILU_HIP(hipGetLastError());
    int flags[2] = {0, 0};
    PILU_HIP(hipMemcpyAsync(flags, e->err.get() + kErrSorted, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    PILU_HIP(hipStreamSynchronize(s));
    if (flags[0]) throw std::runtime_error("rows must have strictly increasing, non-negative column indices");
    if (flags[1]) throw std::runtime_error("missing diag");

    // longest row, longest L part and longest U part (device reduction)
    PILU_HIP(hipMemsetAsync(e->err.get() + kErrExtent, 0, 3 * sizeof(int), s));
    k_row_extents<<<std::min(grid_for(n, 256), 1024), 256, 0, s>>>(n, e->rp.get(), e->dg.get(),
                                                                 e->err.get() + kErrExtent);
    PILU_HIP(hipGetLastError());
    int ext[3] = {0, 0, 0};
    PILU_HIP(hipMemcpyAsync(ext, e->err.get() + kErrExtent, sizeof(ext), hipMemcpyDeviceToHost, s));
    PILU_HIP(hipStreamSynchronize(s));
    e->max_row = ext[0];
    e->max_lo = ext[1];
    e->max_up = ext[2];
    ++e->pattern_version;  // host copies of the pattern are stale from here on
    pt.mark("pattern checks, diagonal, row extents");

    (void)device_scratch();  // the no-allocation apply's scratch on this device
    build_schedules(e);
    pt.mark("L and U level schedules");

    e->vin.reserve(e->nnz);
    e->fac.reserve(e->nnz);
    e->y.reserve(n);
    e->y_armed = false;
    e->factored = false;
    e->mode_l = e->mode_u = 0;
    e->blk_rows = e->nblk = 0;
    e->tile[0] = e->tile[1] = e->tile[2] = 0;
    e->grid[0] = e->grid[1] = e->grid[2] = 0;
    ++e->part_version;
    e->analysed = true;
}

// Packed schedule-ordered factors of one direction for the current
// partition: rowinfo (row, entry count) per schedule position, the slabs
// of every (block, level) group, in-block columns encoded as
// -(1 + schedule position within the block), and kptr (first schedule
// position of every processed block).
// Device side of the pack build: inverse schedule, first position of each
// row's block, group of each position, and the slab fill.
__global__ void k_pack_inv(int n, const int* __restrict__ perm, int* __restrict__ kinv) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) kinv[perm[k]] = k;
}

__global__ void k_pack_block_start(int nblk, const int* __restrict__ kptr, const int* __restrict__ perm,
                                   int* __restrict__ kb0) {
    const int p = blockIdx.x;
    if (p >= nblk) return;
    for (int k = kptr[p] + threadIdx.x; k < kptr[p + 1]; k += blockDim.x) kb0[perm[k]] = kptr[p];
}

// (row, entry count in the direction) of every schedule position.
template <bool UPPER>
__global__ void k_pack_rowinfo(int n, const int* __restrict__ perm, const int* __restrict__ rp,
                               const int* __restrict__ dg, int2* __restrict__ rowinfo) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int row = perm[k];
    rowinfo[k] = make_int2(row, UPPER ? rp[row + 1] - dg[row] - 1 : dg[row] - rp[row]);
}

// Widest row of every level group.
__global__ void k_pack_width(int ngroups, const int* __restrict__ loff, const int2* __restrict__ rowinfo,
                             int* __restrict__ width) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    int w = 0;
    for (int k = loff[g]; k < loff[g + 1]; ++k) w = max(w, rowinfo[k].y);
    width[g] = w;
}

__global__ void k_pack_group_of(int ngroups, const int* __restrict__ loff, int* __restrict__ grp_of) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ngroups)
        for (int k = loff[g]; k < loff[g + 1]; ++k) grp_of[k] = g;
}

// Slab of every schedule position k: entry q of the row sits at
// base[g] + q * count + (k - loff[g]); an in-block column is stored as
// -(1 + its position within the block), padding as column 0, source -1.
template <bool UPPER>
__global__ void k_pack_fill(int n, const int* __restrict__ perm, const int* __restrict__ grp_of,
                            const int* __restrict__ loff, const int* __restrict__ base,
                            const int* __restrict__ width, const int* __restrict__ rp, const int* __restrict__ dg,
                            const int* __restrict__ ci, const int* __restrict__ block_of,
                            const int* __restrict__ kinv, const int* __restrict__ kb0, int* __restrict__ pcol,
                            int* __restrict__ psrc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int g = grp_of[k], lo = loff[g], count = loff[g + 1] - lo, w = width[g];
    const int row = perm[k];
    const int e0 = UPPER ? dg[row] + 1 : rp[row];
    const int cnt = UPPER ? rp[row + 1] - dg[row] - 1 : dg[row] - rp[row];
    const int brow = block_of[row], k0 = kb0[row];
    for (int q = 0; q < w; ++q) {
        const long pos = base[g] + static_cast<long>(q) * count + (k - lo);
        if (q < cnt) {
            const int c = ci[e0 + q];
            pcol[pos] = block_of[c] == brow ? -1 - (kinv[c] - k0) : c;
            psrc[pos] = e0 + q;
        } else {
            pcol[pos] = 0;
            psrc[pos] = -1;
        }
    }
}

void build_pack(Ext* e, bool upper, SweepPack& pk) {
    refuse_if_no_alloc(e, "packed factors");
    const int n = e->n, nblk = e->nblk;
    hipStream_t st = e->stream;
    const DevBuf<int>& perm = upper ? e->bperm_u : e->bperm_l;
    const DevBuf<int>& blev = upper ? e->blev_u : e->blev_l;
    const DevBuf<int>& loff = upper ? e->loff_u : e->loff_l;
    pk.h_blev.assign(static_cast<size_t>(nblk) + 1, 0);
    PILU_HIP(hipMemcpy(pk.h_blev.data(), blev.get(), sizeof(int) * pk.h_blev.size(), hipMemcpyDeviceToHost));
    const int ngroups = pk.h_blev[nblk];
    pk.h_loff.assign(static_cast<size_t>(ngroups) + 1, 0);
    PILU_HIP(hipMemcpy(pk.h_loff.data(), loff.get(), sizeof(int) * pk.h_loff.size(), hipMemcpyDeviceToHost));
    std::vector<int> kptr(static_cast<size_t>(nblk) + 1);
    for (int p = 0; p <= nblk; ++p) kptr[p] = pk.h_loff[pk.h_blev[p]];
    // rows and widths on the device; the slab bases on the host (one pass
    // over the groups, kept there for the pass lists)
    pk.rowinfo.resize(2 * static_cast<size_t>(std::max(n, 1)));
    DevBuf<int> dwidth(std::max(ngroups, 1));
    int2* rowinfo = reinterpret_cast<int2*>(pk.rowinfo.get());
    if (n > 0) {
        if (upper) k_pack_rowinfo<true><<<grid_for(n, 256), 256, 0, st>>>(n, perm.get(), e->rp.get(), e->dg.get(), rowinfo);
        else k_pack_rowinfo<false><<<grid_for(n, 256), 256, 0, st>>>(n, perm.get(), e->rp.get(), e->dg.get(), rowinfo);
    }
    if (ngroups > 0) k_pack_width<<<grid_for(ngroups, 256), 256, 0, st>>>(ngroups, loff.get(), rowinfo, dwidth.get());
    PILU_HIP(hipGetLastError());
    pk.h_width.assign(ngroups, 0);
    if (ngroups > 0)
        PILU_HIP(hipMemcpy(pk.h_width.data(), dwidth.get(), sizeof(int) * ngroups, hipMemcpyDeviceToHost));
    pk.h_base.assign(ngroups, 0);
    long total = 0;
    int maxw = 0;
    for (int g = 0; g < ngroups; ++g) {
        pk.h_base[g] = static_cast<int>(total);
        maxw = std::max(maxw, pk.h_width[g]);
        total += static_cast<long>(pk.h_width[g]) * (pk.h_loff[g + 1] - pk.h_loff[g]);
        if (total >= (1L << 31)) throw std::runtime_error("packed sweep: too many entries");
    }
    pk.ne = total;
    pk.max_w = maxw;
    const size_t ne_alloc = static_cast<size_t>(std::max<long>(total, 1));
    pk.pcol.resize(ne_alloc);
    pk.psrc.resize(ne_alloc);
    pk.kptr.resize(kptr.size());
    PILU_HIP(hipMemcpy(pk.kptr.get(), kptr.data(), sizeof(int) * kptr.size(), hipMemcpyHostToDevice));
    // slabs on the device (every slot is written, padding included)
    {
        DevBuf<int> kinv(std::max(n, 1)), kb0(std::max(n, 1)), grp_of(std::max(n, 1));
        DevBuf<int> dbase(std::max(ngroups, 1));
        PILU_HIP(hipMemcpy(dbase.get(), pk.h_base.data(), sizeof(int) * ngroups, hipMemcpyHostToDevice));
        if (total == 0) {
            const int zero = 0, minus1 = -1;
            PILU_HIP(hipMemcpy(pk.pcol.get(), &zero, sizeof(int), hipMemcpyHostToDevice));
            PILU_HIP(hipMemcpy(pk.psrc.get(), &minus1, sizeof(int), hipMemcpyHostToDevice));
        }
        if (n > 0) {
            k_pack_inv<<<grid_for(n, 256), 256, 0, st>>>(n, perm.get(), kinv.get());
            if (nblk > 0) k_pack_block_start<<<nblk, 256, 0, st>>>(nblk, pk.kptr.get(), perm.get(), kb0.get());
            if (ngroups > 0)
                k_pack_group_of<<<grid_for(ngroups, 256), 256, 0, st>>>(ngroups, loff.get(), grp_of.get());
            if (upper)
                k_pack_fill<true><<<grid_for(n, 256), 256, 0, st>>>(
                    n, perm.get(), grp_of.get(), loff.get(), dbase.get(), dwidth.get(), e->rp.get(), e->dg.get(),
                    e->ci.get(), e->block_of.get(), kinv.get(), kb0.get(), pk.pcol.get(), pk.psrc.get());
            else
                k_pack_fill<false><<<grid_for(n, 256), 256, 0, st>>>(
                    n, perm.get(), grp_of.get(), loff.get(), dbase.get(), dwidth.get(), e->rp.get(), e->dg.get(),
                    e->ci.get(), e->block_of.get(), kinv.get(), kb0.get(), pk.pcol.get(), pk.psrc.get());
            PILU_HIP(hipGetLastError());
        }
        PILU_HIP(hipStreamSynchronize(st));
    }
    pk.version = e->part_version;
    pk.pass_g = 0;
    pk.factors = -1;
}

// Pass list of a pack for G lanes per row: every level group in chunks of
// kPpThreads / G rows, each block's list padded to a multiple of kPpUnroll.
void build_passes(Ext* e, SweepPack& pk, int G) {
    if (pk.pass_g == G) return;
    refuse_if_no_alloc(e, "pass list");
    const int per = kPpThreads / G;
    std::vector<PassDesc> desc;
    std::vector<int> dptr(static_cast<size_t>(e->nblk) + 1, 0);
    pk.max_np = 0;
    for (int p = 0; p < e->nblk; ++p) {
        dptr[p] = static_cast<int>(desc.size());
        int k_any = -1;
        for (int g = pk.h_blev[p]; g < pk.h_blev[p + 1]; ++g) {
            const int lo = pk.h_loff[g], hi = pk.h_loff[g + 1], count = hi - lo, width = pk.h_width[g];
            if (count >= (1 << 16) || width >= (1 << 15))
                throw std::runtime_error("packed sweep: level group too large for the pass descriptor");
            for (int c = lo; c < hi; c += per) {
                const int nact = std::min(per, hi - c);
                // every slab and schedule position the pass can touch lies in the pack
                if (c + nact > e->n ||
                    (width > 0 && pk.h_base[g] + static_cast<long>(width - 1) * count + (c - lo) + nact > pk.ne))
                    throw std::runtime_error("packed sweep: pass descriptor out of range");
                PassDesc d;
                d.k_lo = c;
                d.base = pk.h_base[g];
                d.cb = count | (c + per >= hi ? static_cast<int>(0x80000000u) : 0);
                d.jw = (c - lo) | (width << 16);
                desc.push_back(d);
                k_any = lo;
            }
        }
        if (k_any < 0) k_any = 0;
        while ((desc.size() - dptr[p]) % kPpUnroll) desc.push_back(PassDesc{k_any, 0, 0, 0});
        pk.max_np = std::max(pk.max_np, static_cast<int>(desc.size()) - dptr[p]);
    }
    dptr[e->nblk] = static_cast<int>(desc.size());
    if (desc.empty()) desc.push_back(PassDesc{0, 0, 0, 0});
    pk.pdesc.resize(sizeof(PassDesc) * desc.size());
    pk.desc_ptr.resize(dptr.size());
    PILU_HIP(hipMemcpy(pk.pdesc.get(), desc.data(), sizeof(PassDesc) * desc.size(), hipMemcpyHostToDevice));
    PILU_HIP(hipMemcpy(pk.desc_ptr.get(), dptr.data(), sizeof(int) * dptr.size(), hipMemcpyHostToDevice));
    pk.pass_g = G;
}

// Refresh the packed values from the current factors.
// The value arrays always get at least one element and the diagonal is
// always gathered: the packed sweep reads pval[0] from idle lanes and needs
// the U diagonal even when a direction has no off-diagonal entry at all.
void gather_pack(Ext* e, SweepPack& pk, bool upper) {
    if (pk.version != e->part_version) return;
    refuse_if_no_alloc(e, "packed factor values");
    hipStream_t s = e->stream;
    const int grid = 2048;
    if (e->precision == 32) {
        pk.pval32.reserve(std::max<long>(pk.ne, 1));
        if (upper) pk.pdiag32.reserve(e->n);
        k_pack_gather<float, float><<<grid, 256, 0, s>>>(pk.ne, pk.psrc.get(), e->fac32.get(), pk.pval32.get(), e->n,
                                                         reinterpret_cast<const int2*>(pk.rowinfo.get()), e->dg.get(),
                                                         upper ? pk.pdiag32.get() : nullptr);
    } else {
        pk.pval64.reserve(std::max<long>(pk.ne, 1));
        if (upper) pk.pdiag64.reserve(e->n);
        k_pack_gather<double, double><<<grid, 256, 0, s>>>(pk.ne, pk.psrc.get(), e->fac.get(), pk.pval64.get(), e->n,
                                                           reinterpret_cast<const int2*>(pk.rowinfo.get()),
                                                           e->dg.get(), upper ? pk.pdiag64.get() : nullptr);
    }
    PILU_HIP(hipGetLastError());
    pk.factors = e->factor_version;
}

void ensure_packs(Ext* e) {
    if (e->nblk <= 0) throw std::runtime_error("packed sweep needs a block partition");
    PhaseTimer pt("packs", e->stream);
    if (e->pack_l.version != e->part_version) {
        build_pack(e, false, e->pack_l);
        pt.mark("L pack");
    }
    if (e->pack_u.version != e->part_version) {
        build_pack(e, true, e->pack_u);
        pt.mark("U pack");
    }
    if (e->factored) {
        if (e->pack_l.factors != e->factor_version) gather_pack(e, e->pack_l, false);
        if (e->pack_u.factors != e->factor_version) gather_pack(e, e->pack_u, true);
    }
}

// ---------------------------------------------------------------------
// Record streams of the wave-specialised sweep (mode 5), both directions
// for the current partition and G lanes per row.  Per direction:
//   * the block schedule (rows by processing position of their block,
//     local level, row) comes from the device;
//   * a block's imports are its distinct out-of-block columns, ordered by
//     the first record that reads them; a row some block imports is a face
//     row, and the face rows of each block get consecutive face slots in
//     schedule order (so a block publishes and re-arms contiguous ranges);
//   * each local level is cut into records of at most 64 / G rows; columns
//     become LDS slots (in-block: schedule position within the block;
//     imports: max_rows + import index; padding: the zero slot); the first
//     S * G entries of each row are stored lane-major (S = 8 / G, at least
//     1), the rest row-major; values and diagonal are gathered into the
//     stream after every factorization through (vdst, vsrc).
// ---------------------------------------------------------------------
struct WsDir {
    std::vector<int> perm, blev, loff, kinv, blk_at, face_of;
    std::vector<std::vector<int2>> imp;  // per position: (column, first record)
    std::vector<int> face_lo;            // per position: first face slot (nblk + 1)
};

void ws_schedule(Ext* e, bool upper, WsDir& d, int P) {
    const int n = e->n, nblk = e->nblk;
    const DevBuf<int>& perm = upper ? e->bperm_u : e->bperm_l;
    const DevBuf<int>& blev = upper ? e->blev_u : e->blev_l;
    const DevBuf<int>& loff = upper ? e->loff_u : e->loff_l;
    d.perm.resize(n);
    d.blev.resize(static_cast<size_t>(nblk) + 1);
    PILU_HIP(hipMemcpy(d.perm.data(), perm.get(), sizeof(int) * n, hipMemcpyDeviceToHost));
    PILU_HIP(hipMemcpy(d.blev.data(), blev.get(), sizeof(int) * d.blev.size(), hipMemcpyDeviceToHost));
    d.loff.resize(static_cast<size_t>(d.blev[nblk]) + 1);
    PILU_HIP(hipMemcpy(d.loff.data(), loff.get(), sizeof(int) * d.loff.size(), hipMemcpyDeviceToHost));
    d.kinv.assign(n, 0);
    for (int k = 0; k < n; ++k) d.kinv[d.perm[k]] = k;
    const std::vector<int>& rp = e->h_rp;
    const std::vector<int>& dg = e->h_dg;
    const std::vector<int>& ci = e->h_ci;
    const std::vector<int>& bo = e->h_block_of;
    d.blk_at.assign(nblk, -1);
    d.imp.assign(nblk, {});
    parallel_for(nblk, [&](long lo, long hi) {
        for (long p = lo; p < hi; ++p) {
            const int g0 = d.blev[p], g1 = d.blev[p + 1];
            if (g0 == g1) continue;
            const int blk = bo[d.perm[d.loff[g0]]];
            d.blk_at[p] = blk;
            std::vector<int2> v;
            int rec = 0;
            for (int g = g0; g < g1; ++g)
                for (int c0 = d.loff[g]; c0 < d.loff[g + 1]; c0 += P, ++rec)
                    for (int k = c0; k < std::min(c0 + P, d.loff[g + 1]); ++k) {
                        const int row = d.perm[k];
                        const int e0 = upper ? dg[row] + 1 : rp[row], e1 = upper ? rp[row + 1] : dg[row];
                        for (int q = e0; q < e1; ++q)
                            if (bo[ci[q]] != blk) v.push_back(make_int2(ci[q], rec));
                    }
            std::sort(v.begin(), v.end(), [](int2 a, int2 b) { return a.x != b.x ? a.x < b.x : a.y < b.y; });
            size_t m = 0;
            for (size_t i = 0; i < v.size(); ++i)
                if (m == 0 || v[i].x != v[m - 1].x) v[m++] = v[i];
            v.resize(m);
            std::sort(v.begin(), v.end(), [](int2 a, int2 b) { return a.y != b.y ? a.y < b.y : a.x < b.x; });
            d.imp[p] = std::move(v);
        }
    });
    std::vector<char> is_face(n, 0);
    for (const auto& v : d.imp)
        for (const int2& x : v) is_face[x.x] = 1;
    d.face_of.assign(n, -1);
    d.face_lo.assign(static_cast<size_t>(nblk) + 1, 0);
    int nf = 0;
    for (int p = 0; p < nblk; ++p) {
        d.face_lo[p] = nf;
        const int k0 = d.loff[d.blev[p]], k1 = d.loff[d.blev[p + 1]];
        for (int k = k0; k < k1; ++k)
            if (is_face[d.perm[k]]) d.face_of[d.perm[k]] = nf++;
    }
    d.face_lo[nblk] = nf;
}

inline int a16(int b) { return (b + 15) & ~15; }

// Lays out and uploads one direction's stream; other_face[2 b], [2 b + 1]:
// the other direction's face slots of block b, which this direction
// re-arms.
// Bytes of one record of cnt rows, w entries per row.
int ws_record_bytes(int cnt, int w, int G, int sv, bool upper) {
    const int S = std::max(1, 8 / G), L = cnt * G, wt = std::max(w - S * G, 0);
    return 32 + a16(cnt * 8) + a16(L * S * 4) + a16(L * S * sv) + (wt > 0 ? a16(cnt * wt * 4) + a16(cnt * wt * sv) : 0) +
           (upper ? a16(cnt * sv) : 0);
}

void ws_stream(Ext* e, bool upper, const WsDir& d, const std::vector<int>& other_face, int G, WsPack& wp) {
    const int n = e->n, nblk = e->nblk, P = 64 / G;
    const int S = std::max(1, 8 / G);
    const int sv = e->precision == 32 ? 4 : 8;
    const std::vector<int>& rp = e->h_rp;
    const std::vector<int>& dg = e->h_dg;
    const std::vector<int>& ci = e->h_ci;
    const std::vector<int>& bo = e->h_block_of;
    const int max_rows = e->max_block_rows;
    int max_imp = 0;
    for (const auto& v : d.imp) max_imp = std::max(max_imp, static_cast<int>(v.size()));
    const int zero = max_rows + max_imp;
    auto entries = [&](int row) { return upper ? rp[row + 1] - dg[row] - 1 : dg[row] - rp[row]; };
    // pass 1: bytes and value slots per position
    std::vector<long> bytes(nblk, 0), nval(nblk, 0);
    std::vector<int> maxrec(nblk, 0);
    parallel_for(nblk, [&](long lo, long hi) {
        for (long p = lo; p < hi; ++p) {
            long off = 0, nv = 0;
            int big = 0;
            for (int g = d.blev[p]; g < d.blev[p + 1]; ++g)
                for (int c0 = d.loff[g]; c0 < d.loff[g + 1]; c0 += P) {
                    const int cnt = std::min(P, d.loff[g + 1] - c0);
                    int w = 0;
                    for (int k = c0; k < c0 + cnt; ++k) w = std::max(w, entries(d.perm[k]));
                    w = (w + G - 1) / G * G;
                    const int rb = ws_record_bytes(cnt, w, G, sv, upper);
                    big = std::max(big, rb);
                    if ((off % kWsRing) + rb > kWsRing) off += kWsRing - off % kWsRing;
                    off += rb;
                    nv += static_cast<long>(cnt) * (G * S + std::max(w - S * G, 0)) + (upper ? cnt : 0);
                }
            bytes[p] = (off + kWsChunk - 1) / kWsChunk * kWsChunk;
            nval[p] = nv;
            maxrec[p] = big;
        }
    });
    if (*std::max_element(maxrec.begin(), maxrec.end()) > kWsRing / 2)
        throw std::runtime_error("wave-specialised sweep: a record exceeds half the LDS ring (rows too long)");
    std::vector<long> s0(static_cast<size_t>(nblk) + 1, 0), v0(static_cast<size_t>(nblk) + 1, 0),
        i0(static_cast<size_t>(nblk) + 1, 0);
    for (int p = 0; p < nblk; ++p) {
        s0[p + 1] = s0[p] + bytes[p] / 16;
        v0[p + 1] = v0[p] + nval[p];
        i0[p + 1] = i0[p] + static_cast<long>(d.imp[p].size());
    }
    if (s0[nblk] >= (1L << 31) / static_cast<long>(16 / sv) || v0[nblk] >= (1L << 31) || i0[nblk] >= (1L << 31))
        throw std::runtime_error("wave-specialised sweep: record stream too large");
    std::vector<uint32_t> words(static_cast<size_t>(std::max<long>(s0[nblk], 1)) * 4, 0);
    std::vector<int> vdst(static_cast<size_t>(std::max<long>(v0[nblk], 1))), vsrc(vdst.size(), -1);
    std::vector<int2> imports(static_cast<size_t>(std::max<long>(i0[nblk], 1)));
    std::vector<WsBlock> blocks(nblk);
    std::atomic<bool> count_mismatch{false};
    // pass 2: fill
    parallel_for(nblk, [&](long lo, long hi) {
        std::vector<std::pair<int, int>> imap;
        for (long p = lo; p < hi; ++p) {
            const int blk = d.blk_at[p];
            const int k0 = d.blev[p] < d.blev[p + 1] ? d.loff[d.blev[p]] : 0;
            const int k1 = d.blev[p] < d.blev[p + 1] ? d.loff[d.blev[p + 1]] : 0;
            WsBlock& bk = blocks[p];
            bk.s0 = static_cast<int>(s0[p]);
            bk.s1 = static_cast<int>(s0[p + 1]);
            bk.i0 = static_cast<int>(i0[p]);
            bk.i1 = static_cast<int>(i0[p + 1]);
            bk.r0 = blk >= 0 ? other_face[2 * blk] : 0;
            bk.r1 = blk >= 0 ? other_face[2 * blk + 1] : 0;
            bk.k0 = k0;
            bk.nrows = k1 - k0;
            imap.clear();
            for (size_t i = 0; i < d.imp[p].size(); ++i) {
                const int2 x = d.imp[p][i];
                imports[i0[p] + i] = make_int2(d.face_of[x.x], x.y);
                imap.emplace_back(x.x, static_cast<int>(i));
            }
            std::sort(imap.begin(), imap.end());
            auto import_slot = [&](int c) {
                auto it = std::lower_bound(imap.begin(), imap.end(), std::make_pair(c, -1));
                return max_rows + it->second;
            };
            char* base = reinterpret_cast<char*>(words.data()) + s0[p] * 16;
            const long sbyte0 = s0[p] * 16;
            // the block's records: (first position, rows, width, bytes, offset)
            struct Rec { int c0, cnt, w, rb; long off; };
            std::vector<Rec> recs;
            long off = 0;
            for (int g = d.blev[p]; g < d.blev[p + 1]; ++g)
                for (int c0 = d.loff[g]; c0 < d.loff[g + 1]; c0 += P) {
                    const int cnt = std::min(P, d.loff[g + 1] - c0);
                    int w = 0;
                    for (int k = c0; k < c0 + cnt; ++k) w = std::max(w, entries(d.perm[k]));
                    w = (w + G - 1) / G * G;
                    const int rb = ws_record_bytes(cnt, w, G, sv, upper);
                    if ((off % kWsRing) + rb > kWsRing) {
                        const int pad = static_cast<int>(kWsRing - off % kWsRing);
                        int* h = reinterpret_cast<int*>(base + off);
                        h[0] = 0; h[1] = 0; h[2] = pad; h[3] = 0;
                        off += pad;
                    }
                    recs.push_back({c0, cnt, w, rb, off});
                    off += rb;
                }
            long vi = v0[p];
            for (size_t r = 0; r < recs.size(); ++r) {
                const Rec& rc = recs[r];
                const int cnt = rc.cnt, w = rc.w, wt = std::max(w - S * G, 0), L = cnt * G;
                const int rows_b = 32, lcols_b = rows_b + a16(cnt * 8), lvals_b = lcols_b + a16(L * S * 4);
                const int tcols_b = lvals_b + a16(L * S * sv), tvals_b = tcols_b + (wt > 0 ? a16(cnt * wt * 4) : 0);
                const int diag_b = tvals_b + (wt > 0 ? a16(cnt * wt * sv) : 0);
                int* h = reinterpret_cast<int*>(base + rc.off);
                h[0] = cnt;
                h[1] = w;
                h[2] = rc.rb;
                h[3] = rc.c0 - k0;
                if (r + 1 < recs.size()) {
                    const Rec& nx = recs[r + 1];
                    h[4] = nx.cnt;
                    h[5] = nx.w;
                    h[6] = nx.rb;
                    h[7] = nx.c0 - k0;
                }
                int2* rows = reinterpret_cast<int2*>(base + rc.off + rows_b);
                int* lcols = reinterpret_cast<int*>(base + rc.off + lcols_b);
                int* tcols = reinterpret_cast<int*>(base + rc.off + tcols_b);
                for (int j = 0; j < cnt; ++j) {
                    const int row = d.perm[rc.c0 + j];
                    rows[j] = make_int2(row, d.face_of[row]);
                    const int e0 = upper ? dg[row] + 1 : rp[row];
                    const int m = entries(row);
                    // entry q of row j: (LDS slot, factor index or -1)
                    auto entry = [&](int q, int& slot, int& src) {
                        slot = zero;
                        src = -1;
                        if (q < m) {
                            const int c = ci[e0 + q];
                            slot = bo[c] == blk ? d.kinv[c] - k0 : import_slot(c);
                            src = e0 + q;
                        }
                    };
                    for (int t = 0; t < G; ++t)
                        for (int u = 0; u < S; ++u) {
                            const int li = (j * G + t) * S + u;
                            int slot, src;
                            entry(t + u * G, slot, src);
                            lcols[li] = slot;
                            vdst[vi] = static_cast<int>((sbyte0 + rc.off + lvals_b + static_cast<long>(li) * sv) / sv);
                            vsrc[vi] = src;
                            ++vi;
                        }
                    for (int q = 0; q < wt; ++q) {
                        int slot, src;
                        entry(S * G + q, slot, src);
                        tcols[j * wt + q] = slot;
                        vdst[vi] = static_cast<int>((sbyte0 + rc.off + tvals_b + static_cast<long>(j * wt + q) * sv) / sv);
                        vsrc[vi] = src;
                        ++vi;
                    }
                }
                if (upper)
                    for (int j = 0; j < cnt; ++j) {
                        vdst[vi] = static_cast<int>((sbyte0 + rc.off + diag_b + static_cast<long>(j) * sv) / sv);
                        vsrc[vi] = dg[d.perm[rc.c0 + j]];
                        ++vi;
                    }
            }
            if (vi != v0[p + 1]) count_mismatch = true;
        }
    });
    if (count_mismatch) throw std::logic_error("wave-specialised sweep: value count mismatch");
    wp.stream.resize(words.size() / 4);
    PILU_HIP(hipMemcpy(wp.stream.get(), words.data(), words.size() * 4, hipMemcpyHostToDevice));
    wp.vdst.resize(vdst.size());
    wp.vsrc.resize(vsrc.size());
    PILU_HIP(hipMemcpy(wp.vdst.get(), vdst.data(), vdst.size() * 4, hipMemcpyHostToDevice));
    PILU_HIP(hipMemcpy(wp.vsrc.get(), vsrc.data(), vsrc.size() * 4, hipMemcpyHostToDevice));
    wp.imports.resize(imports.size());
    PILU_HIP(hipMemcpy(wp.imports.get(), imports.data(), imports.size() * sizeof(int2), hipMemcpyHostToDevice));
    wp.blocks.resize(sizeof(WsBlock) * static_cast<size_t>(std::max(nblk, 1)));
    PILU_HIP(hipMemcpy(wp.blocks.get(), blocks.data(), sizeof(WsBlock) * nblk, hipMemcpyHostToDevice));
    wp.nface = d.face_lo[nblk];
    wp.face.resize(static_cast<size_t>(std::max(wp.nface, 1)));
    wp.face_armed = false;
    wp.nvals = v0[nblk];
    wp.nstream = s0[nblk];
    wp.max_imp = max_imp;
    wp.value_bytes = sv;
    wp.g = G;
    wp.version = e->part_version;
    wp.factors = -1;
    (void)n;
}

// LDS of k_sweep_ws for a direction: block x, imports, zero slot, import
// list, ring.
size_t ws_lds_bytes(const Ext* e, const WsPack& wp) {
    return sizeof(double) * static_cast<size_t>((e->max_block_rows + wp.max_imp + 2) & ~1) +
           static_cast<size_t>(a16(wp.max_imp * 8)) + kWsRing;
}

void gather_ws(Ext* e, WsPack& wp) {
    if (wp.version != e->part_version || wp.nvals <= 0) return;
    refuse_if_no_alloc(e, "record stream values");
    hipStream_t s = e->stream;
    const int grid = 2048;
    if (wp.value_bytes == 4)
        k_ws_gather<float, float><<<grid, 256, 0, s>>>(wp.nvals, wp.vdst.get(), wp.vsrc.get(), e->fac32.get(),
                                                       reinterpret_cast<float*>(wp.stream.get()));
    else
        k_ws_gather<double, double><<<grid, 256, 0, s>>>(wp.nvals, wp.vdst.get(), wp.vsrc.get(), e->fac.get(),
                                                         reinterpret_cast<double*>(wp.stream.get()));
    PILU_HIP(hipGetLastError());
    wp.factors = e->factor_version;
}

// (first, end) face slot of every block id in one direction's slot order.
std::vector<int> ws_face_ranges(const Ext* e, const WsDir& d) {
    std::vector<int> out(2 * static_cast<size_t>(e->nblk), 0);
    for (int p = 0; p < e->nblk; ++p)
        if (d.blk_at[p] >= 0) {
            out[2 * d.blk_at[p]] = d.face_lo[p];
            out[2 * d.blk_at[p] + 1] = d.face_lo[p + 1];
        }
    return out;
}

// Record stream of one direction for G lanes per row.  Face slots do not
// depend on G, so the other direction's ranges (re-armed by this one) are
// computed once per partition.
void build_ws(Ext* e, bool upper, int G) {
    refuse_if_no_alloc(e, "record streams");
    if (e->nblk <= 0) throw std::runtime_error("wave-specialised sweep needs a block partition");
    if (G < 1 || G > 16 || (G & (G - 1))) throw std::runtime_error("wave-specialised sweep: 1..16 lanes per row");
    host_pattern(e);
    WsPack& wp = upper ? e->ws_u : e->ws_l;
    WsPack& other = upper ? e->ws_l : e->ws_u;
    WsDir d;
    ws_schedule(e, upper, d, 64 / G);
    wp.face_rng = ws_face_ranges(e, d);
    wp.rng_version = e->part_version;
    if (other.rng_version != e->part_version) {
        WsDir o;
        ws_schedule(e, !upper, o, 64);
        other.face_rng = ws_face_ranges(e, o);
        other.rng_version = e->part_version;
    }
    ws_stream(e, upper, d, other.face_rng, G, wp);
    if (e->factored) gather_ws(e, wp);
}

bool ws_current(const Ext* e, const WsPack& wp, int G) {
    return wp.version == e->part_version && wp.g == G && wp.value_bytes == (e->precision == 32 ? 4 : 8);
}

void ensure_ws(Ext* e, bool upper, int G) {
    WsPack& wp = upper ? e->ws_u : e->ws_l;
    if (!ws_current(e, wp, G)) build_ws(e, upper, G);
    if (e->factored && wp.factors != e->factor_version) gather_ws(e, wp);
}

bool ws_fits(Ext* e, int G) {
    try {
        ensure_ws(e, false, G);
        ensure_ws(e, true, G);
    } catch (const std::runtime_error&) {
        return false;
    }
    return ws_lds_bytes(e, e->ws_l) <= 150 * 1024 && ws_lds_bytes(e, e->ws_u) <= 150 * 1024;
}

template <int G, int S, typename T, bool UPPER>
void sweep_ws_one(Ext* e, const double* b, double* out, double* reset, const int* gate, int* ctr) {
    WsPack& wp = UPPER ? e->ws_u : e->ws_l;
    WsPack& other = UPPER ? e->ws_l : e->ws_u;
    // the other direction's face buffer is re-armed here when it sweeps in
    // this mode over the same partition
    const bool other_ws = (UPPER ? e->mode_l : e->mode_u) == 5 && other.version == e->part_version &&
                          other.face.get() != nullptr;
    const size_t lds = ws_lds_bytes(e, wp);
    if (lds > 150 * 1024) throw std::runtime_error("wave-specialised sweep: block and imports exceed the LDS");
    if (!wp.stream.get() || !wp.blocks.get() || !wp.imports.get() || !wp.face.get() || wp.factors != e->factor_version)
        throw std::runtime_error("wave-specialised sweep: record streams missing (factorize before sweeping)");
    hipStream_t s = e->stream;
    if (!wp.face_armed)
        k_fill_u64<<<grid_for(std::max(wp.nface, 1), 256 * 8), 256, 0, s>>>(std::max(wp.nface, 1),
                                                                            reinterpret_cast<uint64_t*>(wp.face.get()),
                                                                            kPending);
    auto kern = k_sweep_ws<G, S, T, UPPER>;
    PILU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(lds)));
    kern<<<e->nblk, kWsThreads, lds, s>>>(reinterpret_cast<const WsBlock*>(wp.blocks.get()), wp.stream.get(),
                                          wp.imports.get(), UPPER ? e->bperm_u.get() : e->bperm_l.get(), b, out, reset,
                                          wp.face.get(), other_ws ? other.face.get() : nullptr, gate, ctr,
                                          e->err.get() + kErrSweep, watchdog_ticks(e->watchdog_s), e->max_block_rows,
                                          wp.max_imp, e->n, wp.nface);
    wp.face_armed = false;
    if (other_ws) other.face_armed = true;
}

int effective_lanes(const Ext* e, bool upper);

// Lanes per row of the wave-specialised sweep: the configured count, at
// most 16 (records hold 64 / G rows).
int ws_lanes(int lanes) { return std::min(std::max(lanes, 1), 16); }

// Lanes per row the packed sweeps run with for a configured count
// (sweep_pp_dispatch: 1, 2, 4, 8, else 16).
int pp_lanes(int lanes) { return lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 ? lanes : 16; }

// Builds what the configured sweep kinds need (packs, record streams), so
// that the first apply allocates nothing more.
void prepare_modes(Ext* e) {
    const bool pp_l = e->mode_l >= 2 && e->mode_l <= 3, pp_u = e->mode_u >= 2 && e->mode_u <= 3;
    if (pp_l || pp_u) {
        ensure_packs(e);
        // pass lists for the lanes the packed sweeps will run with
        PhaseTimer pt("packs", e->stream);
        if (pp_l) build_passes(e, e->pack_l, pp_lanes(effective_lanes(e, false)));
        if (pp_u) build_passes(e, e->pack_u, pp_lanes(effective_lanes(e, true)));
        pt.mark("pass lists");
    }
    if (e->mode_l == 5) ensure_ws(e, false, ws_lanes(effective_lanes(e, false)));
    if (e->mode_u == 5) ensure_ws(e, true, ws_lanes(effective_lanes(e, true)));
}


template <typename T, bool UPPER>
void sweep_ws_dispatch(int lanes, Ext* e, const double* b, double* out, double* reset, const int* gate, int* ctr) {
    lanes = ws_lanes(lanes);
    ensure_ws(e, UPPER, lanes);
    switch (lanes) {
        case 1: sweep_ws_one<1, 8, T, UPPER>(e, b, out, reset, gate, ctr); break;
        case 2: sweep_ws_one<2, 4, T, UPPER>(e, b, out, reset, gate, ctr); break;
        case 4: sweep_ws_one<4, 2, T, UPPER>(e, b, out, reset, gate, ctr); break;
        case 8: sweep_ws_one<8, 1, T, UPPER>(e, b, out, reset, gate, ctr); break;
        default: sweep_ws_one<16, 1, T, UPPER>(e, b, out, reset, gate, ctr); break;
    }
}

template <int G, typename T, bool UPPER>
void sweep_one(Ext* e, const T* val, const double* b, double* out, double* reset, const int* gate, int* ctr) {
    constexpr int rows_per_block = (64 / G) * kWavesPerBlock;
    const int grid = grid_for(e->n, rows_per_block);
    k_sweep<G, T, UPPER><<<grid, kBlock, 0, e->stream>>>(e->n, UPPER ? e->perm_u.get() : e->perm_l.get(),
                                                        e->rp.get(), e->ci.get(), e->dg.get(), val, b, out,
                                                        reset, gate, ctr, e->err.get() + kErrSweep,
                                                        watchdog_ticks(e->watchdog_s));
}

template <int G, typename T, bool UPPER>
void sweep_blk_one(Ext* e, const T* val, const double* b, double* out, double* reset, const int* gate, int* ctr) {
    const size_t lds = sizeof(double) * static_cast<size_t>(e->blk_rows);
    auto kern = k_sweep_blk<G, T, UPPER>;
    PILU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(lds)));
    kern<<<e->nblk, 512, lds, e->stream>>>(e->n, e->blk_rows, e->nblk, UPPER ? e->bperm_u.get() : e->bperm_l.get(),
                                          UPPER ? e->blev_u.get() : e->blev_l.get(),
                                          UPPER ? e->loff_u.get() : e->loff_l.get(), e->rp.get(), e->ci.get(),
                                          e->dg.get(), val, b, out, reset, gate, ctr, e->err.get() + kErrSweep,
                                          watchdog_ticks(e->watchdog_s));
}

template <typename T, bool UPPER>
void sweep_blk_dispatch(int lanes, Ext* e, const T* val, const double* b, double* out, double* reset,
                        const int* gate, int* ctr) {
    switch (lanes) {
        case 1: sweep_blk_one<1, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
        case 2: sweep_blk_one<2, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
        case 4: sweep_blk_one<4, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
        case 8: sweep_blk_one<8, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
        case 16: sweep_blk_one<16, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
        case 32: sweep_blk_one<32, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
        default: sweep_blk_one<64, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
    }
}

// LDS of the packed sweep: the block's x (b on entry) and its pass list.
size_t pp_lds_bytes(const Ext* e, const SweepPack& pk) {
    return sizeof(double) * static_cast<size_t>(e->max_block_rows) +
           sizeof(PassDesc) * static_cast<size_t>(pk.max_np);
}
constexpr size_t kMaxLds = 160 * 1024;

// True when the packed sweep with G lanes per row fits the LDS in both
// directions for the current partition.
bool pp_fits(Ext* e, int G) {
    ensure_packs(e);
    build_passes(e, e->pack_l, G);
    build_passes(e, e->pack_u, G);
    return pp_lds_bytes(e, e->pack_l) <= kMaxLds && pp_lds_bytes(e, e->pack_u) <= kMaxLds;
}

template <int G, int S, typename T, bool UPPER, bool FLAGS>
void sweep_pp_one(Ext* e, const T* pval, const T* pdiag, const double* b, double* out, double* reset, const int* gate,
                  int* ctr) {
    SweepPack& pk = UPPER ? e->pack_u : e->pack_l;
    build_passes(e, pk, G);
    const size_t lds = pp_lds_bytes(e, pk);
    if (lds > kMaxLds) throw std::runtime_error("packed sweep: block and pass list exceed the LDS");
    // the kernel reads these unconditionally (clamped indices): refuse to
    // launch on missing arrays
    const size_t nv = sizeof(T) == 4 ? pk.pval32.size() : pk.pval64.size();
    const size_t nd = sizeof(T) == 4 ? pk.pdiag32.size() : pk.pdiag64.size();
    if (!pval || nv < static_cast<size_t>(std::max<long>(pk.ne, 1)) || (UPPER && (!pdiag || nd < static_cast<size_t>(e->n))) ||
        !pk.kptr.get() || !pk.desc_ptr.get() || !pk.pdesc.get() || pk.rowinfo.size() < 2 * static_cast<size_t>(e->n) ||
        pk.pcol.size() < static_cast<size_t>(std::max<long>(pk.ne, 1)))
        throw std::runtime_error("packed sweep: pack arrays missing (factorize before sweeping)");
    auto kern = k_sweep_pp<G, S, kPpUnroll, 4, 2, T, UPPER, FLAGS>;
    PILU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(lds)));
    kern<<<e->nblk, kPpThreads, lds, e->stream>>>(pk.kptr.get(), pk.desc_ptr.get(),
                                          reinterpret_cast<const PassDesc*>(pk.pdesc.get()),
                                          reinterpret_cast<const int2*>(pk.rowinfo.get()), pk.pcol.get(), pval, pdiag,
                                          b, out, reset, gate, ctr, e->err.get() + kErrSweep,
                                          watchdog_ticks(e->watchdog_s), e->max_block_rows);
}

template <int G, typename T, bool UPPER>
void sweep_pp_s(Ext* e, const double* b, double* out, double* reset, const int* gate, int* ctr) {
    SweepPack& pk = UPPER ? e->pack_u : e->pack_l;
    const T* pval;
    const T* pdiag;
    if constexpr (sizeof(T) == 4) {
        pval = pk.pval32.get();
        pdiag = UPPER ? pk.pdiag32.get() : nullptr;
    } else {
        pval = pk.pval64.get();
        pdiag = UPPER ? pk.pdiag64.get() : nullptr;
    }
    const int need = (pk.max_w + G - 1) / G;
    if ((UPPER ? e->mode_u : e->mode_l) == 3) {
        if (need <= 1) sweep_pp_one<G, 1, T, UPPER, true>(e, pval, pdiag, b, out, reset, gate, ctr);
        else if (need <= 2) sweep_pp_one<G, 2, T, UPPER, true>(e, pval, pdiag, b, out, reset, gate, ctr);
        else sweep_pp_one<G, 4, T, UPPER, true>(e, pval, pdiag, b, out, reset, gate, ctr);
        return;
    }
    if (need <= 1) sweep_pp_one<G, 1, T, UPPER, false>(e, pval, pdiag, b, out, reset, gate, ctr);
    else if (need <= 2) sweep_pp_one<G, 2, T, UPPER, false>(e, pval, pdiag, b, out, reset, gate, ctr);
    else sweep_pp_one<G, 4, T, UPPER, false>(e, pval, pdiag, b, out, reset, gate, ctr);
}

template <typename T, bool UPPER>
void sweep_pp_dispatch(int lanes, Ext* e, const double* b, double* out, double* reset, const int* gate, int* ctr) {
    switch (lanes) {
        case 1: sweep_pp_s<1, T, UPPER>(e, b, out, reset, gate, ctr); break;
        case 2: sweep_pp_s<2, T, UPPER>(e, b, out, reset, gate, ctr); break;
        case 4: sweep_pp_s<4, T, UPPER>(e, b, out, reset, gate, ctr); break;
        case 8: sweep_pp_s<8, T, UPPER>(e, b, out, reset, gate, ctr); break;
        default: sweep_pp_s<16, T, UPPER>(e, b, out, reset, gate, ctr); break;
    }
}

// Most (level, row slot) descriptors (16 bytes each) the level walk may
// allocate per direction: 64 per row, or 2^26 (1 GiB) for small n.  A
// chain-like schedule (levels close to n) with many row slots per level
// exceeds it and is refused; the autotuner only tries shapes within half
// of the per-row budget.
long lv_desc_cap(long n) { return std::max(64L * n, 1L << 26); }
long lv_desc_tune_cap(long n) { return 32L * n; }

template <int G, typename T, bool UPPER>
void sweep_lv_one(Ext* e, const T* val, const double* b, double* out, double* reset, const int* gate) {
    constexpr int slots = kLvThreads / G;
    const int nlev = UPPER ? e->nlev_u : e->nlev_l;
    if (e->n <= 0 || nlev <= 0) return;
    DevBuf<int4>& desc = UPPER ? e->lvdesc_u : e->lvdesc_l;
    int* key = UPPER ? e->lvkey_u : e->lvkey_l;
    const DevBuf<int>& off = UPPER ? e->lvoff_u : e->lvoff_l;
    const DevBuf<int>& perm = UPPER ? e->perm_u : e->perm_l;
    const long nd = static_cast<long>(nlev) * slots;
    if (off.size() < static_cast<size_t>(nlev) + 1 || perm.size() < static_cast<size_t>(e->n))
        throw std::runtime_error("level-walk sweep: level schedule missing");
    if (nd > lv_desc_cap(e->n))
        throw std::runtime_error("level-walk sweep: levels x row slots exceed the descriptor budget "
                                 "(use fewer lanes per row or another sweep mode)");
    if (e->lvprog.size() < 3) {
        refuse_if_no_alloc(e, "level-walk progress words");
        e->lvprog.reserve(3);
        PILU_HIP(hipMemset(e->lvprog.get(), 0, 3 * sizeof(unsigned long long)));
    }
    int& cached = UPPER ? e->lvcached_u : e->lvcached_l;
    if (key[0] != e->pattern_version || key[1] != slots) {
        refuse_if_no_alloc(e, "level-walk descriptors");
        int* nbad = reinterpret_cast<int*>(e->lvprog.get() + 2);
        e->lvbad.reserve(static_cast<size_t>(nlev));
        PILU_HIP(hipMemsetAsync(nbad, 0, sizeof(int), e->stream));
        k_lv_slots<UPPER><<<nlev, 256, 0, e->stream>>>(off.get(), perm.get(), e->rp.get(), e->dg.get(), e->ci.get(),
                                                      e->lvbad.get(), nbad);
        PILU_HIP(hipGetLastError());
        desc.reserve(static_cast<size_t>(nd));
        k_lv_desc<UPPER><<<grid_for(nd, 256), 256, 0, e->stream>>>(nlev, slots, off.get(), perm.get(), e->rp.get(),
                                                                  e->dg.get(), e->lvbad.get(), desc.get());
        PILU_HIP(hipGetLastError());
        cached = nlev - read_int(nbad, e->stream);
        key[0] = e->pattern_version;
        key[1] = slots;
    }
    // workgroup 0 solves; workgroups 8, 16, ..., 8 * kLvHelpers (on its XCD)
    // warm its L2; the others return at once
    const int grid = 1 + 8 * kLvHelpers;
    const unsigned epoch = ++e->lv_epoch;
    unsigned long long* prog = e->lvprog_scratch ? e->lvprog_scratch : e->lvprog.get();
    auto kern = k_sweep_lv<G, T, UPPER>;
    PILU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(kLvLds)));
    kern<<<grid, kLvThreads, kLvLds, e->stream>>>(nlev, off.get(), desc.get(), perm.get(), e->rp.get(), e->ci.get(),
                                                 e->dg.get(), val, e->nnz, b, out, reset, gate, prog, epoch,
                                                 reinterpret_cast<int*>(prog + 1),
                                                 watchdog_ticks(e->watchdog_s), cached > 0 ? 1 : 0);
}

template <typename T, bool UPPER>
void sweep_dispatch(int lanes, Ext* e, const T* val, const double* b, double* out, double* reset,
                    const int* gate, int* ctr) {
    const int mode = UPPER ? e->mode_u : e->mode_l;
    if (mode == 4) {
        switch (lanes) {
            case 1: sweep_lv_one<1, T, UPPER>(e, val, b, out, reset, gate); break;
            case 2: sweep_lv_one<2, T, UPPER>(e, val, b, out, reset, gate); break;
            case 4: sweep_lv_one<4, T, UPPER>(e, val, b, out, reset, gate); break;
            case 8: sweep_lv_one<8, T, UPPER>(e, val, b, out, reset, gate); break;
            case 16: sweep_lv_one<16, T, UPPER>(e, val, b, out, reset, gate); break;
            case 32: sweep_lv_one<32, T, UPPER>(e, val, b, out, reset, gate); break;
            default: sweep_lv_one<64, T, UPPER>(e, val, b, out, reset, gate); break;
        }
        return;
    }
    if (mode == 5 && e->nblk > 0) {
        sweep_ws_dispatch<T, UPPER>(lanes, e, b, out, reset, gate, ctr);
        return;
    }
    if (mode >= 2 && mode <= 3 && e->nblk > 0) {
        ensure_packs(e);
        sweep_pp_dispatch<T, UPPER>(lanes, e, b, out, reset, gate, ctr);
        return;
    }
    if (mode == 1 && e->blk_rows > 0) {
        sweep_blk_dispatch<T, UPPER>(lanes, e, val, b, out, reset, gate, ctr);
        return;
    }
    switch (lanes) {
        case 1: sweep_one<1, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
        case 2: sweep_one<2, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
        case 4: sweep_one<4, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
        case 8: sweep_one<8, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
        case 16: sweep_one<16, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
        case 32: sweep_one<32, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
        default: sweep_one<64, T, UPPER>(e, val, b, out, reset, gate, ctr); break;
    }
}

int effective_lanes(const Ext* e, bool upper) {
    const int set = upper ? e->lanes_u : e->lanes_l;
    if (set > 0) return set;
    const double tri = e->n > 0 ? static_cast<double>(e->nnz - e->n) / (2.0 * e->n) : 1.0;
    return lanes_for_avg(tri);
}

PivotGuard pivot_guard(Ext* e) {
    return PivotGuard{e->pivot_mode, e->pivot_tiny, e->pivot_repl, e->err.get() + kErrPivotRow};
}

template <int G, int CAP, int P>
void factor_short_one(Ext* e, float* f32) {
    constexpr int groups = kBlock / G;
    const int grid = grid_for(e->n, groups);
    k_factor_short<G, CAP, P><<<grid, kBlock, 0, e->stream>>>(e->n, e->perm_l.get(), e->rp.get(), e->ci.get(),
                                                             e->dg.get(), e->vin.get(), e->fac.get(), f32,
                                                             e->err.get() + kErrFactor, e->ctr.get() + 2,
                                                             watchdog_ticks(e->watchdog_s), e->omega_ptr,
                                                             pivot_guard(e));
}

template <int G, int CAP>
void factor_one(Ext* e, float* f32) {
    constexpr int groups = kBlock / G;
    const int grid = grid_for(e->n, groups);
    k_factor<G, CAP><<<grid, kBlock, 0, e->stream>>>(e->n, e->perm_l.get(), e->rp.get(), e->ci.get(),
                                                    e->dg.get(), e->vin.get(), e->fac.get(), f32,
                                                    e->err.get() + kErrFactor, e->ctr.get() + 2,
                                                    watchdog_ticks(e->watchdog_s), e->omega_ptr, pivot_guard(e));
}

// Runs the factorization kernel on e->vin; returns false on a zero pivot.
bool run_factor(Ext* e) {
    hipStream_t s = e->stream;
    // the factors and packs are overwritten from here on: until this run
    // completes, no sweep may use them
    e->factored = false;
    float* f32 = nullptr;
    if (e->precision == 32) {
        e->fac32.reserve(e->nnz);
        f32 = e->fac32.get();
    }
    PILU_HIP(hipMemsetAsync(e->err.get() + kErrFactor, 0, sizeof(int), s));
    if (e->pivot_mode == 1) k_fill_int<<<1, 64, 0, s>>>(1, e->err.get() + kErrPivotRow, INT_MAX);
    e->factor_status = 0;
    const int mr = e->max_row;
    // short rows (stencils up to ILU(1) of a 3D 7-point operator): every
    // pivot row fetched at once, one entry per lane
    if (mr <= 32 && e->max_lo <= 8 && e->max_up <= 8) factor_short_one<8, 32, 8>(e, f32);
    else if (mr <= 32) factor_one<8, 32>(e, f32);
    else if (mr <= 64) factor_one<16, 64>(e, f32);
    else if (mr <= 256) factor_one<32, 256>(e, f32);
    else if (mr <= 2560) factor_one<64, 2560>(e, f32);
    else throw std::runtime_error("row longer than 2560 entries: not supported by the device factorization");
    PILU_HIP(hipGetLastError());
    const int err = read_int(e->err.get() + kErrFactor, s);
    if (err & kErrTimeout) throw std::runtime_error("device factorization: dependency wait timed out (watchdog)");
    if (err & 2) throw std::runtime_error("device factorization: row exceeded the LDS staging capacity");
    if (e->pivot_mode == 1) {
        const int first = read_int(e->err.get() + kErrPivotRow, s);
        if (first != INT_MAX) e->factor_status = (first & 1) ? -103 : -102;
    }
    e->factored = true;
    ++e->factor_version;
    gather_pack(e, e->pack_l, false);
    gather_pack(e, e->pack_u, true);
    if (ws_current(e, e->ws_l, e->ws_l.g)) gather_ws(e, e->ws_l);
    if (ws_current(e, e->ws_u, e->ws_u.g)) gather_ws(e, e->ws_u);
    return (err & 1) == 0 && e->factor_status == 0;
}

Ext* need_setup(Ext* e) {
    if (!e || e->n <= 0 || !e->analysed)
        throw std::runtime_error("IluSolver: SetupMatrix has not been called or did not complete");
    return e;
}

}  // namespace

// Prints the schedule description when PAR_ILU_SHOW_ALGORITHM is set.
void show_algorithm(const lljbash::IluSolver* solver) {
    const char* v = std::getenv("PAR_ILU_SHOW_ALGORITHM");
    if (v && *v && std::strcmp(v, "0") != 0) std::printf("%s\n", solver->DescribeSchedule().c_str());
}

// Sweep kinds whose output must hold kPending in every slot on entry
// (value-flag synchronisation through the output vector).
bool flags_in_output(int mode) { return mode >= 0 && mode <= 3; }

// The forward sweep b -> y, then the backward sweep y -> x.  Each sweep
// re-arms the other one's output where the other one synchronises through
// it: the forward sweep re-arms x (after reading b, which may alias x),
// the backward sweep re-arms y for the next forward sweep.  y must hold
// kPending on entry when the forward sweep needs it (y_armed tracks that
// for the solver's own y).  ev (optional): events recorded before, between
// and after the two sweeps.
void sweep_pair(Ext* e, const double* d_b, double* d_x, double* y, const int* gate, int* ctr, hipEvent_t* ev) {
    hipStream_t s = e->stream;
    const int gl = effective_lanes(e, false), gu = effective_lanes(e, true);
    double* reset_l = flags_in_output(e->mode_u) ? d_x : nullptr;
    double* reset_u = flags_in_output(e->mode_l) ? y : nullptr;
    if (ev) PILU_HIP(hipEventRecord(ev[0], s));
    if (e->precision == 32) sweep_dispatch<float, false>(gl, e, e->fac32.get(), d_b, y, reset_l, gate, ctr);
    else sweep_dispatch<double, false>(gl, e, e->fac.get(), d_b, y, reset_l, gate, ctr);
    if (ev) PILU_HIP(hipEventRecord(ev[1], s));
    if (e->precision == 32) sweep_dispatch<float, true>(gu, e, e->fac32.get(), y, d_x, reset_u, gate, ctr + 2);
    else sweep_dispatch<double, true>(gu, e, e->fac.get(), y, d_x, reset_u, gate, ctr + 2);
    if (ev) PILU_HIP(hipEventRecord(ev[2], s));
    PILU_HIP(hipGetLastError());
}

// Fills the solver's y with kPending when the forward sweep synchronises
// through it and the last backward sweep did not re-arm it.
void arm_y(Ext* e) {
    if (flags_in_output(e->mode_l) && !e->y_armed)
        k_fill_u64<<<grid_for(e->n, 256 * 8), 256, 0, e->stream>>>(e->n, reinterpret_cast<uint64_t*>(e->y.get()),
                                                                    kPending);
}

void ilu_apply(Ext* e, const double* d_b, double* d_x, const int* gate) {
    need_setup(e);
    if (!e->factored) throw std::runtime_error("IluSolver: Factorize has not been called");
    arm_y(e);
    sweep_pair(e, d_b, d_x, e->y.get(), gate, e->ctr.get() + 4, nullptr);
    e->y_armed = flags_in_output(e->mode_l);
}

__global__ void k_scale_vec(long n, double alpha, double* __restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] *= alpha;
}

// y = alpha T^-1 x for one triangle of the factors: one sweep of the
// configured kind for that direction, then the scale.  The sweep's output
// is armed with kPending first where that kind synchronises through it; x
// aliasing y is staged through the solver's y (which then needs re-arming).
void ilu_trsv(Ext* e, bool upper, double alpha, const double* d_x, double* d_y) {
    need_setup(e);
    if (!e->factored) throw std::runtime_error("IluSolver: Factorize has not been called");
    if (!d_x || !d_y) throw std::runtime_error("Trsv: null vector");
    hipStream_t s = e->stream;
    const double* in = d_x;
    if (d_x == d_y) {
        PILU_HIP(hipMemcpyAsync(e->y.get(), d_x, sizeof(double) * e->n, hipMemcpyDeviceToDevice, s));
        in = e->y.get();
        e->y_armed = false;
    }
    if (flags_in_output(upper ? e->mode_u : e->mode_l))
        k_fill_u64<<<grid_for(e->n, 256 * 8), 256, 0, s>>>(e->n, reinterpret_cast<uint64_t*>(d_y), kPending);
    const int g = effective_lanes(e, upper);
    int* ctr = e->ctr.get() + (upper ? 6 : 4);
    if (upper) {
        if (e->precision == 32) sweep_dispatch<float, true>(g, e, e->fac32.get(), in, d_y, nullptr, nullptr, ctr);
        else sweep_dispatch<double, true>(g, e, e->fac.get(), in, d_y, nullptr, nullptr, ctr);
    } else {
        if (e->precision == 32) sweep_dispatch<float, false>(g, e, e->fac32.get(), in, d_y, nullptr, nullptr, ctr);
        else sweep_dispatch<double, false>(g, e, e->fac.get(), in, d_y, nullptr, nullptr, ctr);
    }
    if (alpha != 1.0) k_scale_vec<<<grid_for(e->n, 256 * 8), 256, 0, s>>>(e->n, alpha, d_y);
    PILU_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------
// Scratch of the no-allocation apply: the forward sweep's output y
// (kScratchRows doubles), then 8 words for the level walk's progress (words
// 0-2) and 16 ticket counters (as ints, words 8-15).  One copy per device,
// held until the process ends; every no-allocation apply on a device uses
// it, so such calls must not overlap on one device.
// ---------------------------------------------------------------------
constexpr long kScratchRows = 1L << 22;
constexpr long kScratchWords = kScratchRows + 16;
static __device__ __align__(256) double scratch[kScratchWords];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current (the host-side symbol is only a
// stand-in and must never be passed to a kernel).
double* device_scratch() {
    int dev = 0;
    PILU_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) throw std::runtime_error("no-allocation apply: device index too large");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        PILU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("no-allocation apply: scratch symbol has no device address");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

// x = U^-1 L^-1 b like ilu_apply, with y, the ticket counters and the level
// walk's progress words in the device scratch instead of the solver's own
// buffers; allocates nothing.  Every sweep structure the configured sweep
// needs (packs, pass lists, level-walk descriptors) must already exist.
void ilu_apply_no_alloc(Ext* e, const double* d_b, double* d_x) {
    need_setup(e);
    if (!e->factored) throw std::runtime_error("IluSolver: Factorize has not been called");
    if (e->n > kScratchRows)
        throw std::runtime_error("no-allocation apply: n exceeds the scratch capacity of 4194304 rows");
    double* base = device_scratch();
    double* y = base;
    unsigned long long* prog = reinterpret_cast<unsigned long long*>(base + kScratchRows);
    int* ctr = reinterpret_cast<int*>(base + kScratchRows + 8);
    struct Guard {
        Ext* e;
        Guard(Ext* x, unsigned long long* p) : e(x) { e->no_alloc = true; e->lvprog_scratch = p; }
        ~Guard() { e->no_alloc = false; e->lvprog_scratch = nullptr; }
    } guard(e, prog);
    hipStream_t s = e->stream;
    // y starts all-pending; progress words and ticket counters start at
    // zero (fill kernels, so the sequence can also be captured in a graph)
    k_fill_u64<<<grid_for(e->n, 256 * 8), 256, 0, s>>>(e->n, reinterpret_cast<uint64_t*>(y), kPending);
    k_fill_u64<<<1, 64, 0, s>>>(16, reinterpret_cast<uint64_t*>(prog), 0);
    sweep_pair(e, d_b, d_x, y, nullptr, ctr, nullptr);
}

// Schedule of independent diagonal blocks given by block_of (the systems of
// a block-diagonal matrix): rows sorted by (block, level within the block,
// row), blev[b] the first (block, level) group of block b, loff[g] the first
// position of group g.
void block_level_schedule(Ext* e, bool upper, const int* block_of, int nblk, DevBuf<int>& perm, DevBuf<int>& blev,
                          DevBuf<int>& loff) {
    std::vector<int> pos(static_cast<size_t>(nblk));
    for (int b = 0; b < nblk; ++b) pos[b] = b;
    build_block_schedule(e, upper, block_of, nblk, pos, perm, blev, loff);
}

}  // namespace pilu

namespace lljbash {

using pilu::DevBuf;

IluSolver::~IluSolver() {
    delete ext_;
    ext_ = nullptr;
    DestroyCsrMatrix(&aMatrix_);
}

void IluSolver::SetupMatrix() {
    pilu::SetupScope scope(nullptr);
    const int n = aMatrix_.size;
    if (n <= 0 || !aMatrix_.row_ptr) throw std::runtime_error("IluSolver: empty matrix");
    const long nnz = aMatrix_.row_ptr[n];
    if (aMatrix_.row_ptr[0] != 0 || nnz < 0 || nnz > aMatrix_.max_nnz)
        throw std::runtime_error("IluSolver: row_ptr inconsistent with the matrix storage");
    if (!ext_) ext_ = new Ext;
    Ext* e = ext_;
    e->analysed = false;
    e->n = n;
    e->nnz = nnz;
    e->rp.reserve(static_cast<size_t>(n) + 1);
    e->ci.reserve(std::max<long>(nnz, 1));
    PILU_HIP(hipMemcpy(e->rp.get(), aMatrix_.row_ptr, sizeof(int) * (static_cast<size_t>(n) + 1),
                       hipMemcpyHostToDevice));
    if (nnz > 0)
        PILU_HIP(hipMemcpy(e->ci.get(), aMatrix_.col_idx, sizeof(int) * nnz, hipMemcpyHostToDevice));
    pilu::analyse(e);
    pilu::show_algorithm(this);
}

void IluSolver::SetupMatrixDevice(int n, int nnz, const int* d_row_ptr, const int* d_col_idx) {
    pilu::SetupScope scope(nullptr);
    if (n <= 0) throw std::runtime_error("IluSolver: empty matrix");
    if (!ext_) ext_ = new Ext;
    Ext* e = ext_;
    e->analysed = false;
    e->n = n;
    e->nnz = nnz;
    e->rp.reserve(static_cast<size_t>(n) + 1);
    e->ci.reserve(std::max(nnz, 1));
    PILU_HIP(hipMemcpyAsync(e->rp.get(), d_row_ptr, sizeof(int) * (static_cast<size_t>(n) + 1),
                            hipMemcpyDeviceToDevice, e->stream));
    PILU_HIP(hipMemcpyAsync(e->ci.get(), d_col_idx, sizeof(int) * static_cast<size_t>(nnz),
                            hipMemcpyDeviceToDevice, e->stream));
    pilu::analyse(e);
}

bool IluSolver::Factorize() {
    Ext* e = pilu::need_setup(ext_);
    if (aMatrix_.size != e->n || aMatrix_.row_ptr[e->n] != e->nnz)
        throw std::runtime_error("IluSolver: matrix changed since SetupMatrix");
    hipStream_t s = e->stream;
    PILU_HIP(hipMemcpyAsync(e->vin.get(), aMatrix_.value, sizeof(double) * e->nnz, hipMemcpyHostToDevice, s));
    pilu::k_prepare_factor<<<pilu::grid_for(e->nnz, 256 * 8), 256, 0, s>>>(
        e->nnz, e->vin.get(), reinterpret_cast<uint64_t*>(e->fac.get()), false);
    const bool ok = pilu::run_factor(e);
    PILU_HIP(hipMemcpyAsync(aMatrix_.value, e->fac.get(), sizeof(double) * e->nnz, hipMemcpyDeviceToHost, s));
    PILU_HIP(hipStreamSynchronize(s));
    return ok;
}

bool IluSolver::FactorizeDevice(const double* d_values) {
    Ext* e = pilu::need_setup(ext_);
    hipStream_t s = e->stream;
    PILU_HIP(hipMemcpyAsync(e->vin.get(), d_values, sizeof(double) * e->nnz, hipMemcpyDeviceToDevice, s));
    pilu::k_prepare_factor<<<pilu::grid_for(e->nnz, 256 * 8), 256, 0, s>>>(
        e->nnz, e->vin.get(), reinterpret_cast<uint64_t*>(e->fac.get()), false);
    return pilu::run_factor(e);
}

int IluSolver::SymbolicDevice(int n, const int* d_row_ptr, const int* d_col_idx, int lof) {
    pilu::SetupScope scope(nullptr);
    if (!ext_) ext_ = new Ext;
    Ext* e = ext_;
    e->analysed = false;  // the previous pattern is gone whether or not this completes
    pilu::SymbolicResult res;
    pilu::PhaseTimer pt("symbolic", e->stream);
    const int nnz_f = pilu::symbolic_device(n, d_row_ptr, d_col_idx, nullptr, lof, res, e->stream,
                                            2.0 * e->watchdog_s);
    pt.mark("symbolic ILU(k)");
    if (nnz_f < 0) return -1;
    int nnz_a = 0;
    PILU_HIP(hipMemcpy(&nnz_a, d_row_ptr + n, sizeof(int), hipMemcpyDeviceToHost));
    e->n = n;
    e->nnz = nnz_f;
    e->nnz_a = nnz_a;
    e->rp = std::move(res.rp);
    e->ci = std::move(res.ci);
    e->nzmap = std::move(res.nzmap);
    e->fill_level = std::move(res.lev);
    pilu::analyse(e);
    pilu::show_algorithm(this);
    return nnz_f;
}

bool IluSolver::FactorizeDeviceFromA(const double* d_a_values) {
    Ext* e = pilu::need_setup(ext_);
    if (!e->nzmap.get() || e->nnz_a <= 0) throw std::runtime_error("IluSolver: no A-to-F map (call SymbolicDevice)");
    hipStream_t s = e->stream;
    pilu::k_prepare_factor<<<pilu::grid_for(e->nnz, 256 * 8), 256, 0, s>>>(
        e->nnz, e->vin.get(), reinterpret_cast<uint64_t*>(e->fac.get()), true);
    pilu::k_scatter_values<<<pilu::grid_for(e->nnz_a, 256 * 8), 256, 0, s>>>(e->nnz_a, e->nzmap.get(), d_a_values,
                                                                             e->vin.get());
    return pilu::run_factor(e);
}

int IluSolver::CopyFillLevels(signed char* out) const {
    if (!ext_ || !ext_->fill_level.get()) return 0;
    PILU_HIP(hipMemcpy(out, ext_->fill_level.get(), sizeof(signed char) * ext_->nnz, hipMemcpyDeviceToHost));
    return static_cast<int>(ext_->nnz);
}

void IluSolver::Substitute(const double* b, double* x) {
    Ext* e = pilu::need_setup(ext_);
    hipStream_t s = e->stream;
    e->hb.reserve(e->n);
    PILU_HIP(hipMemcpyAsync(e->hb.get(), b, sizeof(double) * e->n, hipMemcpyHostToDevice, s));
    pilu::ilu_apply(e, e->hb.get(), e->hb.get(), nullptr);
    PILU_HIP(hipMemcpyAsync(x, e->hb.get(), sizeof(double) * e->n, hipMemcpyDeviceToHost, s));
    pilu::check_sweep_errors(e);
}

void IluSolver::SubstituteDevice(const double* d_b, double* d_x) {
    pilu::ilu_apply(pilu::need_setup(ext_), d_b, d_x, nullptr);
}

void IluSolver::SubstituteDeviceNoAlloc(const double* d_b, double* d_x) {
    pilu::ilu_apply_no_alloc(pilu::need_setup(ext_), d_b, d_x);
}

void IluSolver::SetPivotPolicy(int policy, double tiny, double replacement) {
    if (policy < 0 || policy > 2) throw std::runtime_error("pivot policy must be 0 (none), 1 (stop) or 2 (replace)");
    if (!(tiny >= 0.0) || !std::isfinite(replacement)) throw std::runtime_error("pivot policy: bad tiny or replacement");
    if (!ext_) ext_ = new Ext;
    ext_->pivot_mode = policy;
    ext_->pivot_tiny = tiny;
    ext_->pivot_repl = replacement;
}

int IluSolver::FactorStatus() const { return ext_ ? ext_->factor_status : 0; }

void IluSolver::TrsvDevice(bool upper, double alpha, const double* d_x, double* d_y) {
    pilu::ilu_trsv(pilu::need_setup(ext_), upper, alpha, d_x, d_y);
}

void IluSolver::Trsv(bool upper, double alpha, const double* x, double* y) {
    Ext* e = pilu::need_setup(ext_);
    hipStream_t s = e->stream;
    e->hb.reserve(e->n);
    PILU_HIP(hipMemcpyAsync(e->hb.get(), x, sizeof(double) * e->n, hipMemcpyHostToDevice, s));
    pilu::ilu_trsv(e, upper, alpha, e->hb.get(), e->hb.get());
    PILU_HIP(hipMemcpyAsync(y, e->hb.get(), sizeof(double) * e->n, hipMemcpyDeviceToHost, s));
    pilu::check_sweep_errors(e);
}

void IluSolver::SetWatchdog(double seconds) {
    if (!(seconds >= 0.0) || seconds > 3600.0) throw std::runtime_error("watchdog limit must be in [0, 3600] s");
    if (!ext_) ext_ = new Ext;
    ext_->watchdog_s = seconds;
}

void IluSolver::SetDiagonalCompensation(const float* d_omega) {
    if (!ext_) ext_ = new Ext;
    ext_->omega_ptr = d_omega;
}

}  // namespace lljbash

namespace pilu {
namespace {
// Average ms of the forward and backward sweeps under the current settings.
void time_sweeps(Ext* e, int lanes, const double* b, double* x, int reps, hipEvent_t* ev, double& tl, double& tu) {
    hipStream_t s = e->stream;
    int* ctr = e->ctr.get() + 4;
    e->lanes_l = e->lanes_u = lanes;
    ilu_apply(e, b, x, nullptr);  // warm-up; also arms the buffers
    tl = tu = 0.0;
    (void)s;
    (void)ctr;
    for (int r = 0; r < reps; ++r) {
        arm_y(e);
        sweep_pair(e, b, x, e->y.get(), nullptr, e->ctr.get() + 4, ev);
        e->y_armed = flags_in_output(e->mode_l);
        PILU_HIP(hipEventSynchronize(ev[2]));
        float a = 0.f, c = 0.f;
        PILU_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
        PILU_HIP(hipEventElapsedTime(&c, ev[1], ev[2]));
        tl += a;
        tu += c;
    }
    tl /= reps;
    tu /= reps;
}
}  // namespace
}  // namespace pilu

namespace lljbash {

// Average rows per level up to which the autotuner tries the
// single-workgroup level walk (mode 4).
constexpr int kLvTryRows = 256;

int IluSolver::AutotuneSweeps(int reps, double* table, int cap) {
    pilu::SetupScope scope(nullptr);
    Ext* e = pilu::need_setup(ext_);
    if (!e->factored) throw std::runtime_error("IluSolver: Factorize before AutotuneSweeps");
    reps = std::max(reps, 1);
    hipStream_t s = e->stream;
    static const int kLanes[7] = {1, 2, 4, 8, 16, 32, 64};
    static const int kBlkLanes[4] = {4, 8, 16, 32};
    static const int kPackLanes[4] = {2, 4, 8, 16};
    static const int kWsLanes[3] = {1, 2, 4};
    struct Part { int blk, tx, ty, tz; };
    std::vector<Part> parts;
    for (int blk : {4096, 8192, 16384})
        if (blk <= e->n) parts.push_back({blk, 0, 0, 0});
    if (e->grid[0] > 0) {
        // tiles spanning the whole x extent (up to 128) and several planes:
        // few tiles on the long z and y dependency chains, ~4-8k rows each
        const int tx = std::min(e->grid[0], 128);
        const int kTiles[6][3] = {{tx, 4, 8}, {tx, 8, 6}, {tx, 6, 8}, {tx, 3, 10}, {tx, 8, 4}, {(tx * 3 + 3) / 4, 4, 8}};
        for (const auto& t : kTiles)
            if (t[1] <= e->grid[1] && t[2] <= e->grid[2] && static_cast<long>(t[0]) * t[1] * t[2] < e->n &&
                static_cast<long>(t[0]) * t[1] * t[2] <= pilu::kMaxBlockRows)
                parts.push_back({0, t[0], t[1], t[2]});
    }
    pilu::DevBuf<double> b(e->n), x(e->n);
    pilu::k_fill_u64<<<pilu::grid_for(e->n, 256 * 8), 256, 0, s>>>(e->n, reinterpret_cast<uint64_t*>(b.get()),
                                                                   0x3FF0000000000000ull);  // 1.0
    hipEvent_t ev[3];
    for (auto& v : ev) PILU_HIP(hipEventCreate(&v));
    int rows = 0;
    auto record = [&](int mode, const Part& pt, int lanes, double tl, double tu) {
        if (table && rows < cap) {
            double* r = table + 8 * rows;
            r[0] = mode; r[1] = pt.blk; r[2] = pt.tx; r[3] = pt.ty; r[4] = pt.tz; r[5] = lanes; r[6] = tl; r[7] = tu;
        }
        ++rows;
    };
    // level-scheduled rows
    e->mode_l = e->mode_u = 0;
    double lev_best[2] = {1e300, 1e300};
    int lev_lanes[2] = {8, 8};
    for (int g : kLanes) {
        double tl, tu;
        pilu::time_sweeps(e, g, b.get(), x.get(), reps, ev, tl, tu);
        record(0, Part{0, 0, 0, 0}, g, tl, tu);
        if (tl < lev_best[0]) { lev_best[0] = tl; lev_lanes[0] = g; }
        if (tu < lev_best[1]) { lev_best[1] = tu; lev_lanes[1] = g; }
    }
    // one workgroup walking the levels, where a level averages few rows
    // (a level wider than the workgroup's row slots is read without
    // prefetch, so wide levels are not tried)
    int lev_mode[2] = {0, 0};
    const int deep = std::max(std::max(e->nlev_l, e->nlev_u), 1);
    if (e->n / deep <= kLvTryRows) {
        e->mode_l = e->mode_u = 4;
        for (int g : {4, 8, 16, 32, 64}) {
            if (static_cast<long>(deep) * (pilu::kLvThreads / g) > pilu::lv_desc_tune_cap(e->n)) continue;
            double tl, tu;
            pilu::time_sweeps(e, g, b.get(), x.get(), reps, ev, tl, tu);
            record(4, Part{0, 0, 0, 0}, g, tl, tu);
            if (tl < lev_best[0]) { lev_best[0] = tl; lev_lanes[0] = g; lev_mode[0] = 4; }
            if (tu < lev_best[1]) { lev_best[1] = tu; lev_lanes[1] = g; lev_mode[1] = 4; }
        }
    }
    // block-local sweeps, per partition
    double best_total = lev_best[0] + lev_best[1];
    int pick_mode[2] = {lev_mode[0], lev_mode[1]}, pick_lanes[2] = {lev_lanes[0], lev_lanes[1]};
    int pick = -1;
    for (size_t ip = 0; ip < parts.size(); ++ip) {
        const Part& pt = parts[ip];
        if (pt.blk) pilu::use_blocks(e, pt.blk);
        else pilu::use_tiles(e, pt.tx, pt.ty, pt.tz);
        double bb[2] = {1e300, 1e300};
        int bl[2] = {8, 8}, bm[2] = {1, 1};
        if (pt.blk) {
            e->mode_l = e->mode_u = 1;
            for (int g : kBlkLanes) {
                double tl, tu;
                pilu::time_sweeps(e, g, b.get(), x.get(), reps, ev, tl, tu);
                record(1, pt, g, tl, tu);
                if (tl < bb[0]) { bb[0] = tl; bl[0] = g; bm[0] = 1; }
                if (tu < bb[1]) { bb[1] = tu; bl[1] = g; bm[1] = 1; }
            }
        }
        for (int mode : {2, 3}) {
            e->mode_l = e->mode_u = mode;
            for (int g : kPackLanes) {
                if (!pilu::pp_fits(e, g)) continue;
                double tl, tu;
                pilu::time_sweeps(e, g, b.get(), x.get(), reps, ev, tl, tu);
                record(mode, pt, g, tl, tu);
                if (tl < bb[0]) { bb[0] = tl; bl[0] = g; bm[0] = mode; }
                if (tu < bb[1]) { bb[1] = tu; bl[1] = g; bm[1] = mode; }
            }
        }
        // wave-specialised sweep over record streams (rows of up to a few
        // tens of entries; blocks whose x and imports fit the LDS)
        e->mode_l = e->mode_u = 5;
        for (int g : kWsLanes) {
            if (!pilu::ws_fits(e, g)) continue;
            double tl, tu;
            pilu::time_sweeps(e, g, b.get(), x.get(), reps, ev, tl, tu);
            record(5, pt, g, tl, tu);
            if (tl < bb[0]) { bb[0] = tl; bl[0] = g; bm[0] = 5; }
            if (tu < bb[1]) { bb[1] = tu; bl[1] = g; bm[1] = 5; }
        }
        const double tot = std::min(bb[0], lev_best[0]) + std::min(bb[1], lev_best[1]);
        if (tot < best_total) {
            best_total = tot;
            pick = static_cast<int>(ip);
            for (int d = 0; d < 2; ++d) {
                pick_mode[d] = bb[d] < lev_best[d] ? bm[d] : lev_mode[d];
                pick_lanes[d] = bb[d] < lev_best[d] ? bl[d] : lev_lanes[d];
            }
        }
    }
    for (auto& v : ev) (void)hipEventDestroy(v);
    pilu::check_sweep_errors(e);
    if (pick >= 0) {
        const Part& pt = parts[pick];
        if (pt.blk) pilu::use_blocks(e, pt.blk);
        else pilu::use_tiles(e, pt.tx, pt.ty, pt.tz);
    }
    e->mode_l = pick_mode[0];
    e->mode_u = pick_mode[1];
    e->lanes_l = pick_lanes[0];
    e->lanes_u = pick_lanes[1];
    pilu::prepare_modes(e);
    pilu::show_algorithm(this);
    return rows;
}

void IluSolver::SetSweepMode(int mode_l, int mode_u, int blk_rows) {
    pilu::SetupScope scope(nullptr);
    Ext* e = pilu::need_setup(ext_);
    if (mode_l < 0 || mode_l > 5 || mode_u < 0 || mode_u > 5)
        throw std::runtime_error("sweep mode is 0, 1, 2, 3, 4 or 5");
    auto blocked = [](int m) { return (m >= 1 && m <= 3) || m == 5; };
    if (blocked(mode_l) || blocked(mode_u)) pilu::use_blocks(e, blk_rows);
    e->mode_l = mode_l;
    e->mode_u = mode_u;
    pilu::prepare_modes(e);
    pilu::show_algorithm(this);
}

void IluSolver::SetGridShape(int nx, int ny, int nz) {
    Ext* e = pilu::need_setup(ext_);
    if (nx <= 0 || ny <= 0 || nz <= 0 || static_cast<long>(nx) * ny * nz != e->n)
        throw std::runtime_error("grid shape must satisfy nx * ny * nz == n");
    e->grid[0] = nx;
    e->grid[1] = ny;
    e->grid[2] = nz;
}

void IluSolver::SetSweepTiles(int mode_l, int mode_u, int tx, int ty, int tz) {
    pilu::SetupScope scope(nullptr);
    Ext* e = pilu::need_setup(ext_);
    auto ok = [](int m) { return m == 0 || m == 2 || m == 3 || m == 5; };
    if (!ok(mode_l) || !ok(mode_u)) throw std::runtime_error("tiled sweeps use mode 0, 2, 3 or 5");
    pilu::use_tiles(e, tx, ty, tz);
    e->mode_l = mode_l;
    e->mode_u = mode_u;
    pilu::prepare_modes(e);
    pilu::show_algorithm(this);
}

std::string IluSolver::DescribeSchedule() const {
    const Ext* e = ext_;
    if (!e || e->n <= 0) return "{}";
    char buf[1024];
    auto pack = [](const pilu::SweepPack& pk) {
        char b[256];
        std::snprintf(b, sizeof(b), "{\"entries\": %ld, \"max_width\": %d, \"lanes\": %d, \"max_passes_per_block\": %d}",
                      pk.ne, pk.max_w, pk.pass_g, pk.max_np);
        return std::string(b);
    };
    std::snprintf(buf, sizeof(buf),
                  "{\"n\": %d, \"nnz\": %ld, \"max_row\": %d, \"precision\": %d, \"levels_l\": %d, "
                  "\"levels_u\": %d, \"grid\": [%d, %d, %d], \"blk_rows\": %d, \"tile\": [%d, %d, %d], "
                  "\"skew\": [%d, %d, %d], "
                  "\"blocks\": %d, \"max_block_rows\": %d, \"mode_l\": %d, \"lanes_l\": %d, \"mode_u\": %d, "
                  "\"lanes_u\": %d, ",
                  e->n, e->nnz, e->max_row, e->precision, e->nlev_l, e->nlev_u, e->grid[0], e->grid[1], e->grid[2],
                  e->blk_rows, e->tile[0], e->tile[1], e->tile[2], e->skew[0], e->skew[1], e->skew[2], e->nblk,
                  e->max_block_rows, e->mode_l,
                  pilu::effective_lanes(e, false), e->mode_u, pilu::effective_lanes(e, true));
    std::string js(buf);
    const bool live_l = e->nblk > 0 && e->pack_l.version == e->part_version;
    const bool live_u = e->nblk > 0 && e->pack_u.version == e->part_version;
    js += "\"lv_cached_levels\": [" + std::to_string(e->lvcached_l) + ", " + std::to_string(e->lvcached_u) + "], ";
    js += "\"pack_l\": " + (live_l ? pack(e->pack_l) : std::string("null"));
    js += ", \"pack_u\": " + (live_u ? pack(e->pack_u) : std::string("null"));
    // wave-specialised streams: LDS per workgroup and how many fit a CU's 160 KiB
    auto ws = [&](const pilu::WsPack& wp) {
        if (wp.version != e->part_version || !wp.stream.get()) return std::string("null");
        const size_t lds = pilu::ws_lds_bytes(e, wp);
        char b[256];
        std::snprintf(b, sizeof(b),
                      "{\"lanes\": %d, \"stream_bytes\": %ld, \"max_imports\": %d, \"face_slots\": %d, \"lds_bytes\": %zu, "
                      "\"workgroups_per_cu\": %zu}",
                      wp.g, static_cast<long>(wp.nstream) * 16, wp.max_imp, wp.nface, lds,
                      static_cast<size_t>(160 * 1024) / std::max<size_t>(lds, 1));
        return std::string(b);
    };
    js += ", \"ws_l\": " + ws(e->ws_l) + ", \"ws_u\": " + ws(e->ws_u) + "}";
    return js;
}

void IluSolver::GetSweepConfig(int* cfg) const {
    const Ext* e = ext_;
    cfg[0] = e ? e->mode_l : 0;
    cfg[1] = e ? pilu::effective_lanes(e, false) : 0;
    cfg[2] = e ? e->mode_u : 0;
    cfg[3] = e ? pilu::effective_lanes(e, true) : 0;
    cfg[4] = e ? e->blk_rows : 0;
    cfg[5] = e ? e->tile[0] : 0;
    cfg[6] = e ? e->tile[1] : 0;
    cfg[7] = e ? e->tile[2] : 0;
}

void IluSolver::SetFactorPrecision(int bits) {
    if (bits != 32 && bits != 64) throw std::runtime_error("factor precision must be 32 or 64");
    if (!ext_) ext_ = new Ext;
    if (ext_->precision != bits) ext_->factored = false;
    ext_->precision = bits;
}

int IluSolver::FactorPrecision() const { return ext_ ? ext_->precision : 64; }

void IluSolver::SetSweepLanes(int lanes_l, int lanes_u) {
    auto ok = [](int g) { return g == 0 || (g >= 1 && g <= 64 && (g & (g - 1)) == 0); };
    if (!ok(lanes_l) || !ok(lanes_u)) throw std::runtime_error("sweep lanes must be 0 or a power of two <= 64");
    if (!ext_) ext_ = new Ext;
    ext_->lanes_l = lanes_l;
    ext_->lanes_u = lanes_u;
}

const int* IluSolver::DeviceRowPtr() const { return ext_ ? ext_->rp.get() : nullptr; }
const int* IluSolver::DeviceColIdx() const { return ext_ ? ext_->ci.get() : nullptr; }
const int* IluSolver::DeviceDiagPtr() const { return ext_ ? ext_->dg.get() : nullptr; }
const double* IluSolver::DeviceFactors() const { return ext_ ? ext_->fac.get() : nullptr; }
int IluSolver::NumLevelsL() const { return ext_ ? ext_->nlev_l : 0; }
int IluSolver::NumLevelsU() const { return ext_ ? ext_->nlev_u : 0; }
int IluSolver::Size() const { return ext_ ? ext_->n : 0; }
int IluSolver::Nonzeros() const { return ext_ ? static_cast<int>(ext_->nnz) : 0; }

}  // namespace lljbash
