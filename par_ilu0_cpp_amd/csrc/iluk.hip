// Symbolic ILU(k) on the GPU: IlukSymbolic (include/iluk.h; reference
// include/iluk.h:15, src/iluk.cpp:5-54).
//
// One wave per row, rows taken in ascending order through workgroup
// tickets.  A row is kept sorted in LDS as packed entries (col << 5 | level)
// and grows by merging: for each pivot k < i of level < lof, in ascending
// column order (including entries created by earlier pivots of the same
// row), the wave reads row k's entries right of its diagonal from a global
// pool, forms candidate levels lev(i,k) + lev(k,j) + 1 <= lof, drops the
// columns already present (first level found is kept) and merges the rest
// in.  A finished row is appended to the pool and announced through a
// per-row header; both are value-as-flag words, so a waiting row spins on
// exactly the words it reads.  A compaction pass then lays the rows out as
// CSR with the reference's value placement and nzmap.
//
// Rows that need no other row.  For lof <= 1 a pivot k of row i can only
// be one of A's own entries (level 0), and only level-0 entries of row k
// can give fill of level <= 1 -- exactly A's entries right of k's
// diagonal.  So when every row of A is sorted the pass reads pivot rows
// from A itself (`a_pivots`): no row waits for another, and all rows run at
// full occupancy instead of a window of ~8k ascending rows (on a 3D grid
// the planes would otherwise run nearly one after another).  The result is
// the same pattern, levels and order as the waiting form; lof >= 2, and A
// with a row out of order or without its diagonal, take the waiting form.
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "common.hpp"
#include "iluk.h"
#include "iluk_device.hpp"

namespace pilu {

namespace {

constexpr int kSymWaves = 4;           // waves (rows) per workgroup
constexpr int kSymGridA = 4096;        // workgroups of the form that reads pivot rows from A
// LDS entries per row buffer: a short-row form (more rows resident, so a
// longer stretch of the row order is in flight) and the general form, used
// when a row of F outgrows the short one.
constexpr int kSymCapShort = 64;
constexpr int kSymCapLong = 2048;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint64_t kNoHeader = ~0ull;
constexpr int kLevBits = 5;
constexpr uint32_t kLevMask = (1u << kLevBits) - 1u;

__device__ __forceinline__ uint32_t peek_u32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// First index in [lo, hi) of sorted packed entries whose column is >= col.
__device__ __forceinline__ int lower_col(const uint32_t* a, int lo, int hi, uint32_t col) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((a[mid] >> kLevBits) < col) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// err bits: 1 repeated column in a row of A, 2 row over CAP, 4 pool full.
template <int CAP>
__global__ __launch_bounds__(kSymWaves * 64) void k_symbolic(
    int n, const int* __restrict__ arp, const int* __restrict__ aci, int lof, uint32_t* pool, long pool_cap,
    unsigned long long* pool_top, uint64_t* header, long* row_off, int* row_cnt, int* err, int* ctr,
    uint64_t wd_limit, bool a_pivots) {
    __shared__ uint32_t s_buf[kSymWaves][2][CAP];
    __shared__ uint32_t s_new[kSymWaves][64];
    // waiting form: rows in ticket order, one per wave; rows from A: a
    // fixed grid of waves striding over the rows (no order needed)
    const int ticket = a_pivots ? 0 : take_ticket(ctr);
    const Watchdog wd(wd_limit);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long first = (a_pivots ? static_cast<long>(blockIdx.x) : static_cast<long>(ticket)) * kSymWaves + wave;
    const long stride = a_pivots ? static_cast<long>(gridDim.x) * kSymWaves : static_cast<long>(n);
    for (long ri = first; ri < n; ri += stride) {
        const int i = static_cast<int>(ri);
        wave_sync_lds();  // the previous row's last reads of the buffers
        uint32_t* cur = s_buf[wave][0];
        uint32_t* nxt = s_buf[wave][1];
        uint32_t* fresh = s_new[wave];
        const int ab = arp[i];
        const int r0 = arp[i + 1] - ab;
        bool failed = r0 > CAP;
        int r = 0;
        if (!failed) {
            // Row i of A at level 0, sorted by column (src/iluk.cpp:11-13).
            bool sorted = true;
            for (int x = lane; x + 1 < r0; x += 64) sorted &= aci[ab + x] < aci[ab + x + 1];
            sorted = __ballot(!sorted) == 0;
            if (sorted) {
                for (int x = lane; x < r0; x += 64) cur[x] = static_cast<uint32_t>(aci[ab + x]) << kLevBits;
            } else {
                bool dup = false;
                for (int x = lane; x < r0; x += 64) {
                    const int c = aci[ab + x];
                    int rank = 0;
                    for (int y = 0; y < r0; ++y) {
                        const int d = aci[ab + y];
                        rank += d < c;
                        dup |= (d == c) && (y != x);
                    }
                    if (!dup) cur[rank] = static_cast<uint32_t>(c) << kLevBits;
                }
                if (__ballot(dup) != 0) {
                    failed = true;
                    if (lane == 0) atomicOr(err, 1);
                }
            }
            r = r0;
            wave_sync_lds();
        } else if (lane == 0) {
            atomicOr(err, 2);
        }
        // Pivot walk (src/iluk.cpp:14-28).
        for (int p = 0; !failed && p < r; ++p) {
            const uint32_t pe = cur[p];
            const int k = static_cast<int>(pe >> kLevBits);
            if (k >= i) break;
            const int lk = static_cast<int>(pe & kLevMask);
            if (lk >= lof) continue;
            long uoff = 0;
            int ucnt = 0;
            if (a_pivots) {
                // A's row k (sorted): its entries right of the diagonal
                const int kb = arp[k], ke = arp[k + 1];
                int le = 0;
                for (int x0 = kb; x0 < ke; x0 += 64) {
                    const int x = x0 + lane;
                    le += __popcll(__ballot(x < ke && aci[x] <= k));
                }
                uoff = kb + le;
                ucnt = ke - kb - le;
            } else {
                uint64_t h;
                bool late = false;
                while ((h = peek_raw(header + k)) == kNoHeader) {
                    if (wd.expired()) { late = true; break; }
                    spin_pause();
                }
                if (__ballot(late) != 0) {
                    failed = true;
                    if (lane == 0) atomicOr(err, kErrTimeout);
                    break;
                }
                uoff = static_cast<long>(h >> 24);
                ucnt = static_cast<int>(h & 0xFFFFFFull);
            }
            for (int c0 = 0; c0 < ucnt && !failed; c0 += 64) {
                const int q = c0 + lane;
                bool is_new = false;
                int ins = 0;
                uint32_t cand = 0;
                bool late = false;
                if (q < ucnt) {
                    uint32_t e;
                    if (a_pivots) {
                        e = static_cast<uint32_t>(aci[uoff + q]) << kLevBits;
                    } else {
                        while ((e = peek_u32(pool + uoff + q)) == kEmpty) {
                            if (wd.expired()) { late = true; break; }
                            spin_pause();
                        }
                    }
                    const uint32_t j = e >> kLevBits;
                    const int L = lk + static_cast<int>(e & kLevMask) + 1;
                    if (L <= lof) {
                        ins = lower_col(cur, p + 1, r, j);
                        is_new = !(ins < r && (cur[ins] >> kLevBits) == j);
                        cand = (j << kLevBits) | static_cast<uint32_t>(L);
                    }
                }
                if (__ballot(late) != 0) {
                    failed = true;
                    if (lane == 0) atomicOr(err, kErrTimeout);
                    break;
                }
                const uint64_t m = __ballot(is_new);
                const int nnew = __popcll(m);
                if (nnew == 0) continue;
                if (r + nnew > CAP) {
                    failed = true;
                    if (lane == 0) atomicOr(err, 2);
                    break;
                }
                const int rank = __popcll(m & ((1ull << lane) - 1ull));
                if (is_new) fresh[rank] = cand;
                wave_sync_lds();
                // Merge: old entry x moves right by the number of new columns
                // below it; new entry of rank y lands at ins + y.
                for (int x = lane; x < r; x += 64) {
                    const uint32_t v = cur[x];
                    int shift = 0;
                    if (x > p) {
                        int lo = 0, hi = nnew;
                        const uint32_t c = v >> kLevBits;
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if ((fresh[mid] >> kLevBits) < c) lo = mid + 1; else hi = mid;
                        }
                        shift = lo;
                    }
                    nxt[x + shift] = v;
                }
                if (is_new) nxt[ins + rank] = cand;
                wave_sync_lds();
                uint32_t* t = cur;
                cur = nxt;
                nxt = t;
                r += nnew;
            }
        }
        // Publish the row (src/iluk.cpp:29,34-47 pattern part).
        long off = 0;
        if (!failed) {
            if constexpr (CAP == kSymCapShort) {
                // short rows: a fixed slot of CAP entries per row, no shared counter
                off = static_cast<long>(i) * CAP;
            } else {
                unsigned long long o = 0;
                if (lane == 0) o = atomicAdd(pool_top, static_cast<unsigned long long>(r));
                off = static_cast<long>(__shfl(static_cast<long long>(o), 0));
            }
            if (off + r > pool_cap) {
                failed = true;
                if (lane == 0) atomicOr(err, 4);
            }
        }
        if (failed) {
            r = 0;
            off = 0;
        }
        int dpos = -1;
        for (int x = lane; x < r; x += 64) {
            const uint32_t v = cur[x];
            if ((v >> kLevBits) == static_cast<uint32_t>(i)) dpos = x;
            __hip_atomic_store(pool + off + x, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // wave max of dpos
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const int d = __shfl_xor(dpos, o, 64);
            dpos = d > dpos ? d : dpos;
        }
        if (lane == 0) {
            row_off[i] = off;
            row_cnt[i] = r;
            const long uo = dpos >= 0 ? off + dpos + 1 : off;
            const long uc = dpos >= 0 ? r - dpos - 1 : 0;
            publish_raw(header + i, (static_cast<uint64_t>(uo) << 24) | static_cast<uint64_t>(uc));
        }
    }
    if (!a_pivots) retire_ticket(ctr);
}

// The rows-from-A form for short rows with G lanes per row.  With lof <= 1
// and every row of A sorted and holding its diagonal, row i of F is A's row
// i at level 0 plus, for lof = 1, the entries right of k of A's row k for
// every k < i in A's row i, at level 1: such a candidate can only ever get
// level 1, so the first-found rule has nothing to decide.  The group puts
// the candidates in LDS, drops every candidate whose column has a level-0
// or an earlier candidate, and ranks the rest by column into the row's
// pool slot.  err bit 2: a row has more than CAP candidates (the caller
// then takes the general form with long rows).
template <int G, int CAP>
__global__ __launch_bounds__(256) void k_symbolic_a(int n, const int* __restrict__ arp, const int* __restrict__ aci,
                                                    int lof, uint32_t* __restrict__ pool, long* __restrict__ row_off,
                                                    int* __restrict__ row_cnt, int* err) {
    constexpr int GROUPS = 256 / G;
    __shared__ uint32_t s_c[GROUPS][CAP];
    __shared__ unsigned char s_keep[GROUPS][CAP];
    const int grp = threadIdx.x / G, t = threadIdx.x % G;
    uint32_t* cand = s_c[grp];
    unsigned char* keep = s_keep[grp];
    for (long i0 = static_cast<long>(blockIdx.x) * GROUPS; i0 < n; i0 += static_cast<long>(gridDim.x) * GROUPS) {
        const long ii = i0 + grp;
        if (ii >= n) continue;  // the whole group (a group lies inside one wave)
        const int i = static_cast<int>(ii);
        const int ab = arp[i];
        const int r0 = arp[i + 1] - ab;
        bool over = r0 > CAP;
        int m = 0;
        if (!over) {
            for (int x = t; x < r0; x += G) cand[x] = static_cast<uint32_t>(aci[ab + x]) << kLevBits;
            m = r0;
            for (int p = 0; lof >= 1 && p < r0; ++p) {
                const int k = aci[ab + p];
                if (k >= i) break;
                const int kb = arp[k], ke = arp[k + 1];
                int le = 0;  // entries of row k up to its diagonal
                for (int x = kb + t; x < ke; x += G) le += aci[x] <= k;
#pragma unroll
                for (int o = G / 2; o > 0; o >>= 1) le += __shfl_xor(le, o, 64);
                const int ub = kb + le, uc = ke - ub;
                if (m + uc > CAP) {
                    over = true;
                    break;
                }
                for (int x = t; x < uc; x += G) cand[m + x] = (static_cast<uint32_t>(aci[ub + x]) << kLevBits) | 1u;
                m += uc;
            }
        }
        if (over) {
            if (t == 0) atomicOr(err, 2);
            continue;
        }
        wave_sync_lds();
        // a candidate stays unless its column has a level-0 candidate or an
        // earlier one of the same level
        for (int a = t; a < m; a += G) {
            const uint32_t ca = cand[a];
            bool drop = false;
            for (int b = 0; b < m; ++b) {
                const uint32_t cb = cand[b];
                drop |= b != a && (cb >> kLevBits) == (ca >> kLevBits) &&
                        ((cb & kLevMask) < (ca & kLevMask) || ((cb & kLevMask) == (ca & kLevMask) && b < a));
            }
            keep[a] = drop ? 0 : 1;
        }
        wave_sync_lds();
        const long off = static_cast<long>(i) * CAP;
        int kept = 0;
        for (int a = t; a < m; a += G) {
            if (!keep[a]) continue;
            const uint32_t ca = cand[a];
            int rank = 0;
            for (int b = 0; b < m; ++b) rank += keep[b] && (cand[b] >> kLevBits) < (ca >> kLevBits);
            pool[off + rank] = ca;
            ++kept;
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) kept += __shfl_xor(kept, o, 64);
        if (t == 0) {
            row_off[i] = off;
            row_cnt[i] = kept;
        }
        wave_sync_lds();  // the next row reuses the group's buffers
    }
}

// Lay rows out as CSR: columns, levels, values (A's values at the level-0
// slots in A's entry order, 0 at fill slots) and nzmap.
__global__ __launch_bounds__(256) void k_sym_emit(int n, const int* __restrict__ arp,
                                                  const double* __restrict__ aval, const uint32_t* __restrict__ pool,
                                                  const long* __restrict__ row_off, const int* __restrict__ frp,
                                                  int* __restrict__ fci, signed char* __restrict__ flev,
                                                  double* __restrict__ fval, int* __restrict__ nzmap, int* err) {
    const int lane = threadIdx.x & 63;
    const long i = static_cast<long>(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (i >= n) return;
    const long src = row_off[i];
    const int fb = frp[i];
    const int cnt = frp[i + 1] - fb;
    const int ab = arp[i];
    int zeros = 0;  // level-0 slots seen in earlier chunks
    for (int c0 = 0; c0 < cnt; c0 += 64) {
        const int x = c0 + lane;
        uint32_t v = 0;
        bool z = false;
        if (x < cnt) {
            v = pool[src + x];
            z = (v & kLevMask) == 0;
        }
        const uint64_t m = __ballot(z);
        if (x < cnt) {
            fci[fb + x] = static_cast<int>(v >> kLevBits);
            if (flev) flev[fb + x] = static_cast<signed char>(v & kLevMask);
            double val = 0.0;
            if (z) {
                const int t = zeros + __popcll(m & ((1ull << lane) - 1ull));
                if (aval) val = aval[ab + t];
                if (nzmap) nzmap[ab + t] = fb + x;
            }
            if (fval) fval[fb + x] = val;
        }
        zeros += __popcll(m);
    }
    if (lane == 0 && zeros != arp[i + 1] - ab) atomicOr(err, 1);
}

// bad[0] set when a row of A is not strictly increasing or lacks its
// diagonal (a row of F without its diagonal passes on no fill, which the
// rows of A alone do not tell for lof = 1).
__global__ void k_rows_sorted(int n, const int* __restrict__ rp, const int* __restrict__ ci, int* bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool ok = true, diag = false;
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
        if (p + 1 < rp[i + 1]) ok &= ci[p] < ci[p + 1];
        diag |= ci[p] == i;
    }
    if (!ok || !diag) atomicOr(bad, 1);
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) hip_fail(e, what, __FILE__, __LINE__);
}

}  // namespace

int symbolic_device(int n, const int* d_arp, const int* d_aci, const double* d_aval, int lof, SymbolicResult& out,
                    hipStream_t s, double watchdog_s) {
    if (n <= 0) return -1;
    if (n >= (1 << 27) - 1) throw std::runtime_error("IlukSymbolic: n must be below 2^27 - 1");
    if (lof < 0 || lof > 31) throw std::runtime_error("IlukSymbolic: level of fill must be in [0, 31]");
    int nnz_a = 0;
    PhaseTimer pt("symbolic pass", s);
    check(hipMemcpyAsync(&nnz_a, d_arp + n, sizeof(int), hipMemcpyDeviceToHost, s), "nnz");
    check(hipStreamSynchronize(s), "sync");
    DevBuf<uint64_t> header(n);
    DevBuf<long> row_off(n);
    DevBuf<int> row_cnt(static_cast<size_t>(n) + 1);
    DevBuf<int> misc(8);  // [0..1] ticket counters, [2] err, [3] unsorted row of A
    DevBuf<unsigned long long> top(1);
    long cap = std::max<long>(static_cast<long>(nnz_a) * (lof == 0 ? 1 : 32), 1024);
    DevBuf<uint32_t> pool;
    // first try the short-row form (fixed pool slots, n x 64 entries: only
    // while that stays within 16 GiB); a row over its capacity retries the
    // long form
    bool short_rows = static_cast<long>(n) * kSymCapShort <= (1L << 32);
    // lof <= 1 over sorted rows of A that hold their diagonals: pivot rows
    // come from A (no waits)
    bool a_pivots = false;
    if (lof <= 1) {
        check(hipMemsetAsync(misc.get() + 3, 0, sizeof(int), s), "sorted flag");
        k_rows_sorted<<<ceil_div(n, 256), 256, 0, s>>>(n, d_arp, d_aci, misc.get() + 3);
        check(hipGetLastError(), "k_rows_sorted");
        int bad = 0;
        check(hipMemcpyAsync(&bad, misc.get() + 3, sizeof(int), hipMemcpyDeviceToHost, s), "sorted flag");
        check(hipStreamSynchronize(s), "sync");
        a_pivots = bad == 0;
    }
    for (int attempt = 0;; ++attempt) {
        if (short_rows) cap = std::max<long>(cap, static_cast<long>(n) * kSymCapShort);
        pool.resize(cap);
        if (!a_pivots) {  // the waiting form polls the pool and the headers
            check(hipMemsetAsync(pool.get(), 0xFF, sizeof(uint32_t) * cap, s), "pool fill");
            check(hipMemsetAsync(header.get(), 0xFF, sizeof(uint64_t) * n, s), "header fill");
        }
        check(hipMemsetAsync(misc.get(), 0, 8 * sizeof(int), s), "misc");
        check(hipMemsetAsync(top.get(), 0, sizeof(unsigned long long), s), "top");
        const int grid = a_pivots ? std::min(ceil_div(n, kSymWaves), kSymGridA) : ceil_div(n, kSymWaves);
        if (a_pivots && short_rows)
            k_symbolic_a<16, kSymCapShort><<<std::min(ceil_div(n, 16), 8192), 256, 0, s>>>(
                n, d_arp, d_aci, lof, pool.get(), row_off.get(), row_cnt.get(), misc.get() + 2);
        else if (short_rows)
            k_symbolic<kSymCapShort><<<grid, kSymWaves * 64, 0, s>>>(
                n, d_arp, d_aci, lof, pool.get(), cap, top.get(), header.get(), row_off.get(), row_cnt.get(),
                misc.get() + 2, misc.get(), watchdog_ticks(watchdog_s), a_pivots);
        else
            k_symbolic<kSymCapLong><<<grid, kSymWaves * 64, 0, s>>>(
                n, d_arp, d_aci, lof, pool.get(), cap, top.get(), header.get(), row_off.get(), row_cnt.get(),
                misc.get() + 2, misc.get(), watchdog_ticks(watchdog_s), a_pivots);
        check(hipGetLastError(), "k_symbolic");
        int err = 0;
        check(hipMemcpyAsync(&err, misc.get() + 2, sizeof(int), hipMemcpyDeviceToHost, s), "err");
        check(hipStreamSynchronize(s), "sync");
        if (err & kErrTimeout) throw std::runtime_error("IlukSymbolic: dependency wait timed out (watchdog)");
        if (err & 1) return -1;  // repeated column: the reference exits here
        if (err & 2) {
            if (!short_rows) throw std::runtime_error("IlukSymbolic: a row of F exceeds 2048 entries");
            short_rows = false;
            continue;
        }
        if (err & 4) {
            if (attempt >= 6) throw std::runtime_error("IlukSymbolic: fill pool exhausted");
            cap *= 4;
            continue;
        }
        break;
    }
    pt.mark("pool fill and k_symbolic");
    // row_ptr = exclusive scan of the row counts.
    out.n = n;
    out.rp.resize(static_cast<size_t>(n) + 1);
    check(hipMemsetAsync(row_cnt.get() + n, 0, sizeof(int), s), "cnt tail");
    size_t bytes = 0;
    check(rocprim::exclusive_scan(nullptr, bytes, row_cnt.get(), out.rp.get(), 0, static_cast<size_t>(n) + 1,
                                  rocprim::plus<int>(), s),
          "scan size");
    DevBuf<char> tmp(std::max<size_t>(bytes, 1));
    check(rocprim::exclusive_scan(tmp.get(), bytes, row_cnt.get(), out.rp.get(), 0, static_cast<size_t>(n) + 1,
                                  rocprim::plus<int>(), s),
          "scan");
    int nnz_f = 0;
    check(hipMemcpyAsync(&nnz_f, out.rp.get() + n, sizeof(int), hipMemcpyDeviceToHost, s), "nnzf");
    check(hipStreamSynchronize(s), "sync");
    out.nnz = nnz_f;
    out.ci.resize(std::max(nnz_f, 1));
    out.lev.resize(std::max(nnz_f, 1));
    out.val.resize(std::max(nnz_f, 1));
    out.nzmap.resize(std::max(nnz_a, 1));
    pt.mark("scan and allocation of F");
    check(hipMemsetAsync(misc.get() + 2, 0, sizeof(int), s), "err reset");
    k_sym_emit<<<ceil_div(n, 4), 256, 0, s>>>(n, d_arp, d_aval, pool.get(), row_off.get(), out.rp.get(), out.ci.get(),
                                              out.lev.get(), d_aval ? out.val.get() : nullptr, out.nzmap.get(),
                                              misc.get() + 2);
    check(hipGetLastError(), "k_sym_emit");
    int err = 0;
    check(hipMemcpyAsync(&err, misc.get() + 2, sizeof(int), hipMemcpyDeviceToHost, s), "err");
    check(hipStreamSynchronize(s), "sync");
    if (err) return -1;
    pt.mark("k_sym_emit");
    return nnz_f;
}

}  // namespace pilu

extern "C" int IlukSymbolic(const CsrMatrix* a, int lof, CsrMatrix* f, int* nzmap) {
    try {
        pilu::SetupScope scope(nullptr);
        const int n = a->size;
        if (n <= 0) return -1;
        const int nnz_a = a->row_ptr[n];
        pilu::DevBuf<int> arp(static_cast<size_t>(n) + 1), aci(std::max(nnz_a, 1));
        pilu::DevBuf<double> aval(std::max(nnz_a, 1));
        PILU_HIP(hipMemcpy(arp.get(), a->row_ptr, sizeof(int) * (static_cast<size_t>(n) + 1), hipMemcpyHostToDevice));
        PILU_HIP(hipMemcpy(aci.get(), a->col_idx, sizeof(int) * static_cast<size_t>(nnz_a), hipMemcpyHostToDevice));
        if (a->value)
            PILU_HIP(hipMemcpy(aval.get(), a->value, sizeof(double) * static_cast<size_t>(nnz_a), hipMemcpyHostToDevice));
        pilu::SymbolicResult res;
        const int nnz_f = pilu::symbolic_device(n, arp.get(), aci.get(), a->value ? aval.get() : nullptr, lof, res, nullptr);
        if (nnz_f < 0) return -1;
        SetupCsrMatrix(f, n, nnz_f);
        PILU_HIP(hipMemcpy(f->row_ptr, res.rp.get(), sizeof(int) * (static_cast<size_t>(n) + 1), hipMemcpyDeviceToHost));
        if (nnz_f > 0) {
            PILU_HIP(hipMemcpy(f->col_idx, res.ci.get(), sizeof(int) * static_cast<size_t>(nnz_f), hipMemcpyDeviceToHost));
            PILU_HIP(hipMemcpy(f->value, res.val.get(), sizeof(double) * static_cast<size_t>(nnz_f), hipMemcpyDeviceToHost));
        }
        if (nzmap && nnz_a > 0)
            PILU_HIP(hipMemcpy(nzmap, res.nzmap.get(), sizeof(int) * static_cast<size_t>(nnz_a), hipMemcpyDeviceToHost));
        return nnz_f;
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "IlukSymbolic: %s\n", ex.what());
        return -1;
    }
}
