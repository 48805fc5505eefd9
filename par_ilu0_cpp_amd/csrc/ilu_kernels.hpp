// Device kernels of the ILU preconditioner: level analysis, numeric
// factorization and the L/U triangular sweeps.
//
// Scheduling.  The reference runs these loops on OpenMP threads over an
// elimination-forest partition with per-row busy-wait flags
// (src/ilu_solver.cpp:375-526, src/subtree.hpp).  Here every row is owned by
// a group of G lanes (G = 64 / rows-per-wave); rows are handed out in level
// order through per-workgroup tickets, and a row waits only for the exact
// values it reads: each published value is its own readiness flag
// (pilu::publish / pilu::peek_bits).  A wave that owns several rows keeps
// polling all of them in one loop, so rows of one wave may depend on each
// other without deadlock.
#pragma once

#include "common.hpp"

namespace pilu {

constexpr int kBlock = 256;          // threads per workgroup for ticketed kernels
constexpr int kWavesPerBlock = kBlock / 64;

// ---------------------------------------------------------------------
// Diagonal pointer (src/ilu_solver.cpp:276-289) plus a missing-diagonal
// check: dg[i] = first entry with column >= i; bad[0] counts rows whose
// first such entry is not column i.
// ---------------------------------------------------------------------
__global__ void k_diag_ptr(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                           int* __restrict__ dg, int* __restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = rp[i], e = rp[i + 1];
    int lo = b, hi = e;  // first position with ci >= i (rows are sorted)
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ci[mid] < i) lo = mid + 1; else hi = mid;
    }
    dg[i] = lo;
    if (lo >= e || ci[lo] != i) atomicAdd(bad, 1);
}

// Row pointers: rp[0] == 0, non-decreasing, rp[n] == nnz.  Must pass
// before any kernel reads col_idx through rp.
__global__ void k_check_rowptr(int n, long nnz, const int* __restrict__ rp, int* __restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const bool ok = (i == 0) ? rp[0] == 0 : (rp[i] >= rp[i - 1] && (i < n || rp[n] == nnz));
    if (!ok) atomicAdd(bad, 1);
}

// Columns in [0, n) and strictly increasing within each row.
__global__ void k_check_sorted(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                               int* __restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
        const int c = ci[p];
        if (c < 0 || c >= n || (p > rp[i] && c <= ci[p - 1])) { atomicAdd(bad, 1); return; }
    }
}

__global__ void k_fill_int(long n, int* __restrict__ p, int v) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_u64(long n, uint64_t* __restrict__ p, uint64_t v) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_iota(int n, int* __restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// ---------------------------------------------------------------------
// Level of each row in the L graph (row i needs rows of L(i)) or the U
// graph (row i needs rows of U(i) \ {i}): level = 1 + max level of its
// dependencies, 0 without any.  G lanes per row, each reading every G-th
// dependency (long rows would otherwise make every level a chain of
// dependent reads); rows are visited in natural order for L and reverse
// order for U, which is a topological order, so a row only waits on rows
// owned by earlier tickets or by its own wave.  level[] must be -1 on
// entry.  maxlev[0] receives the deepest level.  block_of (optional)
// restricts the graph to dependencies inside the row's block (local levels
// of the block-local sweeps).
// ---------------------------------------------------------------------
// The body of one direction's analysis for a workgroup taking its tickets
// from ctr; nblocks: the workgroups that share ctr.
template <bool UPPER, int G>
__device__ __forceinline__ void levels_body(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                            const int* __restrict__ dg, int* level, int* maxlev, int* ctr,
                                            int* err, uint64_t wd_limit, const int* __restrict__ block_of,
                                            int nblocks) {
    constexpr int R = 64 / G;              // rows per wave
    constexpr int ROWS = (kBlock / 64) * R;  // rows per workgroup
    const int ticket = take_ticket(ctr);
    const Watchdog wd(wd_limit);
    const int lane = threadIdx.x & 63;
    const int t = lane & (G - 1);
    const long k = static_cast<long>(ticket) * ROWS + (threadIdx.x >> 6) * R + lane / G;
    const bool active = k < n;
    const int row = active ? static_cast<int>(UPPER ? n - 1 - k : k) : 0;
    const int my_block = block_of && active ? block_of[row] : 0;
    int e = 0, end = 0;
    if (active) {
        e = (UPPER ? dg[row] + 1 : rp[row]) + t;
        end = UPPER ? rp[row + 1] : dg[row];
    }
    int lv = 0;
    int top = -1;  // highest level published by this lane
    bool done = !active;
    // One lane per row (G == 1): the row's first kLevE dependencies (columns
    // and block test) are read once.  Those solved by a lane of the same
    // wave (the previous rows in the order) are taken from that lane's
    // register by a cross-lane read, repeated while some lane of the wave
    // finishes, so a chain of rows inside a wave costs no memory round trip;
    // the others are polled together in memory, one round trip per row
    // instead of two per dependency.  Further entries go through the loop
    // below (memory).
    constexpr int kLevE = G == 1 ? 8 : 1;
    int hc[kLevE];  // dependency polled in memory, or -1
    int hl[kLevE];  // lane of this wave solving the dependency, or -1 (compacted to the front)
    bool head = G == 1;
    int pub = -1;   // this lane's level once known (read by later lanes of the wave)
    int nl_max = 0; // most in-wave dependencies of a lane of the wave
    if constexpr (G == 1) {
        const long k0 = k - lane;  // schedule position of the wave's first row
        int nl = 0;
#pragma unroll
        for (int q = 0; q < kLevE; ++q) {
            hc[q] = -1;
            hl[q] = -1;
        }
#pragma unroll
        for (int q = 0; q < kLevE; ++q) {
            if (active && e + q < end) {
                const int c = ci[e + q];
                if (!(block_of && block_of[c] != my_block)) {
                    const long kc = UPPER ? static_cast<long>(n) - 1 - c : static_cast<long>(c);
                    if (kc >= k0 && kc < k0 + 64) {
#pragma unroll
                        for (int x = 0; x < kLevE; ++x)
                            if (x == nl) hl[x] = static_cast<int>(kc - k0);
                        ++nl;
                    } else {
                        hc[q] = c;
                    }
                }
            }
        }
        e = e + kLevE < end ? e + kLevE : end;
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_xor(nl, off, 64);
            nl = o > nl ? o : nl;
        }
        nl_max = nl;
    }
    while (true) {
        if (!done) {
            if constexpr (G == 1) {
                if (head) {
                    bool ok = true;
#pragma unroll
                    for (int q = 0; q < kLevE; ++q)
                        if (hc[q] >= 0) {
                            const int d = peek_int(level + hc[q]);
                            if (d < 0) {
                                ok = false;
                            } else {
                                lv = d + 1 > lv ? d + 1 : lv;
                                hc[q] = -1;
                            }
                        }
                    head = !ok;
                }
            }
            // lane t reads entries t, t+G, ...; a pending level stops it until the next round
            while (!head && e < end) {
                const int c = ci[e];
                if (block_of && block_of[c] != my_block) { e += G; continue; }  // outside the row's block
                const int d = peek_int(level + c);
                if (d < 0) break;
                lv = d + 1 > lv ? d + 1 : lv;
                e += G;
            }
            if ((head || e < end) && wd.expired()) {
                atomicOr(err, kErrTimeout);
                e = end;
                head = false;
#pragma unroll
                for (int q = 0; q < kLevE; ++q) hl[q] = -1;
            }
        }
        if constexpr (G == 1) {
            // in-wave dependencies from the solving lanes' registers, while
            // some lane of the wave finishes (at most 64 passes)
            for (int it = 0; it < 64; ++it) {
                bool fin = false;
#pragma unroll
                for (int q = 0; q < kLevE; ++q) {
                    if (q < nl_max) {  // uniform over the wave
                        const int d = __shfl(pub, hl[q] >= 0 ? hl[q] : lane, 64);
                        if (hl[q] >= 0 && d >= 0) {
                            lv = d + 1 > lv ? d + 1 : lv;
                            hl[q] = -1;
                        }
                    }
                }
                bool inw = true;
#pragma unroll
                for (int q = 0; q < kLevE; ++q) inw = inw && hl[q] < 0;
                if (!done && !head && e >= end && inw) {
                    pub = lv;
                    publish_int(level + row, lv);
                    top = lv > top ? lv : top;
                    done = true;
                    fin = true;
                }
                if (nl_max == 0 || __ballot(fin) == 0) break;
            }
        } else {
            if (!done && group_all<G>(!head && e >= end)) {
                const int l = group_max<G>(lv);
                if (t == 0) {
                    publish_int(level + row, l);
                    top = l > top ? l : top;
                }
                done = true;
            }
        }
        if (__ballot(!done) == 0) break;
        spin_pause();
    }
    // one atomic per wave for the level count (one per row serialises on
    // the single counter word)
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_xor(top, off, 64);
        top = o > top ? o : top;
    }
    if (lane == 0 && top >= 0) atomicMax(maxlev, top);
    retire_ticket(ctr, nblocks);
}

template <bool UPPER, int G>
__global__ __launch_bounds__(kBlock) void k_levels(int n, const int* __restrict__ rp,
                                                   const int* __restrict__ ci,
                                                   const int* __restrict__ dg, int* level,
                                                   int* maxlev, int* ctr, int* err, uint64_t wd_limit,
                                                   const int* __restrict__ block_of) {
    levels_body<UPPER, G>(n, rp, ci, dg, level, maxlev, ctr, err, wd_limit, block_of, static_cast<int>(gridDim.x));
}

// Both directions in one launch, so that the two dependency chains run at
// the same time: even workgroups analyse L (tickets ctr[0..1], maxlev[0]),
// odd ones U (ctr[2..3], maxlev[1]); nb workgroups per direction.  Each
// direction's rows wait only on rows of lower tickets of the same
// direction, taken by workgroups dispatched before them.
template <int G>
__global__ __launch_bounds__(kBlock) void k_levels_lu(int n, const int* __restrict__ rp,
                                                      const int* __restrict__ ci,
                                                      const int* __restrict__ dg, int* level_l, int* level_u,
                                                      int* maxlev, int* ctr, int* err, uint64_t wd_limit, int nb) {
    if (blockIdx.x & 1)
        levels_body<true, G>(n, rp, ci, dg, level_u, maxlev + 1, ctr + 2, err, wd_limit, nullptr, nb);
    else
        levels_body<false, G>(n, rp, ci, dg, level_l, maxlev, ctr, err, wd_limit, nullptr, nb);
}

// ---------------------------------------------------------------------
// Triangular sweep.  UPPER == false: forward sweep with unit L,
//   out[i] = b[i] - sum_{j in L(i)} l_ij out[j]          (src/ilu_solver.cpp:783-788)
// UPPER == true: backward sweep,
//   out[i] = (b[i] - sum_{j in U(i), j > i} u_ij out[j]) / u_ii   (:789-795)
// `out` must hold kPending in every slot on entry.  Each row also stores
// kPending into reset[row] after reading b[row]: the forward sweep uses it
// to prepare the backward sweep's output, the backward sweep to re-arm its
// input vector for the next forward sweep.  b may alias reset.
// Lane t of a group accumulates entries t, t+G, ... in order (lane 0 starts
// from b[i]); the group total is a fixed butterfly, so results are
// deterministic, and bit-identical to the one-thread reference order for
// G == 1.
// gate (optional): the kernel does nothing when *gate == 0.
// ---------------------------------------------------------------------
template <int G, typename T, bool UPPER>
__global__ __launch_bounds__(kBlock) void k_sweep(int n, const int* __restrict__ perm,
                                                  const int* __restrict__ rp,
                                                  const int* __restrict__ ci,
                                                  const int* __restrict__ dg,
                                                  const T* __restrict__ val, const double* b,
                                                  double* out, double* reset,
                                                  const int* gate, int* ctr, int* err,
                                                  uint64_t wd_limit) {
    if (gate && *gate == 0) return;
    constexpr int R = 64 / G;
    const int ticket = take_ticket(ctr);
    const Watchdog wd(wd_limit);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int slot = lane / G;
    const int t = lane & (G - 1);
    const long pos = (static_cast<long>(ticket) * kWavesPerBlock + wave) * R + slot;
    const bool active = pos < n;
    int row = 0, e = 0, end = 0;
    double acc = 0.0, dval = 1.0;
    if (active) {
        row = perm[pos];
        const int d = dg[row];
        e = (UPPER ? d + 1 : rp[row]) + t;
        end = UPPER ? rp[row + 1] : d;
        if (UPPER) dval = static_cast<double>(val[d]);
        if (t == 0) {
            acc = b[row];
            if (reset) reinterpret_cast<uint64_t*>(reset)[row] = kPending;
        }
    }
    // The lane's first kSwH entries (t, t+G, ...): columns and values read
    // once, their x polled together; later entries one at a time below.
    // The lane sums them in the same order either way.
    constexpr int kSwH = 4;
    int hc[kSwH];
    T hv[kSwH];
    uint64_t hb[kSwH];
    int nh = 0;
#pragma unroll
    for (int q = 0; q < kSwH; ++q) {
        hc[q] = 0;
        hv[q] = T(0);
        hb[q] = kPending;
        if (active && e + q * G < end) {
            hc[q] = ci[e + q * G];
            hv[q] = val[e + q * G];
            nh = q + 1;
        }
    }
    bool head = nh > 0;
    e = e + kSwH * G < end ? e + kSwH * G : end;
    bool done = !active;
    while (true) {
        if (!done) {
            if (head) {
                bool ok = true;
#pragma unroll
                for (int q = 0; q < kSwH; ++q)
                    if (q < nh) {
                        if (hb[q] == kPending) hb[q] = peek_bits(out + hc[q]);
                        ok = ok && hb[q] != kPending;
                    }
                if (ok) {
#pragma unroll
                    for (int q = 0; q < kSwH; ++q)
                        if (q < nh) acc = fma(-__longlong_as_double(hb[q]), static_cast<double>(hv[q]), acc);
                    head = false;
                }
            }
            while (!head && e < end) {
                const uint64_t bits = peek_bits(out + ci[e]);
                if (bits == kPending) break;
                acc = fma(-__longlong_as_double(bits), static_cast<double>(val[e]), acc);
                e += G;
            }
            if ((head || e < end) && wd.expired()) {
                atomicOr(err, kErrTimeout);
                acc = __longlong_as_double(kQuietNan);
                e = end;
                head = false;
            }
        }
        const bool ready = group_all<G>(done || (!head && e >= end));
        if (ready && !done) {
            const double s = group_sum<G>(acc);
            if (t == 0) publish(out + row, UPPER ? s / dval : s);
            done = true;
        }
        if (__ballot(!done) == 0) break;
        spin_pause();
    }
    retire_ticket(ctr);
}

// ---------------------------------------------------------------------
// Block-local triangular sweep.  Rows are cut into blocks of `blk`
// consecutive rows; one workgroup owns a block and keeps the block's results
// in LDS.  Inside the block it walks the block's local levels (levels of the
// dependency graph restricted to the block), separated by workgroup
// barriers, so dependencies inside the block cost an LDS read instead of a
// device-scope round trip.  Dependencies on earlier blocks (later ones for
// the backward sweep) are value-as-flag waits on the published results;
// blocks are taken in dependency order through tickets, so every such wait
// is on a resident or finished workgroup.  Per row, lane t of a G-lane group
// accumulates entries t, t+G, ... in order and the group sum is the same
// butterfly as k_sweep: results are bit-identical to k_sweep with the same G.
// perm: rows ordered by (processing block, local level, row); groups of one
// (block, level) are perm[lev_off[g] .. lev_off[g+1]); the groups of
// processing block p are g in [blk_lev[p], blk_lev[p+1]).
// ---------------------------------------------------------------------
template <int G, typename T, bool UPPER>
__global__ __launch_bounds__(512) void k_sweep_blk(int n, int blk, int nblk, const int* __restrict__ perm,
                                                   const int* __restrict__ blk_lev, const int* __restrict__ lev_off,
                                                   const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const int* __restrict__ dg, const T* __restrict__ val,
                                                   const double* b, double* out, double* reset, const int* gate,
                                                   int* ctr, int* err, uint64_t wd_limit) {
    extern __shared__ double xs[];
    if (gate && *gate == 0) return;
    const int ticket = take_ticket(ctr);
    const Watchdog wd(wd_limit);
    const int bid = UPPER ? nblk - 1 - ticket : ticket;
    const int bs = bid * blk;
    const int be = bs + blk < n ? bs + blk : n;
    const int ng = blockDim.x / G;
    const int gid = threadIdx.x / G;
    const int t = threadIdx.x & (G - 1);
    bool timed_out = false;
    for (int g = blk_lev[ticket]; g < blk_lev[ticket + 1]; ++g) {
        const int lo = lev_off[g], hi = lev_off[g + 1];
        for (int k = lo + gid; k < hi; k += ng) {
            const int row = perm[k];
            const int d = dg[row];
            int e = (UPPER ? d + 1 : rp[row]) + t;
            const int end = UPPER ? rp[row + 1] : d;
            double acc = 0.0;
            if (t == 0) {
                acc = b[row];
                if (reset) reset[row] = __longlong_as_double(kPending);
            }
            for (; e < end; e += G) {
                const int c = ci[e];
                double v;
                if (c >= bs && c < be) {
                    v = xs[c - bs];
                } else {
                    uint64_t bits;
                    while ((bits = peek_bits(out + c)) == kPending) {
                        if (timed_out || wd.expired()) { timed_out = true; bits = kQuietNan; break; }
                        spin_pause();
                    }
                    v = __longlong_as_double(bits);
                }
                acc = fma(-v, static_cast<double>(val[e]), acc);
            }
            const double s = group_sum<G>(acc);
            if (t == 0) {
                const double r = UPPER ? s / static_cast<double>(val[d]) : s;
                xs[row - bs] = r;
                publish(out + row, r);
            }
        }
        __syncthreads();
    }
    if (timed_out) atomicOr(err, kErrTimeout);
    retire_ticket(ctr);
}

// ---------------------------------------------------------------------
// Single-workgroup level walk.  One workgroup of kLvThreads threads solves
// every row, level by level (perm sorted by level: level l is
// perm[lev_off[l] .. lev_off[l+1])), with a workgroup barrier between
// levels.  All dependencies of a row sit in earlier levels and were written
// by this workgroup before a barrier, so there are no value-flag waits and
// no hand-offs between workgroups: a level costs one barrier and one read of
// its dependencies' x, which hit lines this workgroup wrote.  It suits deep
// dependency graphs with few rows per level (random sparsity with fill),
// where the hand-off latency of the multi-workgroup sweeps dominates.
//
// Slot j of level l (its j-th row, j < kLvThreads / G) is described by
// desc[l * slots + j] = {row or -1, first entry, end, rows in level l}.  The
// descriptor is read four levels ahead and the slot's row data (the lane's first kLvH
// entries, columns and values; b; the diagonal) two levels ahead, in two
// register sets used alternately, from clamped addresses so that the issue
// sequence has no branches: loads complete in order, and a level then
// waits only for its x reads (the loop is unrolled four times, so the
// register sets are carried around the loop, which waits for the
// outstanding reads, once per four levels).  Rows past the first `slots` of
// a level are read directly.  Lane t of a G-lane group sums entries t, t+G, ... in order
// from b (lane 0) and the group total is k_sweep's butterfly, so results
// are bit-identical to k_sweep with the same G.  Like k_sweep, each row
// stores kPending into reset[row] after reading b[row].
// ---------------------------------------------------------------------
constexpr int kLvThreads = 1024;
constexpr int kLvH = 4;
// With `cache` set, the x of recently solved rows are also kept in LDS, in
// slot row % kLvCache tagged with the row (96 KB of dynamic LDS); a read
// whose slot holds another row goes to memory.  A level uses the ring only
// if no slot is both written and read, or written twice, within it
// (k_lv_slots; bit 30 of the descriptor's row count marks the others, whose
// rows are then neither read from nor written to the ring), so a slot's
// value and tag always come from the same row.
constexpr int kLvCache = 8192;
constexpr size_t kLvLds = kLvCache * (sizeof(double) + sizeof(int));

__device__ __forceinline__ double lv_x(int c, const double* out, const double* xs, const int* xt, bool cache) {
    if (cache) {
        const int sl = c & (kLvCache - 1);
        if (xt[sl] == c) return xs[sl];
    }
    return out[c];
}

__device__ __forceinline__ void lv_put(int row, double v, double* out, double* xs, int* xt, bool cache) {
    out[row] = v;
    if (cache) {
        const int sl = row & (kLvCache - 1);
        xs[sl] = v;
        xt[sl] = row;
    }
}

template <typename T>
struct LvRow {
    int row, e, end, w;  // e: this lane's first entry; w: rows in the level
    int c[kLvH];
    T v[kLvH];
    double bv;
    T dv;
};

template <int G, typename T, bool UPPER>
__device__ __forceinline__ void lv_fetch(const int4& d, int t, long last, const int* __restrict__ ci,
                                         const T* __restrict__ val, const double* b, LvRow<T>& r) {
    r.row = d.x;
    r.e = d.y + t;
    r.end = d.z;
    r.w = d.w;
    // positions past the row's end read its last entry again (a line already
    // requested), not the next rows' entries: the single workgroup's
    // requests in flight are the limit, so none are spent on unused data
    long hi = d.z > d.y ? d.z - 1 : d.y;
    hi = hi < last ? hi : last;
#pragma unroll
    for (int q = 0; q < kLvH; ++q) {
        long k = static_cast<long>(r.e) + q * G;
        k = k < hi ? k : hi;
        r.c[q] = ci[k];
        r.v[q] = val[k];
    }
    r.bv = b[d.x >= 0 ? d.x : 0];
    if (UPPER) r.dv = val[d.y > 0 ? d.y - 1 : 0];
    else r.dv = T(1);
}

// One level: solves the slot's row from `r` (fetched two levels earlier),
// refills `r` with level l + 2 from `d` and `d` with the descriptor of
// level l + 4, then the level's further rows, then the barrier.
template <int G, typename T, bool UPPER>
__device__ __forceinline__ void lv_level(int l, int nlev, LvRow<T>& r, int4& d, const int* __restrict__ lev_off,
                                         const int4* __restrict__ desc, const int* __restrict__ perm,
                                         const int* __restrict__ rp, const int* __restrict__ ci,
                                         const int* __restrict__ dg, const T* __restrict__ val, long last,
                                         const double* b, double* out, double* reset, double* xs, int* xt,
                                         bool cache) {
    constexpr int kSlots = kLvThreads / G;
    const int j = static_cast<int>(threadIdx.x) / G;
    const int t = threadIdx.x & (G - 1);
    const LvRow<T> cur = r;
    const int width = cur.w & 0x3fffffff;
    cache = cache && (cur.w >> 30) == 0;
    double x[kLvH];
#pragma unroll
    for (int q = 0; q < kLvH; ++q) x[q] = lv_x(cur.c[q], out, xs, xt, cache);
    lv_fetch<G, T, UPPER>(d, t, last, ci, val, b, r);
    {
        const int l4 = l + 4 < nlev ? l + 4 : nlev - 1;
        d = desc[static_cast<long>(l4) * kSlots + j];
    }
    double acc = t == 0 ? cur.bv : 0.0;
#pragma unroll
    for (int q = 0; q < kLvH; ++q)
        if (cur.e + q * G < cur.end) acc = fma(-x[q], static_cast<double>(cur.v[q]), acc);
    for (int e = cur.e + kLvH * G; e < cur.end; e += G)
        acc = fma(-lv_x(ci[e], out, xs, xt, cache), static_cast<double>(val[e]), acc);
    const double s = group_sum<G>(acc);
    if (t == 0 && cur.row >= 0) {
        lv_put(cur.row, UPPER ? s / static_cast<double>(cur.dv) : s, out, xs, xt, cache);
        if (reset) reinterpret_cast<uint64_t*>(reset)[cur.row] = kPending;
    }
    if (width > kSlots) {
        const int p0 = lev_off[l], p1 = p0 + width;
        for (int p = p0 + kSlots + j; p < p1; p += kSlots) {
            const int row = perm[p];
            const int dd = dg[row];
            const int e0 = UPPER ? dd + 1 : rp[row], end = UPPER ? rp[row + 1] : dd;
            double a = t == 0 ? b[row] : 0.0;
            for (int e = e0 + t; e < end; e += G)
                a = fma(-lv_x(ci[e], out, xs, xt, cache), static_cast<double>(val[e]), a);
            const double sm = group_sum<G>(a);
            if (t == 0) {
                lv_put(row, UPPER ? sm / static_cast<double>(val[dd]) : sm, out, xs, xt, cache);
                if (reset) reinterpret_cast<uint64_t*>(reset)[row] = kPending;
            }
        }
    }
    __syncthreads();
}

// L2 warming for k_sweep_lv.  The solving workgroup's reads of the factors
// are limited by the requests one compute unit keeps in flight, each
// waiting a full memory latency.  Helper workgroups on the solver's XCD
// (workgroups are dealt to the 8 XCDs round-robin, so workgroups 8, 16, ...
// share workgroup 0's L2) read the lines of the levels kLvAheadMin to
// kLvAheadMax ahead of the solver's progress (published every four
// levels), helper k taking every helpers-th level; the solver then finds
// them in L2.  They read only, and write `sink` only if a checksum takes
// an impossible value (keeping the reads), so results do not depend on
// them; each leaves when the solver is done, when the next level is past
// the last, or when its watchdog expires.  `epoch` tells this launch's
// progress word from an earlier launch's.
constexpr int kLvHelpers = 7;
constexpr int kLvAheadMin = 4;
constexpr int kLvAheadMax = 64;

template <int G, bool UPPER, typename T>
__device__ void lv_helper(int k, int nlev, const int* __restrict__ lev_off, const int4* __restrict__ desc,
                          const int* __restrict__ perm, const int* __restrict__ rp, const int* __restrict__ dg,
                          const int* __restrict__ ci, const T* __restrict__ val, const double* b,
                          const unsigned long long* prog, unsigned epoch, int* sink, uint64_t wd_limit) {
    constexpr int kSlots = kLvThreads / G;
    __shared__ int s_lvl;
    const Watchdog wd(wd_limit);
    const int tid = threadIdx.x;
    int next = k;
    int acc = 0;
    double accd = 0.0;
    while (next < nlev) {
        if (tid == 0) {
            const unsigned long long p = __hip_atomic_load(prog, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_lvl = static_cast<unsigned>(p >> 32) == epoch ? static_cast<int>(p & 0xffffffffu) : 0;
            if (wd.expired()) s_lvl = nlev;
        }
        __syncthreads();
        const int lvl = s_lvl;
        __syncthreads();
        if (lvl >= nlev) break;
        while (next < lvl + kLvAheadMin) next += kLvHelpers;
        if (next >= nlev) break;
        if (next >= lvl + kLvAheadMax) {
            spin_pause();
            continue;
        }
        // the level's descriptors, and per row one read per 64 bytes of its
        // columns and values (16 lanes per row, rows 64 at a time)
        if (tid < kSlots) acc ^= desc[static_cast<long>(next) * kSlots + tid].x;
        const int p0 = lev_off[next], p1 = lev_off[next + 1];
        for (int p = p0 + tid / 16; p < p1; p += kLvThreads / 16) {
            const int row = perm[p];
            const int d = dg[row];
            const int e0 = UPPER ? d : rp[row], e1 = UPPER ? rp[row + 1] : d;
            if ((tid & 15) == 0) accd += b[row];
            for (int e = e0 + (tid & 15) * 16; e < e1; e += 256) {
                acc ^= ci[e];
                accd += static_cast<double>(val[e]);
                if (sizeof(T) == 8 && e + 8 < e1) accd += static_cast<double>(val[e + 8]);
            }
        }
        next += kLvHelpers;
    }
    if (acc == 0x7fffffff && accd == -1.0) sink[0] = 1;
}

template <int G, typename T, bool UPPER>
__global__ __launch_bounds__(kLvThreads) void k_sweep_lv(int nlev, const int* __restrict__ lev_off,
                                                         const int4* __restrict__ desc, const int* __restrict__ perm,
                                                         const int* __restrict__ rp, const int* __restrict__ ci,
                                                         const int* __restrict__ dg, const T* __restrict__ val,
                                                         long nnz, const double* b, double* out, double* reset,
                                                         const int* gate, unsigned long long* prog, unsigned epoch,
                                                         int* sink, uint64_t wd_limit, int cache) {
    extern __shared__ double lv_lds[];
    if (gate && *gate == 0) return;
    if (nlev <= 0 || nnz <= 0) return;
    if (blockIdx.x != 0) {
        if (blockIdx.x % 8 == 0 && static_cast<int>(blockIdx.x / 8) <= kLvHelpers)
            lv_helper<G, UPPER, T>(static_cast<int>(blockIdx.x / 8) - 1, nlev, lev_off, desc, perm, rp, dg, ci, val,
                                   b, prog, epoch, sink, wd_limit);
        return;
    }
    constexpr int kSlots = kLvThreads / G;
    const int j = static_cast<int>(threadIdx.x) / G;
    const int t = threadIdx.x & (G - 1);
    const long last = nnz - 1;
    double* xs = lv_lds;
    int* xt = reinterpret_cast<int*>(lv_lds + kLvCache);
    const bool use = cache != 0;
    if (use) {
        for (int i = threadIdx.x; i < kLvCache; i += kLvThreads) xt[i] = -1;
        __syncthreads();
    }
    auto dsc = [&](int l) { return desc[static_cast<long>(l < nlev ? l : nlev - 1) * kSlots + j]; };
    LvRow<T> r0, r1;
    lv_fetch<G, T, UPPER>(dsc(0), t, last, ci, val, b, r0);
    lv_fetch<G, T, UPPER>(dsc(1), t, last, ci, val, b, r1);
    int4 d0 = dsc(2), d1 = dsc(3);
    const unsigned long long ep = static_cast<unsigned long long>(epoch) << 32;
    for (int l = 0; l < nlev; l += 4) {
        if (threadIdx.x == 0) __hip_atomic_store(prog, ep | static_cast<unsigned>(l), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        lv_level<G, T, UPPER>(l, nlev, r0, d0, lev_off, desc, perm, rp, ci, dg, val, last, b, out, reset, xs, xt, use);
        if (l + 1 >= nlev) break;
        lv_level<G, T, UPPER>(l + 1, nlev, r1, d1, lev_off, desc, perm, rp, ci, dg, val, last, b, out, reset, xs, xt, use);
        if (l + 2 >= nlev) break;
        lv_level<G, T, UPPER>(l + 2, nlev, r0, d0, lev_off, desc, perm, rp, ci, dg, val, last, b, out, reset, xs, xt, use);
        if (l + 3 >= nlev) break;
        lv_level<G, T, UPPER>(l + 3, nlev, r1, d1, lev_off, desc, perm, rp, ci, dg, val, last, b, out, reset, xs, xt, use);
    }
    if (threadIdx.x == 0) __hip_atomic_store(prog, ep | static_cast<unsigned>(nlev), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// desc[l * slots + j] of k_sweep_lv: the j-th row of level l, its first
// entry and end in the sweep's direction (row -1, no entries, past the
// level's last row), and the level's row count, with bit 30 set when the
// level may not use the LDS cache (bad[l]).
template <bool UPPER>
__global__ void k_lv_desc(int nlev, int slots, const int* __restrict__ lev_off, const int* __restrict__ perm,
                          const int* __restrict__ rp, const int* __restrict__ dg, const int* __restrict__ bad,
                          int4* __restrict__ desc) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<long>(nlev) * slots) return;
    const int l = static_cast<int>(i / slots), j = static_cast<int>(i % slots);
    const int p = lev_off[l] + j;
    int4 d = make_int4(-1, 0, 0, (lev_off[l + 1] - lev_off[l]) | (bad[l] ? 1 << 30 : 0));
    if (p < lev_off[l + 1]) {
        const int row = perm[p];
        d.x = row;
        d.y = UPPER ? dg[row] + 1 : rp[row];
        d.z = UPPER ? rp[row + 1] : dg[row];
    }
    desc[i] = d;
}

// lev_bad[l] = 1 unless the rows of level l have distinct k_sweep_lv cache
// slots and none is the slot of a dependency of a row of the level (the
// slots written during such a level are exactly its rows', none of them
// read during it, so a read never meets a half-written slot); nbad[0]
// counts the others.  One workgroup per level (grid = nlev) marks its rows'
// slots in an LDS bitmap, then tests every dependency's slot against it.
template <bool UPPER>
__global__ __launch_bounds__(256) void k_lv_slots(const int* __restrict__ lev_off, const int* __restrict__ perm,
                                                  const int* __restrict__ rp, const int* __restrict__ dg,
                                                  const int* __restrict__ ci, int* __restrict__ lev_bad, int* nbad) {
    constexpr int M = kLvCache - 1;
    __shared__ unsigned bits[kLvCache / 32];
    const int l = blockIdx.x;
    for (int i = threadIdx.x; i < kLvCache / 32; i += 256) bits[i] = 0u;
    __syncthreads();
    const int p0 = lev_off[l], p1 = lev_off[l + 1];
    bool bad = false;
    for (int p = p0 + static_cast<int>(threadIdx.x); p < p1; p += 256) {
        const int sl = perm[p] & M;
        const unsigned bit = 1u << (sl & 31);
        if (atomicOr(&bits[sl >> 5], bit) & bit) bad = true;
    }
    __syncthreads();
    for (int p = p0 + static_cast<int>(threadIdx.x); p < p1 && !bad; p += 256) {
        const int row = perm[p];
        const int e0 = UPPER ? dg[row] + 1 : rp[row], e1 = UPPER ? rp[row + 1] : dg[row];
        for (int e = e0; e < e1; ++e) {
            const int sc = ci[e] & M;
            if ((bits[sc >> 5] >> (sc & 31)) & 1u) {
                bad = true;
                break;
            }
        }
    }
    const bool any = __syncthreads_or(bad ? 1 : 0) != 0;
    if (threadIdx.x == 0) {
        lev_bad[l] = any ? 1 : 0;
        if (any) atomicAdd(nbad, 1);
    }
}

// off[l] = first position of level l in a level-sorted order (every level
// from 0 to nlev - 1 holds at least one row); off[nlev] = n.
__global__ void k_level_starts(int n, int nlev, const int* __restrict__ lev_sorted, int* __restrict__ off) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (k == 0) off[0] = 0;
    else if (lev_sorted[k] != lev_sorted[k - 1]) off[lev_sorted[k]] = k;
    if (k == n - 1) off[nlev] = n;
}

// ---------------------------------------------------------------------
// Packed block-local sweep.  The rows are split into blocks (contiguous row
// ranges, or any partition whose block dependency graph is acyclic, such
// as 3D tiles of a grid); each workgroup takes one block in a topological
// order of that graph and solves it level by level, the block's results
// held in LDS.  The factors are stored in the sweep's schedule order: for
// every (block, local level) group the entries of its rows form a
// column-major ELL slab (entry q of the group's j-th row at
// base + q * count + j).  A column inside the row's own block is stored as
// -(1 + its schedule position within the block), its LDS slot; any other
// column as itself, read from the output vector as a value-flag.
//
// Everything a pass needs is requested before it runs:
//   * at block start the right-hand side of the block's rows is staged in
//     LDS (xs[] then holds b_i until row i is solved and x_i afterwards;
//     the companion buffer is re-armed for those rows at the same time),
//     and so are the block's compact pass descriptors (PassDesc);
//   * the loads of pass t+DA (row info, column indices) and of pass t+DB
//     (values, diagonal, and a first read of every out-of-block x_c) are
//     issued right after pass t computes, in the order they will be
//     consumed, so the in-order vector-memory counter never makes a pass
//     wait for a younger request; addresses are clamped into the pack so
//     that the issue sequence has no branches;
//   * an out-of-block x_c that was still pending when read early is polled
//     again when its pass runs.
// Workgroup barriers, LDS-only, sit at level boundaries.  Every row sum
// follows k_sweep's lane order, so results are bit-identical to it for
// the same lanes per row G.
// ---------------------------------------------------------------------

// Passes per unrolled step of k_sweep_pp; every block's descriptor list is
// padded to a multiple of it with empty passes (cb = 0: no rows).
constexpr int kPpUnroll = 16;
constexpr int kPpThreads = 512;  // threads per workgroup of the packed sweep

struct PassDesc {
    int k_lo;   // schedule position of the pass's first row
    int base;   // slab offset of the level group
    int cb;     // bits 0..30: rows in the group (slab column stride); bit 31: level boundary follows
    int jw;     // bits 0..15: index of the pass's first row within its group; bits 16..31: group width
};

__device__ __forceinline__ int pd_count(const PassDesc& d) { return d.cb & 0x7fffffff; }
__device__ __forceinline__ int pd_j0(const PassDesc& d) { return d.jw & 0xffff; }
__device__ __forceinline__ int pd_width(const PassDesc& d) { return d.jw >> 16; }

// Per-lane data of one pass, requested DA passes ahead.
template <int S>
struct PpA {
    int row, cnt, kk;  // row, its entry count, schedule position (clamped into the pass)
    bool act;
    int c[S];
};

// Values of one pass, requested DB passes ahead.
template <int S, typename T>
struct PpB {
    T v[S];
    T dv;
    double bv;  // right-hand side of the row (flag-synchronised form only)
};

template <int G, int S>
__device__ __forceinline__ void pp_issue_a(const PassDesc& d, int gid, int t, int per,
                                           const int2* __restrict__ rowinfo, const int* __restrict__ pcol,
                                           PpA<S>& a) {
    const int count = pd_count(d), j0 = pd_j0(d), width = pd_width(d);
    const int nact = count - j0 < per ? count - j0 : per;
    a.act = gid < nact;
    const int off = a.act ? gid : 0;
    a.kk = d.k_lo + off;
    const int2 ri = rowinfo[a.kk];
    a.row = ri.x;
    a.cnt = ri.y;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int q = t + s * G;
        a.c[s] = pcol[q < width ? d.base + q * count + j0 + off : 0];
    }
}

template <int G, int S, typename T, bool UPPER, bool FLAGS>
__device__ __forceinline__ void pp_issue_b(const PassDesc& d, int t, const PpA<S>& a, const T* __restrict__ pval,
                                           const T* __restrict__ pdiag, const double* rhs, PpB<S, T>& b) {
    const int count = pd_count(d), width = pd_width(d);
    const int jj = pd_j0(d) + (a.kk - d.k_lo);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int q = t + s * G;
        b.v[s] = pval[q < width ? d.base + q * count + jj : 0];
    }
    if (UPPER) b.dv = pdiag[a.kk];
    if (FLAGS) b.bv = rhs[a.row];
}

// Early reads of the out-of-block x of one pass, requested DX passes ahead.
template <int G, int S>
__device__ __forceinline__ void pp_issue_x(int t, int idle, const PpA<S>& a, const double* out, uint64_t* xg) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int q = t + s * G;
        const int c = a.c[s];
        // lanes with nothing to read early all read one slot of the block
        // (a single request per wave)
        const bool need = a.act && q < a.cnt && c >= 0;
        int idx = need ? c : idle;
        asm volatile("" : "+v"(idx));  // keep the read unconditional (a select, not a branch)
        xg[s] = peek_bits(out + idx);
    }
}

// kptr[t]: schedule position of the first row of the t-th block processed;
// pass_ptr[t]: its first PassDesc (each list padded to a multiple of DA).
// Per pass t the requests are issued in the order they are consumed: the
// early x reads of pass t+DX, the values of pass t+DB, the row data of pass
// t+DA (DX < DB < DA).  The first two only depend on row data requested
// DA-DB or more passes earlier, so no issue waits on a young request.
template <int G, int S, int DA, int DB, int DX, typename T, bool UPPER, bool FLAGS>
__global__ __launch_bounds__(kPpThreads) void k_sweep_pp(const int* __restrict__ kptr, const int* __restrict__ pass_ptr,
                                                  const PassDesc* __restrict__ desc,
                                                  const int2* __restrict__ rowinfo, const int* __restrict__ pcol,
                                                  const T* __restrict__ pval, const T* __restrict__ pdiag,
                                                  const double* b, double* out, double* reset, const int* gate,
                                                  int* ctr, int* err, uint64_t wd_limit, int max_rows) {
    static_assert(DA % DB == 0 && DA % DX == 0 && DX < DB && DB < DA && DA == kPpUnroll, "ring sizes");
    extern __shared__ double xs[];
    if (gate && *gate == 0) return;
    const int ticket = take_ticket(ctr);
    const Watchdog wd(wd_limit);
    const int kb0 = kptr[ticket], nrows = kptr[ticket + 1] - kb0;
    PassDesc* sd = reinterpret_cast<PassDesc*>(xs + max_rows);
    const int p0 = pass_ptr[ticket], np = pass_ptr[ticket + 1] - p0;
    // b may alias the companion buffer (in-place apply): the barrier form
    // reads each b_i into LDS before re-arming it; the flag form reads b_i
    // through the prefetch ring and re-arms row i once it has used it.
    for (int i = threadIdx.x; i < nrows; i += kPpThreads) {
        const int row = rowinfo[kb0 + i].x;
        if (FLAGS) {
            xs[i] = __longlong_as_double(kPending);
        } else {
            xs[i] = b[row];
            if (reset) reset[row] = __longlong_as_double(kPending);
        }
    }
    for (int i = threadIdx.x; i < np; i += kPpThreads) sd[i] = desc[p0 + i];
    __syncthreads();
    if (np == 0) {
        retire_ticket(ctr);
        return;
    }
    const int idle = rowinfo[kb0].x;
    constexpr int per = kPpThreads / G;
    const int gid = threadIdx.x / G;
    const int t = threadIdx.x & (G - 1);
    bool timed_out = false;
    PpA<S> ra[DA];
    PpB<S, T> rb[DB];
    uint64_t rx[DX][S];
    auto pass_desc = [&](int q) -> PassDesc { return sd[q < np ? q : np - 1]; };
#pragma unroll
    for (int u = 0; u < DA; ++u) pp_issue_a<G, S>(pass_desc(u), gid, t, per, rowinfo, pcol, ra[u]);
#pragma unroll
    for (int u = 0; u < DX; ++u) pp_issue_x<G, S>(t, idle, ra[u], out, rx[u]);
#pragma unroll
    for (int u = 0; u < DB; ++u) pp_issue_b<G, S, T, UPPER, FLAGS>(pass_desc(u), t, ra[u], pval, pdiag, b, rb[u]);
    // np is a multiple of DA: the unrolled body below has no pass guard, so
    // the issue sequence (and the compiler's count of outstanding requests)
    // is the same on every path.
    for (int p = 0; p < np; p += DA) {
#pragma unroll
        for (int u = 0; u < DA; ++u) {
            const int q = p + u;
            // descriptors used by this pass and its issues, read now so that
            // the LDS latency overlaps the row work
            const PassDesc d = sd[q];
            const PassDesc db = pass_desc(q + DB);
            const PassDesc da = pass_desc(q + DA);
            PpA<S>& a = ra[u];
            PpB<S, T>& r = rb[u % DB];
            uint64_t* xg = rx[u % DX];
            if (a.act) {
                const int slot = a.kk - kb0;
                double acc = t == 0 ? (FLAGS ? r.bv : xs[slot]) : 0.0;
                if (FLAGS && t == 0 && reset) reset[a.row] = __longlong_as_double(kPending);
                // x of an in-block row: LDS (a value-flag in the flag-synchronised form)
                auto local_x = [&](int c) -> double {
                    if (!FLAGS) return xs[-1 - c];
                    uint64_t bits;
                    while ((bits = __hip_atomic_load(reinterpret_cast<uint64_t*>(xs - 1 - c), __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP)) == kPending) {
                        if (timed_out || wd.expired()) { timed_out = true; bits = kQuietNan; break; }
                        spin_pause();
                    }
                    return __longlong_as_double(bits);
                };
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (t + s * G < a.cnt) {
                        const int c = a.c[s];
                        double x;
                        if (c < 0) {
                            x = local_x(c);
                        } else {
                            uint64_t bits = xg[s];
                            while (bits == kPending) {
                                if (timed_out || wd.expired()) { timed_out = true; bits = kQuietNan; break; }
                                spin_pause();
                                bits = peek_bits(out + c);
                            }
                            x = __longlong_as_double(bits);
                        }
                        acc = fma(-x, static_cast<double>(r.v[s]), acc);
                    }
                }
                const int count = pd_count(d);
                const int jj = pd_j0(d) + (a.kk - d.k_lo);
                for (int qq = t + S * G; qq < a.cnt; qq += G) {
                    const int pos = d.base + qq * count + jj;
                    const int c = pcol[pos];
                    double x;
                    if (c < 0) {
                        x = local_x(c);
                    } else {
                        uint64_t bits;
                        while ((bits = peek_bits(out + c)) == kPending) {
                            if (timed_out || wd.expired()) { timed_out = true; bits = kQuietNan; break; }
                            spin_pause();
                        }
                        x = __longlong_as_double(bits);
                    }
                    acc = fma(-x, static_cast<double>(pval[pos]), acc);
                }
                const double sum = group_sum<G>(acc);
                if (t == 0) {
                    const double res = UPPER ? sum / static_cast<double>(r.dv) : sum;
                    if (FLAGS) {
                        uint64_t rb_ = __double_as_longlong(res);
                        if (res != res) rb_ = kQuietNan;
                        __hip_atomic_store(reinterpret_cast<uint64_t*>(xs + slot), rb_, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    } else {
                        xs[slot] = res;
                    }
                    publish(out + a.row, res);
                }
            }
            if (!FLAGS && d.cb < 0) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
                __builtin_amdgcn_s_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
            }
            // requests in consumption order: x of pass q+DX, values of pass
            // q+DB, row data of pass q+DA (into the slot pass q just freed)
            pp_issue_x<G, S>(t, idle, ra[(u + DX) % DA], out, rx[u % DX]);
            pp_issue_b<G, S, T, UPPER, FLAGS>(db, t, ra[(u + DB) % DA], pval, pdiag, b, rb[u % DB]);
            pp_issue_a<G, S>(da, gid, t, per, rowinfo, pcol, ra[u]);
        }
    }
    if (timed_out) atomicOr(err, kErrTimeout);
    retire_ticket(ctr);
}

// Values of the packed sweep from the factors: pval[e] = fac[src[e]] (src
// < 0: padding), pdiag[k] = fac[dg[row of k]].
template <typename TI, typename TO>
__global__ void k_pack_gather(long ne, const int* __restrict__ src, const TI* __restrict__ fac, TO* __restrict__ pval,
                              int n, const int2* __restrict__ rowinfo, const int* __restrict__ dg,
                              TO* __restrict__ pdiag) {
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    for (long e = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; e < ne; e += stride) {
        const int s = src[e];
        pval[e] = s >= 0 ? static_cast<TO>(fac[s]) : TO(0);
    }
    if (pdiag)
        for (long k = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; k < n; k += stride)
            pdiag[k] = static_cast<TO>(fac[dg[rowinfo[k].x]]);
}

// Check of every row's final pivot u_ii before the row is published, as
// MKL's dcsrilu0 documents it for its ipar[30], dpar[30] and dpar[31] (the
// reference's USE_MKL_ILU build, src/ilu_solver.cpp:248-259,631-639):
// mode 0 no check (the reference's default build); mode 1 (ipar[30] = 0)
// a pivot that is zero or smaller in magnitude than `tiny` fails the
// factorization, and the smallest such row is recorded in *first as 2 row
// (zero) or 2 row + 1 (small); mode 2 (ipar[30] != 0) such a pivot is
// replaced by `repl` and the factorization continues.
struct PivotGuard {
    int mode;
    double tiny;
    double repl;
    int* first;
};

// Applies the guard to row `row`'s pivot in LDS (lane t == 0 of the group).
__device__ __forceinline__ void guard_pivot(const PivotGuard& pg, double* piv, int row, int t) {
    if (pg.mode != 0 && t == 0) {
        const double d = *piv;
        if (d == 0.0 || fabs(d) < pg.tiny) {
            if (pg.mode == 1) atomicMin(pg.first, 2 * row + (d == 0.0 ? 0 : 1));
            else *piv = pg.repl;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ---------------------------------------------------------------------
// Numeric ILU, row IKJ form of left_looking_col (src/ilu_solver.cpp:553-589):
//   for pivots k of row i in ascending column order:
//     a_ik /= u_kk;  a_ij -= u_kj * a_ik  for j in U(k), j > k, j in row i
// Updates to a row entry arrive in ascending-k order, exactly as in the
// reference; each uses one fma, the contraction the reference's
// -ffast-math -march=native build applies.  Row i's values are staged in
// LDS (capacity CAP per group); entries of U(k) are matched into row i by
// binary search over its sorted columns.  vin holds the values to factor,
// vout (all kPending on entry) receives the factors, published entry by
// entry as value-flags; f32 (optional) receives a float copy.
// err bit 0: a zero pivot was divided by; bit 1: a row exceeded CAP (its
// factors are published as NaN so that dependent rows still finish).
// omega (optional): per-row diagonal compensation.  Updates u_kj * a_ik that
// fall outside row i's pattern are summed (lane-strided, then a fixed
// butterfly) and omega[i] times that sum is added to a_ii before the row is
// published -- modified ILU with a per-row relaxation factor.  pg: the
// pivot check above, after the compensation.
// ---------------------------------------------------------------------
template <int G, int CAP>
__global__ __launch_bounds__(kBlock) void k_factor(int n, const int* __restrict__ perm,
                                                   const int* __restrict__ rp,
                                                   const int* __restrict__ ci,
                                                   const int* __restrict__ dg,
                                                   const double* __restrict__ vin, double* vout,
                                                   float* f32, int* err, int* ctr, uint64_t wd_limit,
                                                   const float* __restrict__ omega, const PivotGuard pg) {
    constexpr int R = 64 / G;
    constexpr int GROUPS = kBlock / G;
    __shared__ int s_col[GROUPS][CAP];
    __shared__ double s_val[GROUPS][CAP];
    const int ticket = take_ticket(ctr);
    const Watchdog wd(wd_limit);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int slot = lane / G;
    const int grp = wave * R + slot;
    const int t = lane & (G - 1);
    const long pos = static_cast<long>(ticket) * GROUPS + grp;
    const bool active = pos < n;
    int* col = s_col[grp];
    double* val = s_val[grp];
    int row = 0, rb = 0, len = 0, ndiag = 0;
    bool overflow = false;
    if (active) {
        row = perm[pos];
        rb = rp[row];
        len = rp[row + 1] - rb;
        ndiag = dg[row] - rb;
        overflow = len > CAP;
        if (!overflow) {
            for (int q = t; q < len; q += G) {
                col[q] = ci[rb + q];
                val[q] = vin[rb + q];
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    int p = 0;  // next pivot (uniform within the group)
    double dropped = 0.0;  // this lane's share of the fill updates outside the pattern
    bool done = !active;
    if (active && overflow) {
        for (int q = t; q < len; q += G) {
            publish(vout + rb + q, __longlong_as_double(kQuietNan));
            if (f32) f32[rb + q] = __builtin_nanf("");
        }
        if (t == 0) atomicOr(err, 2);
        done = true;
    }
    while (true) {
        if (!done) {
            while (p < ndiag) {
                const int k = col[p];
                const int dk = dg[k];
                const int ke = rp[k + 1];
                const uint64_t ub = peek_bits(vout + dk);
                bool ok = ub != kPending;
                for (int q = dk + 1 + t; ok && q < ke; q += G) ok = peek_bits(vout + q) != kPending;
                if (!group_all<G>(ok)) break;
                const double ukk = __longlong_as_double(ub);
                const double lik = val[p] / ukk;
                if (ukk == 0.0 && t == 0) atomicOr(err, 1);
                for (int q = dk + 1 + t; q < ke; q += G) {
                    const int j = ci[q];
                    int lo = p + 1, hi = len;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (col[mid] < j) lo = mid + 1; else hi = mid;
                    }
                    const double u = __longlong_as_double(peek_bits(vout + q));
                    if (lo < len && col[lo] == j) val[lo] = fma(-u, lik, val[lo]);
                    else dropped = fma(-u, lik, dropped);
                }
                if (t == 0) val[p] = lik;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                ++p;
            }
            if (p < ndiag && group_all<G>(wd.expired())) {
                if (t == 0) atomicOr(err, kErrTimeout);
                for (int q = t; q < len; q += G) val[q] = __longlong_as_double(kQuietNan);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                p = ndiag;
            }
            if (p >= ndiag) {
                if (omega) {
                    // diagonal compensation: u_ii += omega_i * (sum of dropped fill)
                    const double dsum = group_sum<G>(dropped);
                    if (t == 0) val[ndiag] = fma(static_cast<double>(omega[row]), dsum, val[ndiag]);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                }
                if (pg.mode != 0) guard_pivot(pg, val + ndiag, row, t);
                for (int q = t; q < len; q += G) {
                    const double v = val[q];
                    if (f32) f32[rb + q] = static_cast<float>(v);
                    publish(vout + rb + q, v);
                }
                done = true;
            }
        }
        if (__ballot(!done) == 0) break;
        spin_pause();
    }
    retire_ticket(ctr);
}

// The same factorization for short rows: at most P pivots per row, at most G
// strictly upper entries per pivot row (one per lane), rows of at most CAP
// entries.  The arithmetic and its order are k_factor's (lane t owns entry
// dk+1+t of every pivot row, pivots in ascending order, the same fma and
// the same lane-strided sum of dropped updates), so the factors are
// identical.  What changes is the waiting: the pivot rows' bounds, columns
// and target slots come from the static pattern and are all fetched before
// any wait, and the values of every pivot row are polled together, so a row
// costs one round trip after its last pivot row is published instead of two
// round trips per pivot.
template <int G, int CAP, int P>
__global__ __launch_bounds__(kBlock) void k_factor_short(int n, const int* __restrict__ perm,
                                                         const int* __restrict__ rp,
                                                         const int* __restrict__ ci,
                                                         const int* __restrict__ dg,
                                                         const double* __restrict__ vin, double* vout,
                                                         float* f32, int* err, int* ctr, uint64_t wd_limit,
                                                         const float* __restrict__ omega, const PivotGuard pg) {
    constexpr int R = 64 / G;
    constexpr int GROUPS = kBlock / G;
    __shared__ int s_col[GROUPS][CAP];
    __shared__ double s_val[GROUPS][CAP];
    const int ticket = take_ticket(ctr);
    const Watchdog wd(wd_limit);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int slot = lane / G;
    const int grp = wave * R + slot;
    const int t = lane & (G - 1);
    const long pos = static_cast<long>(ticket) * GROUPS + grp;
    const bool active = pos < n;
    int* col = s_col[grp];
    double* val = s_val[grp];
    int row = 0, rb = 0, len = 0, ndiag = 0;
    if (active) {
        row = perm[pos];
        rb = rp[row];
        len = rp[row + 1] - rb;
        ndiag = dg[row] - rb;
        for (int q = t; q < len; q += G) {
            col[q] = ci[rb + q];
            val[q] = vin[rb + q];
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    // pattern of the pivot rows: diagonal position, this lane's entry and
    // its slot in row i (-1: outside the pattern, -2: no entry)
    int dk[P], qe[P], tgt[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        dk[p] = 0;
        qe[p] = 0;
        tgt[p] = -2;
        if (active && p < ndiag) {
            const int k = col[p];
            dk[p] = dg[k];
            const int ke = rp[k + 1];
            const int q = dk[p] + 1 + t;
            if (q < ke) {
                qe[p] = q;
                const int j = ci[q];
                int lo = p + 1, hi = len;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (col[mid] < j) lo = mid + 1; else hi = mid;
                }
                tgt[p] = (lo < len && col[lo] == j) ? lo : -1;
            }
        }
    }
    uint64_t ub[P], uu[P];
#pragma unroll
    for (int p = 0; p < P; ++p) ub[p] = uu[p] = kPending;
    bool done = !active;
    int pn = 0;              // pivots applied so far (uniform in the group)
    double dropped = 0.0;    // this lane's share of the fill updates outside the pattern
    while (true) {
        if (!done) {
            // poll every pivot not applied yet, then apply the ready ones in
            // ascending order: only the last pivot to arrive stays on the
            // critical path
#pragma unroll
            for (int p = 0; p < P; ++p) {
                if (p >= pn && p < ndiag) {
                    if (ub[p] == kPending) ub[p] = peek_bits(vout + dk[p]);
                    if (tgt[p] != -2 && uu[p] == kPending) uu[p] = peek_bits(vout + qe[p]);
                }
            }
#pragma unroll
            for (int p = 0; p < P; ++p) {
                if (p == pn && p < ndiag) {
                    if (!group_all<G>(ub[p] != kPending && (tgt[p] == -2 || uu[p] != kPending))) break;
                    const double ukk = __longlong_as_double(ub[p]);
                    const double lik = val[p] / ukk;
                    if (ukk == 0.0 && t == 0) atomicOr(err, 1);
                    const double u = __longlong_as_double(uu[p]);
                    if (tgt[p] >= 0) val[tgt[p]] = fma(-u, lik, val[tgt[p]]);
                    else if (tgt[p] == -1) dropped = fma(-u, lik, dropped);
                    if (t == 0) val[p] = lik;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    ++pn;
                }
            }
            if (pn < ndiag && group_all<G>(wd.expired())) {
                if (t == 0) atomicOr(err, kErrTimeout);
                for (int q = t; q < len; q += G) val[q] = __longlong_as_double(kQuietNan);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                pn = ndiag;
            }
            if (pn >= ndiag) {
                if (omega) {
                    const double dsum = group_sum<G>(dropped);
                    if (t == 0) val[ndiag] = fma(static_cast<double>(omega[row]), dsum, val[ndiag]);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                }
                if (pg.mode != 0) guard_pivot(pg, val + ndiag, row, t);
                for (int q = t; q < len; q += G) {
                    const double v = val[q];
                    if (f32) f32[rb + q] = static_cast<float>(v);
                    publish(vout + rb + q, v);
                }
                done = true;
            }
        }
        if (__ballot(!done) == 0) break;
        spin_pause();
    }
    retire_ticket(ctr);
}

// Zero the values to factor and arm the factor output with kPending.
__global__ void k_prepare_factor(long nnz, double* __restrict__ vin, uint64_t* __restrict__ vout,
                                 bool zero_in) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nnz; i += (long)gridDim.x * blockDim.x) {
        if (zero_in) vin[i] = 0.0;
        vout[i] = kPending;
    }
}

// F.value[nzmap[k]] = A.value[k] (test/benchmark.cpp:116-119 on the device).
__global__ void k_scatter_values(long nnz_a, const int* __restrict__ nzmap, const double* __restrict__ a,
                                 double* __restrict__ f) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < nnz_a; k += (long)gridDim.x * blockDim.x)
        f[nzmap[k]] = a[k];
}

}  // namespace pilu
