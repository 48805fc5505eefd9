// Wave-specialised block sweep (sweep mode 5).
//
// Same partition and schedule as the block-local sweeps (blocks processed
// in a topological order of the block graph, rows of a block in local
// level order), but each workgroup splits the work of one block over three
// waves so that no wave waits on memory requests issued for another
// purpose:
//   * wave 0 solves.  Its inner loop touches LDS only: the block's x (b on
//     entry) and the imported out-of-block x live in LDS, and the factor
//     data arrive in LDS through a ring.  One lane group of G lanes per
//     row, up to 64 / G rows per record (one record = one chunk of a local
//     level).  Results go to `out` with plain stores and, for rows other
//     blocks read ("face" rows), also to a compact face buffer as
//     value-flags.
//   * wave 1 streams the block's record stream -- headers, row ids,
//     columns already translated to LDS slots, values, diagonal -- from HBM
//     into the LDS ring in groups of eight 1 KiB chunks, two groups in
//     flight, flow-controlled by two LDS counters (bytes loaded / bytes
//     consumed).
//   * wave 2 imports: it polls the face buffer for the block's out-of-block
//     dependencies, in order of first use and at most kWsLook records ahead
//     of the solver, and writes each value into its LDS slot as a
//     value-flag (kPending until imported).
// Row sums follow the lane order of every other sweep kind (lane t of a
// row's group accumulates entries t, t+G, ... with fma, lane 0 from b_i,
// then the group butterfly), so results are bit-identical to k_sweep with
// the same G.  Padding entries (column = a slot holding 0.0, value 0) add
// an exact -0.0.
//
// Record layout (16-byte aligned sections): header int4 {rows, entries per
// row w (a multiple of G), record bytes, first slot}; int2 {row, face slot
// or -1} per row; int columns [rows][w] as LDS double indices; T values
// [rows][w]; T diagonal [rows] (backward sweep).  A record never wraps the
// ring (the stream is padded at ring boundaries) and is at most half the
// ring.
#pragma once

#include "common.hpp"

namespace pilu {

constexpr int kWsThreads = 192;   // solver, loader and importer waves
constexpr int kWsChunk = 1024;    // bytes the loader moves per step (64 lanes x 16 B)
constexpr int kWsRing = 16384;    // LDS ring of the record stream (bytes)
constexpr int kWsGroup = 8;       // chunks the loader writes per ring-space check (two groups in flight)
constexpr int kWsImp = 4;         // imports each importer lane polls at once
constexpr int kWsLook = 8;        // records ahead of the solver the importer polls
constexpr int kErrStream = 1 << 9;  // error word: a record stream row or face id out of range

struct WsBlock {
    int s0, s1;     // record stream of the block, in 16-byte units (a multiple of kWsChunk bytes)
    int i0, i1;     // its imports
    int r0, r1;     // face slots of the other sweep direction re-armed by this block
    int k0, nrows;  // first schedule position, rows
};

__device__ __forceinline__ int ws_align16(int b) { return (b + 15) & ~15; }

__device__ __forceinline__ int lds_load_i(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_store_i(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint64_t lds_load_u64(const double* p) {
    return __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_store_u64(double* p, uint64_t v) {
    __hip_atomic_store(reinterpret_cast<uint64_t*>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local"); }
__device__ __forceinline__ void lds_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local"); }

// Waits until the loader has put `need` bytes of the stream into the ring;
// false on a watchdog expiry.
__device__ __forceinline__ bool ws_wait_loaded(const int* ctl, int need, const Watchdog& wd) {
    while (lds_load_i(&ctl[0]) < need) {
        if (wd.expired()) return false;
        __builtin_amdgcn_s_sleep(0);
    }
    lds_acquire();
    return true;
}

// One record as the solver holds it: header, this lane's row data and the
// columns and values of its first S entries (read from the ring one record
// ahead of their use).
template <int S, typename T>
struct WsRec {
    int cnt, w, rbytes, slot0, off, vals_b;
    int2 rf;
    double dv, b;
    int c[S];
    T vr[S];
};

// Reads the next record at stream offset `off` (skipping ring padding);
// false at the block's end or on a watchdog expiry (timed_out set).
template <int G, int S, typename T, bool UPPER>
__device__ __forceinline__ bool ws_fetch(int off, int bytes, const char* ring, const int* ctl, const double* xl,
                                         int zero_slot, const Watchdog& wd, bool& timed_out, WsRec<S, T>& r) {
    const int lane = threadIdx.x & 63;
    const int j = lane / G, t = lane & (G - 1);
    int cnt = 0, w = 0, rbytes = 0, slot0 = 0, base = 0;
    while (true) {
        if (off >= bytes) return false;
        if (!ws_wait_loaded(ctl, off + 16, wd)) { timed_out = true; return false; }
        base = off & (kWsRing - 1);
        const int4 h = *reinterpret_cast<const int4*>(ring + base);
        cnt = __builtin_amdgcn_readfirstlane(h.x);
        w = __builtin_amdgcn_readfirstlane(h.y);
        rbytes = __builtin_amdgcn_readfirstlane(h.z);
        slot0 = __builtin_amdgcn_readfirstlane(h.w);
        if (cnt > 0) break;
        if (rbytes <= 0) return false;  // padding up to the block's end
        off += rbytes;                   // padding up to a ring boundary
    }
    if (!ws_wait_loaded(ctl, off + rbytes, wd)) { timed_out = true; return false; }
    const int rows_b = base + 16;
    const int cols_b = rows_b + ws_align16(cnt * 8);
    const int vals_b = cols_b + ws_align16(cnt * w * 4);
    const int diag_b = vals_b + ws_align16(cnt * w * static_cast<int>(sizeof(T)));
    const int jj = j < cnt ? j : 0;
    r.cnt = cnt;
    r.w = w;
    r.rbytes = rbytes;
    r.slot0 = slot0;
    r.off = off;
    r.vals_b = vals_b + jj * w * static_cast<int>(sizeof(T));
    r.rf = reinterpret_cast<const int2*>(ring + rows_b)[jj];
    r.dv = UPPER ? static_cast<double>(reinterpret_cast<const T*>(ring + diag_b)[jj]) : 1.0;
    r.b = xl[slot0 + jj];
    const int* cp = reinterpret_cast<const int*>(ring + cols_b) + jj * w;
    const T* vp = reinterpret_cast<const T*>(ring + vals_b) + jj * w;
    // branch-free: terms past the row's width read its last entry and are
    // turned into (zero slot, 0), which adds -0.0
    const int wl = w > 0 ? w - 1 : 0;
#pragma unroll
    for (int s = 0; s < S; ++s) r.c[s] = w > 0 ? cp[t + s * G < w ? t + s * G : wl] : zero_slot;
#pragma unroll
    for (int s = 0; s < S; ++s) r.vr[s] = w > 0 ? vp[t + s * G < w ? t + s * G : wl] : T(0);
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (!(s * G < w)) {
            r.c[s] = zero_slot;
            r.vr[s] = T(0);
        }
    return true;
}

// Wave 0: walk the block's records, two stages deep: while record r's x
// are read and summed, record r+1's header, row data, columns and values
// are read from the ring, so only the x reads (which depend on the rows
// just solved) stay on the critical path.
//
// Ring flow control: the solver holds at most one record's ring bytes
// unconsumed while it waits for the loader.  A record whose entries all sit
// in registers after ws_fetch (w <= S * G) is released before record r+1
// is fetched; a longer row's record is read again by the tail loop, so it
// is released after its rows are solved and record r+1 is fetched then.
// Holding r while waiting for r+1 could deadlock: the loader refills a
// ring half only once the half before it is consumed, and two records can
// together exceed the free half.
template <int G, int S, typename T, bool UPPER>
__device__ void ws_solve(const WsBlock& bk, const char* ring, int* ctl, double* xl, int zero_slot, double* out,
                         double* face, int n, int nface, const Watchdog& wd, bool& timed_out, bool& bad) {
    const int lane = threadIdx.x & 63;
    const int j = lane / G;
    const int t = lane & (G - 1);
    const int bytes = (bk.s1 - bk.s0) * 16;
    WsRec<S, T> cur, nxt;
    bool have = ws_fetch<G, S, T, UPPER>(0, bytes, ring, ctl, xl, zero_slot, wd, timed_out, cur);
    int rec = 0;
    while (have) {
        const bool act = j < cur.cnt;
        const bool tail = cur.w > S * G;
        const int next_off = cur.off + cur.rbytes;
        if (!tail) {
            lds_release();  // this record's ring reads have returned
            if (lane == 0) lds_store_i(&ctl[1], next_off);
        }
        // x of this record's first S entries: every read at once
        uint64_t xb[S];
#pragma unroll
        for (int s = 0; s < S; ++s) xb[s] = lds_load_u64(xl + cur.c[s]);
        // the next record, while those reads are in flight
        bool more = false;
        if (!tail) more = ws_fetch<G, S, T, UPPER>(next_off, bytes, ring, ctl, xl, zero_slot, wd, timed_out, nxt);
        bool pend = false;
#pragma unroll
        for (int s = 0; s < S; ++s) pend |= xb[s] == kPending;
        if (pend) {
            // an import not there yet: wait for it
#pragma unroll
            for (int s = 0; s < S; ++s)
                while (xb[s] == kPending) {
                    if (timed_out || wd.expired()) { timed_out = true; xb[s] = kQuietNan; break; }
                    __builtin_amdgcn_s_sleep(0);
                    xb[s] = lds_load_u64(xl + cur.c[s]);
                }
        }
        double acc = act && t == 0 ? cur.b : 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) acc = fma(-__longlong_as_double(xb[s]), static_cast<double>(cur.vr[s]), acc);
        if (act && tail) {
            // entries past the first S per lane (rows longer than S * G)
            const int rows_b = (cur.off & (kWsRing - 1)) + 16;
            const int* cp = reinterpret_cast<const int*>(ring + rows_b + ws_align16(cur.cnt * 8)) + j * cur.w;
            const T* vp = reinterpret_cast<const T*>(ring + cur.vals_b);
            for (int q = t + S * G; q < cur.w; q += G) {
                const int cq = cp[q];
                uint64_t bits = lds_load_u64(xl + cq);
                while (bits == kPending) {
                    if (timed_out || wd.expired()) { timed_out = true; bits = kQuietNan; break; }
                    __builtin_amdgcn_s_sleep(0);
                    bits = lds_load_u64(xl + cq);
                }
                acc = fma(-__longlong_as_double(bits), static_cast<double>(vp[q]), acc);
            }
        }
        const double sum = group_sum<G>(acc);
        if (act && t == 0) {
            const double res = UPPER ? sum / cur.dv : sum;
            uint64_t rb = __double_as_longlong(res);
            if (res != res) rb = kQuietNan;
            lds_store_u64(xl + cur.slot0 + j, rb);
            // row ids come from the stream the host built; an id out of
            // range means a corrupted stream and is reported, not stored
            if (static_cast<unsigned>(cur.rf.x) < static_cast<unsigned>(n)) out[cur.rf.x] = __longlong_as_double(rb);
            else bad = true;
            if (static_cast<unsigned>(cur.rf.y) < static_cast<unsigned>(nface))
                publish_raw(reinterpret_cast<uint64_t*>(face) + cur.rf.y, rb);
            else if (cur.rf.y != -1) bad = true;
        }
        ++rec;
        if (tail) {
            lds_release();  // the tail loop's ring reads have returned
            if (lane == 0) lds_store_i(&ctl[1], next_off);
        }
        if (lane == 0) lds_store_i(&ctl[2], rec);
        if (tail) more = ws_fetch<G, S, T, UPPER>(next_off, bytes, ring, ctl, xl, zero_slot, wd, timed_out, nxt);
        cur = nxt;
        have = more;
    }
    if (lane == 0) {
        lds_store_i(&ctl[1], bytes);
        lds_store_i(&ctl[2], 1 << 30);  // the importer may stop waiting for progress
    }
}

// Wave 1: stream the block's records into the ring.  The ring holds two
// groups of kWsGroup chunks; the loads of group g+2 are issued as group g
// is written, so two groups are always in flight.  Inside a group there is
// no control flow, which lets the compiler wait for each chunk with an
// exact count of the younger loads instead of draining them all.
//
// The loader has no watchdog of its own and never writes a ring half
// before the solver has consumed it: its only wait is for the consumed
// count, which the solver sets to the whole stream when it leaves (after
// its last record or on its own watchdog expiry), so the wait always ends.
__device__ __forceinline__ void ws_load(const WsBlock& bk, const uint4* __restrict__ stream, uint4* ring, int* ctl) {
    constexpr int kGroupBytes = kWsGroup * kWsChunk;
    static_assert(2 * kGroupBytes == kWsRing, "the ring holds two groups");
    const int lane = threadIdx.x & 63;
    const int nch = (bk.s1 - bk.s0) / (kWsChunk / 16);
    if (nch <= 0) return;
    const int ngroups = (nch + kWsGroup - 1) / kWsGroup;
    const long src0 = bk.s0;
    auto src = [&](int c) -> uint4 {
        const int cc = c < nch ? c : nch - 1;
        return stream[src0 + static_cast<long>(cc) * (kWsChunk / 16) + lane];
    };
    uint4 d0[kWsGroup], d1[kWsGroup];
    // sched_barrier(0) keeps the loads in issue order (group 0 before group
    // 1, chunk u before u+1): the in-order memory counter then lets each
    // write wait for its own group only
#pragma unroll
    for (int u = 0; u < kWsGroup; ++u) d0[u] = src(u);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < kWsGroup; ++u) d1[u] = src(kWsGroup + u);
    __builtin_amdgcn_sched_barrier(0);
    // No exit leaves the loop between the two groups' loads.
    auto put = [&](int g, uint4 (&d)[kWsGroup]) {
        // group g reuses the ring half of group g-2: wait until it is
        // consumed (need <= the stream's bytes for every g <= ngroups)
        const int need = (g + 1) * kGroupBytes - kWsRing;
        while (lds_load_i(&ctl[1]) < need) __builtin_amdgcn_s_sleep(1);
        lds_acquire();  // the solver's reads of that half have returned
        uint4* half = ring + (g & 1) * (kGroupBytes / 16);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kWsGroup; ++u) half[u * (kWsChunk / 16) + lane] = d[u];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kWsGroup; ++u) d[u] = src((g + 2) * kWsGroup + u);
        __builtin_amdgcn_sched_barrier(0);
        lds_release();
        if (lane == 0) lds_store_i(&ctl[0], min((g + 1) * kWsGroup, nch) * kWsChunk);
    };
    // both puts run on every iteration (a put past the last group rewrites a
    // ring half the solver has consumed and does not move the loaded count):
    // with no path that skips the second one, the compiler's count of the
    // loads in flight at the loop head stays exact
    for (int g = 0; g < ngroups; g += 2) {
        put(g, d0);
        put(g + 1, d1);
    }
}

// Wave 2: import the block's out-of-block dependencies into their slots.
__device__ __forceinline__ void ws_import(const WsBlock& bk, const int2* __restrict__ imports, double* imp,
                                          const double* face, const int* ctl, const Watchdog& wd,
                                          bool& timed_out) {
    const int lane = threadIdx.x & 63;
    const int n = bk.i1 - bk.i0;
    int idx[kWsImp];
    int2 ent[kWsImp];
#pragma unroll
    for (int k = 0; k < kWsImp; ++k) {
        idx[k] = lane + 64 * k;
        ent[k] = idx[k] < n ? imports[bk.i0 + idx[k]] : make_int2(0, 0);
    }
    while (true) {
        const int rec = lds_load_i(&ctl[2]);
        uint64_t bits[kWsImp];
        bool want[kWsImp];
#pragma unroll
        for (int k = 0; k < kWsImp; ++k) {
            want[k] = idx[k] < n && ent[k].y <= rec + kWsLook;
            bits[k] = want[k] ? peek_raw(reinterpret_cast<const uint64_t*>(face) + ent[k].x) : kPending;
        }
        bool progress = false, left = false;
#pragma unroll
        for (int k = 0; k < kWsImp; ++k) {
            if (want[k] && bits[k] != kPending) {
                lds_store_u64(imp + idx[k], bits[k]);
                idx[k] += 64 * kWsImp;
                ent[k] = idx[k] < n ? imports[bk.i0 + idx[k]] : make_int2(0, 0);
                progress = true;
            }
            left |= idx[k] < n;
        }
        if (__ballot(left) == 0) break;
        if (__ballot(progress) == 0) {
            if (timed_out || wd.expired()) {
                // give up: every remaining slot gets NaN so the solver drains
                timed_out = true;
#pragma unroll
                for (int k = 0; k < kWsImp; ++k)
                    for (int i = idx[k]; i < n; i += 64 * kWsImp) lds_store_u64(imp + i, kQuietNan);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
}

// One workgroup per block, blocks taken in processing order by ticket.
// LDS (dynamic): the block's x (max_rows), the import slots (max_imp), one
// slot holding 0.0 (padding entries), then the ring.  reset (optional):
// the other sweep's output, re-armed (kPending) row by row after b is
// staged; face_rearm: the other direction's face buffer, whose slots
// [r0, r1) this block re-arms.
template <int G, int S, typename T, bool UPPER>
__global__ __launch_bounds__(kWsThreads) void k_sweep_ws(const WsBlock* __restrict__ blocks,
                                                        const uint4* __restrict__ stream,
                                                        const int2* __restrict__ imports, const int* __restrict__ perm,
                                                        const double* b, double* out, double* reset, double* face,
                                                        double* face_rearm, const int* gate, int* ctr, int* err,
                                                        uint64_t wd_limit, int max_rows, int max_imp, int n,
                                                        int nface) {
    extern __shared__ double ws_lds[];
    __shared__ int ctl[4];  // bytes loaded, bytes consumed, records solved
    if (gate && *gate == 0) return;
    const int ticket = take_ticket(ctr);
    const WsBlock bk = blocks[ticket];
    const Watchdog wd(wd_limit);
    double* xl = ws_lds;
    double* imp = ws_lds + max_rows;
    char* ring = reinterpret_cast<char*>(ws_lds + ((max_rows + max_imp + 1 + 1) & ~1));
    const int tid = threadIdx.x;
    if (tid < 4) ctl[tid] = 0;
    // stage b (eight rows per thread in flight), arm the import slots, the
    // zero slot, and the other direction's face slots of this block
    for (int i0 = 0; i0 < bk.nrows; i0 += 8 * kWsThreads) {
        int r[8];
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + u * kWsThreads + tid;
            r[u] = i < bk.nrows ? perm[bk.k0 + i] : -1;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = r[u] >= 0 ? b[r[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + u * kWsThreads + tid;
            if (r[u] >= 0) {
                xl[i] = v[u];
                if (reset) reset[r[u]] = __longlong_as_double(kPending);
            }
        }
    }
    for (int i = tid; i < bk.i1 - bk.i0; i += kWsThreads) imp[i] = __longlong_as_double(kPending);
    if (tid == 0) imp[max_imp] = 0.0;
    if (face_rearm)
        for (int f = bk.r0 + tid; f < bk.r1; f += kWsThreads) face_rearm[f] = __longlong_as_double(kPending);
    __syncthreads();
    bool timed_out = false, bad = false;
    const int wave = tid >> 6;
    if (wave == 0) {
        // the solver is the critical path: the SIMD issues its instructions
        // ahead of spinning loader and importer waves
        __builtin_amdgcn_s_setprio(3);
        ws_solve<G, S, T, UPPER>(bk, ring, ctl, xl, max_rows + max_imp, out, face, n, nface, wd, timed_out, bad);
        __builtin_amdgcn_s_setprio(0);
    }
    else if (wave == 1) ws_load(bk, stream, reinterpret_cast<uint4*>(ring), ctl);
    else ws_import(bk, imports, imp, face, ctl, wd, timed_out);
    if (timed_out) atomicOr(err, kErrTimeout);
    if (bad) atomicOr(err, kErrStream);
    retire_ticket(ctr);
}

// Values of the record streams from the factors: stream_T[dst[e]] =
// fac[src[e]] (src < 0: padding, 0).
template <typename TI, typename TO>
__global__ void k_ws_gather(long ne, const int* __restrict__ dst, const int* __restrict__ src,
                            const TI* __restrict__ fac, TO* __restrict__ stream) {
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    for (long e = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; e < ne; e += stride) {
        const int s = src[e];
        stream[dst[e]] = s >= 0 ? static_cast<TO>(fac[s]) : TO(0);
    }
}

}  // namespace pilu
