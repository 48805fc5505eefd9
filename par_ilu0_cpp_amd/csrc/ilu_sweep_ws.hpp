// Wave-specialised block sweep (sweep mode 5).
//
// Same partition and schedule as the block-local sweeps (blocks processed
// in a topological order of the block graph, rows of a block in local
// level order), but each workgroup splits the work of one block over four
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
//     into a two-group LDS ring in groups of eight 1 KiB chunks, four
//     groups in flight, flow-controlled by two LDS counters (bytes loaded /
//     bytes consumed).
//   * waves 2 and 3 import: each polls the face buffer for the block's
//     out-of-block dependencies, in order of first use and at most kWsLook
//     records ahead of the solver, and writes each value into its LDS slot
//     as a value-flag (kPending until imported).  The two poll the same
//     imports half a poll round apart, which shortens the time from a
//     producer's publish to the import (a poll round is one memory round
//     trip).
// Row sums follow the lane order of every other sweep kind (lane t of a
// row's group accumulates entries t, t+G, ... with fma, lane 0 from b_i,
// then the group butterfly), so results are bit-identical to k_sweep with
// the same G.  Padding entries (column = a slot holding 0.0, value 0) add
// an exact -0.0.
//
// Record layout (16-byte aligned sections): header int4 {rows, entries per
// row w (a multiple of G), record bytes, first slot}; the header of the
// record that follows (zeros after the block's last record), so that its
// body can be read without first reading its header; int2 {row, face slot
// or -1} per row; the first S * G entries of every row lane-major -- lane
// (j, t) of row j holds entries t, t+G, ..., t+(S-1)G as S int columns (LDS
// double indices) and, in the next section, S values -- so each lane reads
// its operands with a few wide, conflict-free LDS reads; entries past S * G
// row-major (rows longer than S * G only), columns then values; T diagonal
// [rows] (backward sweep).  Entries past a row's own count are (zero slot,
// 0).  A record never wraps the ring (the stream is padded at ring
// boundaries) and is at most half the ring.
#pragma once

#include "common.hpp"

namespace pilu {

constexpr int kWsThreads = 256;   // solver, loader and two importer waves
constexpr int kWsChunk = 1024;    // bytes the loader moves per step (64 lanes x 16 B)
constexpr int kWsRing = 16384;    // LDS ring of the record stream (bytes)
constexpr int kWsGroup = 8;       // chunks the loader writes per ring-space check (the ring holds two groups)
constexpr int kWsFlight = 4;      // groups the loader keeps in flight (in registers) ahead of the ring
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

// One record as the solver holds it: header, the next record's header,
// this lane's row data and the columns and values of its first S entries.
template <int S, typename T>
struct WsRec {
    int cnt, w, rbytes, slot0, off;
    int4 nh;
    int2 rf;
    double dv, b;
    int c[S];
    T vr[S];
};

__device__ __forceinline__ int4 ws_uniform(int4 v) {
    return make_int4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                     __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
}

// Issues the reads of a record's body whose header h is known (no wait:
// the values are used later).
template <int G, int S, typename T, bool UPPER>
__device__ __forceinline__ void ws_body(int off, int4 h, const char* ring, const double* xl, WsRec<S, T>& r) {
    const int lane = threadIdx.x & 63;
    const int j = lane / G, t = lane & (G - 1);
    const int cnt = h.x, w = h.y;
    const int base = off & (kWsRing - 1);
    const int rows_b = base + 32;
    const int lcols_b = rows_b + ws_align16(cnt * 8);
    const int lvals_b = lcols_b + ws_align16(cnt * G * S * 4);
    const int jj = j < cnt ? j : 0;
    const int li = jj * G + t;
    r.cnt = cnt;
    r.w = w;
    r.rbytes = h.z;
    r.slot0 = h.w;
    r.off = off;
    r.nh = *reinterpret_cast<const int4*>(ring + base + 16);
    r.rf = reinterpret_cast<const int2*>(ring + rows_b)[jj];
    r.b = xl[h.w + jj];
    const int* cp = reinterpret_cast<const int*>(ring + lcols_b) + li * S;
    const T* vp = reinterpret_cast<const T*>(ring + lvals_b) + li * S;
#pragma unroll
    for (int s = 0; s < S; ++s) r.c[s] = cp[s];
#pragma unroll
    for (int s = 0; s < S; ++s) r.vr[s] = vp[s];
    if (UPPER) {
        const int wt = w > S * G ? w - S * G : 0;
        const int tcols_b = lvals_b + ws_align16(cnt * G * S * static_cast<int>(sizeof(T)));
        const int diag_b = tcols_b + (wt > 0 ? ws_align16(cnt * wt * 4) + ws_align16(cnt * wt * static_cast<int>(sizeof(T))) : 0);
        r.dv = static_cast<double>(reinterpret_cast<const T*>(ring + diag_b)[jj]);
    }
    else r.dv = 1.0;
}

// Wave 0: walk the block's records, two stages deep: record r+1's body
// (its header is in record r) is read from the ring together with record
// r's x, so each record costs one LDS round trip, the x reads (which
// depend on the rows just solved).  Record r is never held back waiting
// for r+1's bytes: if the loader has not put them in the ring yet, r+1 is
// read after r is solved.  The loaded count is re-read only when a record
// ends past the last count seen.
//
// Ring flow control: the solver holds at most one record's ring bytes
// unconsumed while it waits for the loader.  A record whose entries all sit
// in registers after ws_body (w <= S * G) is released before record r+1 is
// read; a longer row's record is read again by the tail loop, so it is
// released after its rows are solved and record r+1 is read then.  Holding
// r while waiting for r+1 could deadlock: the loader refills a ring half
// only once the half before it is consumed, and two records can together
// exceed the free half.
template <int G, int S, typename T, bool UPPER>
__device__ void ws_solve(const WsBlock& bk, const char* ring, int* ctl, double* xl, double* out, double* face, int n,
                         int nface, const Watchdog& wd, bool& timed_out, bool& bad) {
    const int lane = threadIdx.x & 63;
    const int j = lane / G;
    const int t = lane & (G - 1);
    const int bytes = (bk.s1 - bk.s0) * 16;
    int loaded = 0;  // last loaded count seen
    // waits until `need` bytes are in the ring; false on a watchdog expiry
    auto wait_loaded = [&](int need) -> bool {
        while (loaded < need) {
            loaded = __builtin_amdgcn_readfirstlane(lds_load_i(&ctl[0]));
            if (loaded >= need) break;
            if (wd.expired()) return false;
            __builtin_amdgcn_s_sleep(0);
        }
        lds_acquire();
        return true;
    };
    WsRec<S, T> cur, nxt;
    bool live = bytes > 0;
    if (live) {
        if (!wait_loaded(16)) { timed_out = true; live = false; }
    }
    if (live) {
        const int4 h = ws_uniform(*reinterpret_cast<const int4*>(ring));
        live = h.x > 0;
        if (live && !wait_loaded(h.z)) { timed_out = true; live = false; }
        if (live) ws_body<G, S, T, UPPER>(0, h, ring, xl, cur);
    }
    int rec = 0;
    while (live) {
        const int4 nh = ws_uniform(cur.nh);
        const bool has_next = nh.x > 0;
        const int end = cur.off + cur.rbytes;
        // the next record starts at the next ring boundary if it would wrap
        const int nxt_off = has_next && (end & (kWsRing - 1)) + nh.z > kWsRing ? (end + kWsRing - 1) & ~(kWsRing - 1)
                                                                                : end;
        const int nxt_end = nxt_off + nh.z;
        const bool act = j < cur.cnt;
        const bool tail = cur.w > S * G;
        if (!tail) {
            lds_release();  // this record's ring reads have returned
            if (lane == 0) lds_store_i(&ctl[1], end);
        }
        // x of this record's first S entries: every read at once
        uint64_t xb[S];
#pragma unroll
        for (int s = 0; s < S; ++s) xb[s] = lds_load_u64(xl + cur.c[s]);
        // the next record's body, while those reads are in flight, if loaded
        bool fetched = false;
        if (has_next && !tail) {
            if (loaded < nxt_end) {
                loaded = __builtin_amdgcn_readfirstlane(lds_load_i(&ctl[0]));
                lds_acquire();
            }
            if (loaded >= nxt_end) {
                ws_body<G, S, T, UPPER>(nxt_off, nh, ring, xl, nxt);
                fetched = true;
            }
        }
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
            const int wt = cur.w - S * G;
            const int base = cur.off & (kWsRing - 1);
            const int tcols_b = base + 32 + ws_align16(cur.cnt * 8) + ws_align16(cur.cnt * G * S * 4) +
                                ws_align16(cur.cnt * G * S * static_cast<int>(sizeof(T)));
            const int* cp = reinterpret_cast<const int*>(ring + tcols_b) + j * wt;
            const T* vp = reinterpret_cast<const T*>(ring + tcols_b + ws_align16(cur.cnt * wt * 4)) + j * wt;
            for (int q = t; q < wt; q += G) {
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
            if (lane == 0) lds_store_i(&ctl[1], end);
        }
        if (lane == 0) lds_store_i(&ctl[2], rec);
        if (!has_next) break;
        if (!fetched) {
            if (!wait_loaded(nxt_end)) { timed_out = true; break; }
            ws_body<G, S, T, UPPER>(nxt_off, nh, ring, xl, nxt);
        }
        cur = nxt;
    }
    if (lane == 0) {
        lds_store_i(&ctl[1], bytes);
        lds_store_i(&ctl[2], 1 << 30);  // the importer may stop waiting for progress
    }
}

// Wave 1: stream the block's records into the ring.  The ring holds two
// groups of kWsGroup chunks; kWsFlight groups are in flight in registers
// ahead of it (the loads of group g+kWsFlight are issued as group g is
// written), so a block streams kWsFlight * 8 KiB per memory latency.
// Inside a group there is no control flow, which lets the compiler wait
// for each chunk with an exact count of the younger loads instead of
// draining them all.
//
// The loader has no watchdog of its own and never writes a ring half
// before the solver has consumed it: its only wait is for the consumed
// count, which the solver sets to the whole stream when it leaves (after
// its last record or on its own watchdog expiry), so the wait always ends.
__device__ __forceinline__ void ws_load(const WsBlock& bk, const uint4* __restrict__ stream, uint4* ring, int* ctl) {
    constexpr int kGroupBytes = kWsGroup * kWsChunk;
    static_assert(2 * kGroupBytes == kWsRing, "the ring holds two groups");
    static_assert(kWsFlight % 2 == 0, "groups in flight alternate over the two ring halves");
    const int lane = threadIdx.x & 63;
    const int nch = (bk.s1 - bk.s0) / (kWsChunk / 16);
    if (nch <= 0) return;
    const int ngroups = (nch + kWsGroup - 1) / kWsGroup;
    const long src0 = bk.s0;
    auto src = [&](int c) -> uint4 {
        const int cc = c < nch ? c : nch - 1;
        return stream[src0 + static_cast<long>(cc) * (kWsChunk / 16) + lane];
    };
    uint4 d[kWsFlight][kWsGroup];
    // sched_barrier(0) keeps the loads in issue order (group before group,
    // chunk u before u+1): the in-order memory counter then lets each write
    // wait for its own group only
#pragma unroll
    for (int q = 0; q < kWsFlight; ++q) {
#pragma unroll
        for (int u = 0; u < kWsGroup; ++u) d[q][u] = src(q * kWsGroup + u);
        __builtin_amdgcn_sched_barrier(0);
    }
    // No exit leaves the loop between the groups' loads.
    auto put = [&](int g, uint4 (&dq)[kWsGroup]) {
        // group g reuses the ring half of group g-2: wait until it is
        // consumed.  Puts past the last group wait for the whole stream
        // (the solver's final count), never for more.
        const int need = min((g + 1) * kGroupBytes - kWsRing, nch * kWsChunk);
        while (lds_load_i(&ctl[1]) < need) __builtin_amdgcn_s_sleep(1);
        lds_acquire();  // the solver's reads of that half have returned
        uint4* half = ring + (g & 1) * (kGroupBytes / 16);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kWsGroup; ++u) half[u * (kWsChunk / 16) + lane] = dq[u];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kWsGroup; ++u) dq[u] = src((g + kWsFlight) * kWsGroup + u);
        __builtin_amdgcn_sched_barrier(0);
        lds_release();
        if (lane == 0) lds_store_i(&ctl[0], min((g + 1) * kWsGroup, nch) * kWsChunk);
    };
    // every put of an iteration runs (a put past the last group rewrites a
    // ring half the solver has consumed and does not move the loaded
    // count): with no path that skips one, the compiler's count of the
    // loads in flight at the loop head stays exact
    for (int g = 0; g < ngroups; g += kWsFlight) {
#pragma unroll
        for (int q = 0; q < kWsFlight; ++q) put(g + q, d[q]);
    }
}

// Waves 2 and 3: import the block's out-of-block dependencies into their
// slots (both write the same value when both see it).
// imports: the block's (face slot, first record) list, staged in LDS, so a
// lane moves on to its next import without a memory round trip.
__device__ __forceinline__ void ws_import(const WsBlock& bk, const int2* imports, double* imp,
                                          const double* face, const int* ctl, const Watchdog& wd,
                                          bool& timed_out) {
    const int lane = threadIdx.x & 63;
    const int n = bk.i1 - bk.i0;
    int idx[kWsImp];
    int2 ent[kWsImp];
#pragma unroll
    for (int k = 0; k < kWsImp; ++k) {
        idx[k] = lane + 64 * k;
        ent[k] = idx[k] < n ? imports[idx[k]] : make_int2(0, 0);
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
                ent[k] = idx[k] < n ? imports[idx[k]] : make_int2(0, 0);
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
// slot holding 0.0 (padding entries), the import list (max_imp int2), then
// the ring.  reset (optional):
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
    int2* imp_ent = reinterpret_cast<int2*>(ws_lds + ((max_rows + max_imp + 1 + 1) & ~1));
    char* ring = reinterpret_cast<char*>(imp_ent) + ws_align16(max_imp * 8);
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
    for (int i = tid; i < bk.i1 - bk.i0; i += kWsThreads) {
        imp[i] = __longlong_as_double(kPending);
        imp_ent[i] = imports[bk.i0 + i];
    }
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
        ws_solve<G, S, T, UPPER>(bk, ring, ctl, xl, out, face, n, nface, wd, timed_out, bad);
        __builtin_amdgcn_s_setprio(0);
    }
    else if (wave == 1) ws_load(bk, stream, reinterpret_cast<uint4*>(ring), ctl);
    else {
        // the second importer starts about half a poll round later
        if (wave == 3) __builtin_amdgcn_s_sleep(10);
        ws_import(bk, imp_ent, imp, face, ctl, wd, timed_out);
    }
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
