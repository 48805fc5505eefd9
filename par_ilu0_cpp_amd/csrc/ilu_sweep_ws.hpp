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
//     into the LDS ring in 1 KiB chunks, 16 chunks in flight, flow-
//     controlled by two LDS counters (bytes loaded / bytes consumed).
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
constexpr int kWsInflight = 16;   // loader chunks in flight
constexpr int kWsImp = 4;         // imports each importer lane polls at once
constexpr int kWsLook = 8;        // records ahead of the solver the importer polls

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

// Wave 0: walk the block's records.
template <int G, int S, typename T, bool UPPER>
__device__ void ws_solve(const WsBlock& bk, const char* ring, int* ctl, double* xl, double* out, double* face,
                         const Watchdog& wd, bool& timed_out) {
    const int lane = threadIdx.x & 63;
    const int j = lane / G;
    const int t = lane & (G - 1);
    const int bytes = (bk.s1 - bk.s0) * 16;
    int off = 0;
    int rec = 0;
    while (off < bytes) {
        if (!ws_wait_loaded(ctl, off + 16, wd)) { timed_out = true; break; }
        const int base = off & (kWsRing - 1);
        const int4 h = *reinterpret_cast<const int4*>(ring + base);
        const int cnt = __builtin_amdgcn_readfirstlane(h.x);
        const int w = __builtin_amdgcn_readfirstlane(h.y);
        const int rbytes = __builtin_amdgcn_readfirstlane(h.z);
        const int slot0 = __builtin_amdgcn_readfirstlane(h.w);
        if (cnt <= 0) {
            // padding: up to a ring boundary (bytes > 0) or to the block's end
            if (rbytes <= 0) break;
            off += rbytes;
            lds_release();
            if (lane == 0) lds_store_i(&ctl[1], off);
            continue;
        }
        if (!ws_wait_loaded(ctl, off + rbytes, wd)) { timed_out = true; break; }
        const int rows_b = base + 16;
        const int cols_b = rows_b + ws_align16(cnt * 8);
        const int vals_b = cols_b + ws_align16(cnt * w * 4);
        const int diag_b = vals_b + ws_align16(cnt * w * static_cast<int>(sizeof(T)));
        const bool act = j < cnt;
        const int jj = act ? j : 0;
        const int* cp = reinterpret_cast<const int*>(ring + cols_b) + jj * w;
        const T* vp = reinterpret_cast<const T*>(ring + vals_b) + jj * w;
        double acc = 0.0;
        if (act && t == 0) acc = xl[slot0 + j];
        if (act) {
            auto term = [&](int q) {
                const int c = cp[q];
                const double v = static_cast<double>(vp[q]);
                uint64_t bits = lds_load_u64(xl + c);
                while (bits == kPending) {
                    if (timed_out || wd.expired()) { timed_out = true; bits = kQuietNan; break; }
                    __builtin_amdgcn_s_sleep(0);
                    bits = lds_load_u64(xl + c);
                }
                acc = fma(-__longlong_as_double(bits), v, acc);
            };
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (s * G < w) term(t + s * G);
            for (int q = t + S * G; q < w; q += G) term(q);
        }
        const double sum = group_sum<G>(acc);
        if (act && t == 0) {
            const int2 rf = reinterpret_cast<const int2*>(ring + rows_b)[j];
            double res = sum;
            if (UPPER) res = sum / static_cast<double>(reinterpret_cast<const T*>(ring + diag_b)[j]);
            uint64_t rb = __double_as_longlong(res);
            if (res != res) rb = kQuietNan;
            lds_store_u64(xl + slot0 + j, rb);
            out[rf.x] = __longlong_as_double(rb);
            if (rf.y >= 0) publish_raw(reinterpret_cast<uint64_t*>(face) + rf.y, rb);
        }
        off += rbytes;
        ++rec;
        lds_release();  // this record's ring reads have returned
        if (lane == 0) {
            lds_store_i(&ctl[1], off);
            lds_store_i(&ctl[2], rec);
        }
    }
    if (lane == 0) lds_store_i(&ctl[2], 1 << 30);  // the importer may stop waiting for progress
}

// Wave 1: stream the block's records into the ring.
__device__ __forceinline__ void ws_load(const WsBlock& bk, const uint4* __restrict__ stream, uint4* ring, int* ctl,
                                        const Watchdog& wd, bool& timed_out) {
    const int lane = threadIdx.x & 63;
    const int nch = (bk.s1 - bk.s0) / (kWsChunk / 16);
    if (nch <= 0) return;
    const long src0 = bk.s0;
    auto src = [&](int c) -> uint4 {
        const int cc = c < nch ? c : nch - 1;
        return stream[src0 + static_cast<long>(cc) * (kWsChunk / 16) + lane];
    };
    uint4 data[kWsInflight];
#pragma unroll
    for (int u = 0; u < kWsInflight; ++u) data[u] = src(u);
    for (int c0 = 0; c0 < nch; c0 += kWsInflight) {
#pragma unroll
        for (int u = 0; u < kWsInflight; ++u) {
            const int c = c0 + u;
            if (c < nch) {
                const int need = (c + 1) * kWsChunk - kWsRing;
                while (lds_load_i(&ctl[1]) < need) {
                    if (timed_out || wd.expired()) { timed_out = true; return; }
                    __builtin_amdgcn_s_sleep(0);
                }
                lds_acquire();  // the solver's reads of these ring bytes have returned
                ring[((c * kWsChunk) & (kWsRing - 1)) / 16 + lane] = data[u];
                lds_release();
                if (lane == 0) lds_store_i(&ctl[0], (c + 1) * kWsChunk);
                data[u] = src(c + kWsInflight);
            }
        }
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
            __builtin_amdgcn_s_sleep(1);
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
                                                        uint64_t wd_limit, int max_rows, int max_imp) {
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
    bool timed_out = false;
    const int wave = tid >> 6;
    if (wave == 0) ws_solve<G, S, T, UPPER>(bk, ring, ctl, xl, out, face, wd, timed_out);
    else if (wave == 1) ws_load(bk, stream, reinterpret_cast<uint4*>(ring), ctl, wd, timed_out);
    else ws_import(bk, imports, imp, face, ctl, wd, timed_out);
    if (timed_out) atomicOr(err, kErrTimeout);
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
