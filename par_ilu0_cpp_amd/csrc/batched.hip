// Batched ILU(k)-GMRES(m): many independent systems, several right-hand
// sides each, solved together.  No counterpart in the reference, which
// solves one system per call (src/gmres.cpp:138-288); every (system, RHS)
// pair here runs exactly that iteration, with its own Krylov basis,
// Hessenberg matrix, tolerance and step count.
//
// Layout.  The systems are one block-diagonal CSR matrix (system s owns rows
// sys_ptr[s] .. sys_ptr[s+1]); ILU(k) of a block-diagonal matrix is the
// per-block ILU(k), so one symbolic pass, one factorization launch and one
// launch per sweep serve every system.  Vectors are stored RHS-interleaved,
// X[row * K + r], so the K values of a row are one 64-byte line at K = 8.
//
// Work decomposition.  Rows are cut into chunks of at most kChunkRows that
// never cross a system boundary; reductions are per chunk and per RHS, and
// an instance's total is the fixed-order sum over its system's chunks.  The
// scalar part (step condition, Givens, back-solve) runs one thread per
// instance.  One restart cycle is a fixed kernel sequence, captured once in
// a HIP graph and replayed until every instance has finished.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "ilu_internal.hpp"
#include "par_ilu_ex.h"

namespace pilu {

namespace {

constexpr int kChunkRows = 2048;
constexpr int kBThreads = 256;

struct BCtl {
    int its, maxits, cycle_active, step_active, i, status, breakdown, done;
    double tol, beta, inv, pad;
};

// ---- multi-RHS sweep: one workgroup per system ----
// Workgroup s walks the (level, row) groups of system s in order (perm,
// sys_grp, grp_off from block_level_schedule with the systems as blocks);
// the K right-hand sides of a row are K adjacent lanes.  Rows of one level
// are independent and the levels are separated by a workgroup barrier, so
// no cross-workgroup waits are needed.  Each lane sums its row in entry
// order from b, one fma per entry, as the single-system sweep with one lane
// per row does.  A system with no instance running the current Arnoldi step
// returns at once.
//
// Software pipeline: everything a row needs except the x of its
// dependencies (its position, bounds, first kSsE columns and values, b and
// the diagonal) comes from the pattern, the factors or the input vector, so
// it is requested one level ahead for the first kSsSlots row slots of every
// thread; a level then waits only for its x reads, which hit the lines this
// workgroup wrote before the barrier.  Further slots and entries are read
// directly.
constexpr int kSThreads = 1024;
constexpr int kSsE = 4;
constexpr int kSsSlots = 2;

template <typename T>
struct SsRow {
    int row, e, end;
    int c[kSsE];
    T v[kSsE];
    T dv;
    double bv;
};

template <int K, typename T, bool UPPER>
__device__ __forceinline__ void ss_fetch(int p, int p1, int r, const int* __restrict__ perm, const int* __restrict__ rp,
                                         const int* __restrict__ ci, const int* __restrict__ dg,
                                         const T* __restrict__ val, const double* b, SsRow<T>& t) {
    t.row = -1;
    t.e = t.end = 0;
    if (p >= p1) return;
    t.row = perm[p];
    const int d = dg[t.row];
    t.e = UPPER ? d + 1 : rp[t.row];
    t.end = UPPER ? rp[t.row + 1] : d;
#pragma unroll
    for (int q = 0; q < kSsE; ++q)
        if (t.e + q < t.end) {
            t.c[q] = ci[t.e + q];
            t.v[q] = val[t.e + q];
        }
    if (UPPER) t.dv = val[d];
    t.bv = b[static_cast<long>(t.row) * K + r];
}

template <int K, typename T, bool UPPER>
__device__ __forceinline__ void ss_solve(const SsRow<T>& t, int r, const int* __restrict__ ci,
                                         const T* __restrict__ val, double* out) {
    if (t.row < 0) return;
    double acc = t.bv;
#pragma unroll
    for (int q = 0; q < kSsE; ++q)
        if (t.e + q < t.end) acc = fma(-out[static_cast<long>(t.c[q]) * K + r], static_cast<double>(t.v[q]), acc);
    for (int e = t.e + kSsE; e < t.end; ++e)
        acc = fma(-out[static_cast<long>(ci[e]) * K + r], static_cast<double>(val[e]), acc);
    out[static_cast<long>(t.row) * K + r] = UPPER ? acc / static_cast<double>(t.dv) : acc;
}

template <int K, typename T, bool UPPER>
__global__ __launch_bounds__(kSThreads) void k_ssweep(const int* __restrict__ sys_grp, const int* __restrict__ grp_off,
                                                      const int* __restrict__ perm, const int* __restrict__ rp,
                                                      const int* __restrict__ ci, const int* __restrict__ dg,
                                                      const T* __restrict__ val, const double* b, double* out,
                                                      const BCtl* ctl) {
    constexpr int kRowsPerPass = kSThreads / K;
    const int s = blockIdx.x;
    if (ctl) {
        bool any = false;
#pragma unroll
        for (int r = 0; r < K; ++r) any = any || ctl[s * K + r].step_active != 0;
        if (!any) return;
    }
    const int r = threadIdx.x % K;
    const int slot = static_cast<int>(threadIdx.x) / K;
    const int g0 = sys_grp[s], g1 = sys_grp[s + 1];
    SsRow<T> nxt[kSsSlots];
#pragma unroll
    for (int k = 0; k < kSsSlots; ++k) nxt[k].row = -1;
    if (g0 < g1) {
        const int q0 = grp_off[g0], q1 = grp_off[g0 + 1];
#pragma unroll
        for (int k = 0; k < kSsSlots; ++k)
            ss_fetch<K, T, UPPER>(q0 + slot + k * kRowsPerPass, q1, r, perm, rp, ci, dg, val, b, nxt[k]);
    }
    for (int g = g0; g < g1; ++g) {
        const int p0 = grp_off[g], p1 = grp_off[g + 1];
        SsRow<T> cur[kSsSlots];
#pragma unroll
        for (int k = 0; k < kSsSlots; ++k) cur[k] = nxt[k];
        if (g + 1 < g1) {
            const int n1 = grp_off[g + 2];
#pragma unroll
            for (int k = 0; k < kSsSlots; ++k)
                ss_fetch<K, T, UPPER>(p1 + slot + k * kRowsPerPass, n1, r, perm, rp, ci, dg, val, b, nxt[k]);
        }
#pragma unroll
        for (int k = 0; k < kSsSlots; ++k) ss_solve<K, T, UPPER>(cur[k], r, ci, val, out);
        for (int p = p0 + slot + kSsSlots * kRowsPerPass; p < p1; p += kRowsPerPass) {
            SsRow<T> t;
            ss_fetch<K, T, UPPER>(p, p1, r, perm, rp, ci, dg, val, b, t);
            ss_solve<K, T, UPPER>(t, r, ci, val, out);
        }
        __syncthreads();
    }
}

// Sum over the threads of a block that share t % K; result valid in
// threads t < K (one per RHS).  Fixed order.
template <int K>
__device__ double rhs_reduce(double v, double* s_red) {
#pragma unroll
    for (int off = K; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane < K) s_red[wave * K + lane] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x < K)
        for (int w = 0; w < kBThreads / 64; ++w) t += s_red[w * K + threadIdx.x];
    return t;
}

// Instance total of chunk partials part[c*K + r] over the system's chunks.
__device__ __forceinline__ double system_sum(const double* part, const int* sys_chunk, int s, int r, int K) {
    double t = 0.0;
    for (int c = sys_chunk[s]; c < sys_chunk[s + 1]; ++c) t += part[static_cast<long>(c) * K + r];
    return t;
}

// True when some instance of system s still needs this chunk: with
// step == true, an instance running the current Arnoldi step; otherwise an
// instance not yet finished.  Uniform over the block, so a chunk of a system
// whose right-hand sides have all converged costs one small read.
template <int K>
__device__ __forceinline__ bool system_live(const BCtl* ctl, int s, bool step) {
    bool any = false;
#pragma unroll
    for (int r = 0; r < K; ++r) {
        const BCtl& q = ctl[s * K + r];
        any = any || (step ? q.step_active != 0 : q.done == 0);
    }
    return any;
}

// Y = alpha A X + beta Yin on one chunk per block; part[c*K+r] = partial of
// dotv.Y (dotv != null) or Y.Y (nrm).  With ctl, chunks of systems that
// need no work (system_live) are skipped.
template <int K>
__global__ __launch_bounds__(kBThreads) void k_bspmv(const int* __restrict__ chunk_beg, const int* __restrict__ rp,
                                                     const int* __restrict__ ci, const double* __restrict__ val,
                                                     double alpha, const double* __restrict__ x, double beta,
                                                     const double* yin, double* y, const double* __restrict__ dotv,
                                                     bool nrm, double* part, const int* __restrict__ chunk_sys,
                                                     const BCtl* ctl, bool step) {
    __shared__ double s_red[(kBThreads / 64) * K];
    const int c = blockIdx.x;
    if (ctl && !system_live<K>(ctl, chunk_sys[c], step)) return;
    const int r = threadIdx.x % K;
    const int rb = chunk_beg[c], re = chunk_beg[c + 1];
    double acc = 0.0;
    for (int row = rb + threadIdx.x / K; row < re; row += kBThreads / K) {
        double s = 0.0;
        for (int e = rp[row]; e < rp[row + 1]; ++e) s = fma(val[e], x[static_cast<long>(ci[e]) * K + r], s);
        const long o = static_cast<long>(row) * K + r;
        const double v = beta == 0.0 ? alpha * s : beta * yin[o] + alpha * s;
        y[o] = v;
        if (dotv) acc = fma(dotv[o], v, acc);
        else if (nrm) acc = fma(v, v, acc);
    }
    if (part) {
        const double t = rhs_reduce<K>(acc, s_red);
        if (threadIdx.x < K) part[static_cast<long>(c) * K + threadIdx.x] = t;
    }
}

// MGS term j of step i for every instance: h = (v_j, w); w -= h v_j; next partials.
template <int K>
__global__ __launch_bounds__(kBThreads) void k_bmgs(const int* __restrict__ chunk_beg, const int* __restrict__ chunk_sys,
                                                    const int* __restrict__ sys_chunk, const double* __restrict__ vj,
                                                    double* w, const double* __restrict__ vnext,
                                                    const double* part_in, double* part_out, double* hh, long hstride,
                                                    int hpos, const BCtl* ctl) {
    __shared__ double s_red[(kBThreads / 64) * K];
    const int c = blockIdx.x;
    const int s = chunk_sys[c];
    if (!system_live<K>(ctl, s, true)) return;
    const int r = threadIdx.x % K;
    const double h = system_sum(part_in, sys_chunk, s, r, K);
    if (c == sys_chunk[s] && threadIdx.x < K) hh[(static_cast<long>(s) * K + r) * hstride + hpos] = h;
    const int rb = chunk_beg[c], re = chunk_beg[c + 1];
    double acc = 0.0;
    for (long o = static_cast<long>(rb) * K + threadIdx.x; o < static_cast<long>(re) * K; o += kBThreads) {
        const double wr = fma(-h, vj[o], w[o]);
        w[o] = wr;
        acc = fma(vnext ? vnext[o] : wr, wr, acc);
    }
    const double t = rhs_reduce<K>(acc, s_red);
    if (threadIdx.x < K) part_out[static_cast<long>(c) * K + threadIdx.x] = t;
}

// part[c*K+r] = partial sum of b^2 over chunk c, RHS r.
template <int K>
__global__ __launch_bounds__(kBThreads) void k_bnorm(const int* __restrict__ chunk_beg, const double* __restrict__ b,
                                                     double* part) {
    __shared__ double s_red[(kBThreads / 64) * K];
    const int c = blockIdx.x;
    double acc = 0.0;
    for (long o = static_cast<long>(chunk_beg[c]) * K + threadIdx.x; o < static_cast<long>(chunk_beg[c + 1]) * K;
         o += kBThreads)
        acc = fma(b[o], b[o], acc);
    const double t = rhs_reduce<K>(acc, s_red);
    if (threadIdx.x < K) part[static_cast<long>(c) * K + threadIdx.x] = t;
}

// v *= inv per instance (mode 0: cycle start, mode 1: after a step).
template <int K>
__global__ __launch_bounds__(kBThreads) void k_bscale(const int* __restrict__ chunk_beg, const int* __restrict__ chunk_sys,
                                                      double* v, const BCtl* ctl, int mode) {
    const int c = blockIdx.x;
    const int s = chunk_sys[c];
    const int r = threadIdx.x % K;
    const BCtl& q = ctl[s * K + r];
    const bool on = mode == 0 ? q.cycle_active != 0 : (q.step_active != 0 && q.breakdown == 0);
    if (!on) return;
    const double a = q.inv;
    for (long o = static_cast<long>(chunk_beg[c]) * K + threadIdx.x; o < static_cast<long>(chunk_beg[c + 1]) * K;
         o += kBThreads)
        v[o] *= a;
}

// sol += sum_j rs_j z_j per instance.
template <int K>
__global__ __launch_bounds__(kBThreads) void k_bupdate(const int* __restrict__ chunk_beg, const int* __restrict__ chunk_sys,
                                                       const double* __restrict__ z, long nk, const BCtl* ctl,
                                                       const double* hh, long hstride, int m, double* sol) {
    const int c = blockIdx.x;
    const int s = chunk_sys[c];
    const int r = threadIdx.x % K;
    const BCtl& q = ctl[s * K + r];
    if (q.status == 2 || q.i < 0) return;
    const double* rs = hh + (static_cast<long>(s) * K + r) * hstride + static_cast<long>(m + 1) * m + 2 * (m + 1);
    for (long o = static_cast<long>(chunk_beg[c]) * K + threadIdx.x; o < static_cast<long>(chunk_beg[c + 1]) * K;
         o += kBThreads) {
        double x = sol[o];
        for (int j = 0; j <= q.i; ++j) x = fma(rs[j], z[j * nk + o], x);
        sol[o] = x;
    }
}

// ---- one thread per instance ----
__global__ void k_bcycle_begin(int ni, int K, BCtl* ctl, const double* part, const int* sys_chunk, double* hh,
                               long hstride, int m) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= ni) return;
    BCtl& q = ctl[inst];
    q.i = -1;
    q.breakdown = 0;
    q.step_active = 0;
    q.cycle_active = 0;
    if (q.done) return;
    const double beta = std::sqrt(system_sum(part, sys_chunk, inst / K, inst % K, K));
    q.beta = beta;
    if (beta <= q.tol) {
        q.status = 1;
        q.done = 1;
        return;
    }
    if (std::isnan(beta)) {
        q.status = 2;
        q.done = 1;
        return;
    }
    q.inv = 1.0 / beta;
    hh[static_cast<long>(inst) * hstride + static_cast<long>(m + 1) * m + 2 * (m + 1)] = beta;
    q.cycle_active = 1;
}

__global__ void k_bgate(int ni, BCtl* ctl, int i) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= ni) return;
    BCtl& q = ctl[inst];
    if (q.cycle_active && q.beta > q.tol) {
        const int before = q.its++;
        if (before < q.maxits) {
            q.step_active = 1;
            q.i = i;
            return;
        }
    }
    q.step_active = 0;
    q.cycle_active = 0;
}

__global__ void k_bfinish(int ni, int K, BCtl* ctl, const double* part, const int* sys_chunk, double* hh, long hstride,
                          int m) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= ni) return;
    BCtl& q = ctl[inst];
    if (!q.step_active) return;
    const double t = std::sqrt(system_sum(part, sys_chunk, inst / K, inst % K, K));
    const int im1 = m + 1, i = q.i, i1 = i + 1;
    double* base = hh + static_cast<long>(inst) * hstride;
    double* col = base + static_cast<long>(i) * im1;
    double* cc = base + static_cast<long>(im1) * m;
    double* ss = cc + im1;
    double* rs = ss + im1;
    col[i1] = t;
    if (t <= 1e-16) {
        q.breakdown = 1;
        q.cycle_active = 0;
        return;
    }
    q.inv = 1.0 / t;
    for (int k = 1; k <= i; ++k) {
        const int k1 = k - 1;
        const double tk = col[k1];
        col[k1] = cc[k1] * tk + ss[k1] * col[k];
        col[k] = -ss[k1] * tk + cc[k1] * col[k];
    }
    const double a = col[i], b = col[i1];
    double gam = std::sqrt(a * a + b * b);
    if (gam == 0.0) gam = 1e-16;
    cc[i] = col[i] / gam;
    ss[i] = col[i1] / gam;
    rs[i1] = -ss[i] * rs[i];
    rs[i] = cc[i] * rs[i];
    col[i] = cc[i] * col[i] + ss[i] * col[i1];
    q.beta = std::fabs(rs[i1]);
    if (std::isnan(q.beta)) {
        q.status = 2;
        q.done = 1;
        q.cycle_active = 0;
    }
}

// Back-solve and end-of-cycle decision (src/gmres.cpp:265-283).
__global__ void k_bbacksolve(int ni, BCtl* ctl, double* hh, long hstride, int m) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= ni) return;
    BCtl& q = ctl[inst];
    if (q.done || q.i < 0) return;
    const int im1 = m + 1, i = q.i;
    double* base = hh + static_cast<long>(inst) * hstride;
    double* rs = base + static_cast<long>(im1) * m + 2 * im1;
    rs[i] = rs[i] / base[static_cast<long>(i) * im1 + i];
    for (int ii = i - 1; ii >= 0; --ii) {
        double t = rs[ii];
        for (int j = ii + 1; j <= i; ++j) t -= base[static_cast<long>(j) * im1 + ii] * rs[j];
        rs[ii] = t / base[static_cast<long>(ii) * im1 + ii];
    }
}

__global__ void k_bend_cycle(int ni, BCtl* ctl, int* unfinished) {
    // single block: decide per instance, then count the unfinished ones
    __shared__ int s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    int cnt = 0;
    for (int inst = threadIdx.x; inst < ni; inst += blockDim.x) {
        BCtl& q = ctl[inst];
        if (!q.done) {
            if (q.beta <= q.tol) { q.status = 1; q.done = 1; }
            else if (q.its > q.maxits) { q.status = 3; q.done = 1; }
            else if (q.its >= q.maxits) { q.status = 4; q.done = 1; }  // exhausted at a restart boundary
        }
        cnt += q.done ? 0 : 1;
    }
    atomicAdd(&s_cnt, cnt);
    __syncthreads();
    if (threadIdx.x == 0) *unfinished = s_cnt;
}

__global__ void k_binit(int ni, BCtl* ctl, const double* part, const int* sys_chunk, int K, double rel_tol, int maxits) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= ni) return;
    BCtl q{};
    q.maxits = maxits;
    q.tol = rel_tol * std::sqrt(system_sum(part, sys_chunk, inst / K, inst % K, K));
    ctl[inst] = q;
}

__global__ void k_bresult(int ni, int K, const BCtl* ctl, const double* part, const double* part_b, const int* sys_chunk,
                          int* its, int* conv, double* rel) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= ni) return;
    const BCtl& q = ctl[inst];
    its[inst] = q.its;
    conv[inst] = (q.status == 1 || q.status == 4) ? 1 : 0;
    const double rb = std::sqrt(system_sum(part_b, sys_chunk, inst / K, inst % K, K));
    const double rr = std::sqrt(system_sum(part, sys_chunk, inst / K, inst % K, K));
    rel[inst] = rb > 0 ? rr / rb : rr;
}

}  // namespace

class BatchedGmres {
public:
    ~BatchedGmres() {
        drop_graph();
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    int setup(const CsrMatrix* d_a, int lof, int nsys, const int* sys_ptr, int nrhs) {
        if (nrhs != 1 && nrhs != 2 && nrhs != 4 && nrhs != 8) throw std::runtime_error("nrhs must be 1, 2, 4 or 8");
        if (nsys <= 0 || sys_ptr[0] != 0 || sys_ptr[nsys] != d_a->size) throw std::runtime_error("bad sys_ptr");
        if (!stream_) PILU_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        drop_graph();
        K_ = nrhs;
        nsys_ = nsys;
        n_ = d_a->size;
        // chunks
        std::vector<int> beg, sys, sc(nsys + 1);
        for (int s = 0; s < nsys; ++s) {
            if (sys_ptr[s + 1] <= sys_ptr[s]) throw std::runtime_error("empty system");
            sc[s] = static_cast<int>(beg.size());
            for (int r = sys_ptr[s]; r < sys_ptr[s + 1]; r += kChunkRows) {
                beg.push_back(r);
                sys.push_back(s);
            }
        }
        sc[nsys] = static_cast<int>(beg.size());
        nchunks_ = static_cast<int>(beg.size());
        beg.push_back(n_);
        chunk_beg_.resize(beg.size());
        chunk_sys_.resize(sys.size());
        sys_chunk_.resize(sc.size());
        PILU_HIP(hipMemcpy(chunk_beg_.get(), beg.data(), beg.size() * sizeof(int), hipMemcpyHostToDevice));
        PILU_HIP(hipMemcpy(chunk_sys_.get(), sys.data(), sys.size() * sizeof(int), hipMemcpyHostToDevice));
        PILU_HIP(hipMemcpy(sys_chunk_.get(), sc.data(), sc.size() * sizeof(int), hipMemcpyHostToDevice));
        // system of every row (the sweeps' schedule is per system)
        std::vector<int> rs(static_cast<size_t>(n_));
        for (int q = 0; q < nsys; ++q) std::fill(rs.begin() + sys_ptr[q], rs.begin() + sys_ptr[q + 1], q);
        row_sys_.resize(rs.size());
        PILU_HIP(hipMemcpy(row_sys_.get(), rs.data(), rs.size() * sizeof(int), hipMemcpyHostToDevice));
        // pattern: symbolic ILU(lof) of the block-diagonal matrix
        const int nnz_f = ilu_.SymbolicDevice(n_, d_a->row_ptr, d_a->col_idx, lof);
        if (nnz_f < 0) return nnz_f;
        // per-system level schedules of the sweeps
        Ext* ie = ilu_.DeviceState();
        block_level_schedule(ie, false, row_sys_.get(), nsys, sperm_l_, sgrp_l_, soff_l_);
        block_level_schedule(ie, true, row_sys_.get(), nsys, sperm_u_, sgrp_u_, soff_u_);
        PILU_HIP(hipDeviceSynchronize());
        ready_ = false;
        return nnz_f;
    }

    bool factorize(const double* d_avals) {
        PILU_HIP(hipDeviceSynchronize());
        const bool ok = ilu_.FactorizeDeviceFromA(d_avals);
        PILU_HIP(hipDeviceSynchronize());
        ready_ = true;
        return ok;
    }

    void set_precision(int bits) { ilu_.SetFactorPrecision(bits); drop_graph(); }

    int solve(const CsrMatrix* d_a, const double* d_rhs, double* d_sol, double rel_tol, int m, int maxits, int* its,
              int* conv, double* rel) {
        if (!ready_) throw std::runtime_error("batched: factorize first");
        if (d_a->size != n_) throw std::runtime_error("batched: matrix size changed");
        PILU_HIP(hipDeviceSynchronize());  // caller's null-stream work is complete
        const int ni = nsys_ * K_;
        const long nk = static_cast<long>(n_) * K_;
        const long hstride = static_cast<long>(m + 1) * (m + 3);
        bool realloc = m != m_ || maxits != maxits_;
        v_.reserve(static_cast<size_t>(m + 1) * nk);
        z_.reserve(static_cast<size_t>(m) * nk);
        hh_.reserve(static_cast<size_t>(ni) * hstride);
        part_.reserve(3 * static_cast<size_t>(nchunks_) * K_);
        ctl_.reserve(ni);
        y_.reserve(nk);
        flags_.reserve(4);
        res_.reserve(nk);
        int_out_.reserve(2 * static_cast<size_t>(ni));
        dbl_out_.reserve(ni);
        if (v_.get() != gv_ || z_.get() != gz_ || d_rhs != grhs_ || d_sol != gsol_ || d_a->value != gval_ ||
            d_a->row_ptr != grp_ || d_a->col_idx != gci_ || realloc)
            drop_graph();
        m_ = m;
        maxits_ = maxits;
        Ext* ie = ilu_.DeviceState();
        // tolerances: tol = rel_tol * ||b_inst||
        double* pb = part_.get() + 2L * nchunks_ * K_;
        norm_launch(d_rhs, pb);
        k_binit<<<ceil_div(ni, 256), 256, 0, stream_>>>(ni, ctl_.get(), pb, sys_chunk_.get(), K_, rel_tol, maxits);
        PILU_HIP(hipGetLastError());
        if (!graph_exec_) {
            PILU_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            enqueue_cycle(d_a, d_rhs, d_sol, ie, m, hstride, nk);
            hipGraph_t g;
            PILU_HIP(hipStreamEndCapture(stream_, &g));
            PILU_HIP(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0));
            PILU_HIP(hipGraphDestroy(g));
            gv_ = v_.get(); gz_ = z_.get(); grhs_ = d_rhs; gsol_ = d_sol; gval_ = d_a->value;
            grp_ = d_a->row_ptr; gci_ = d_a->col_idx;
        }
        int unfinished = ni;
        cycles_ = 0;
        while (unfinished > 0) {
            PILU_HIP(hipGraphLaunch(graph_exec_, stream_));
            PILU_HIP(hipMemcpyAsync(&unfinished, flags_.get(), sizeof(int), hipMemcpyDeviceToHost, stream_));
            PILU_HIP(hipStreamSynchronize(stream_));
            ++cycles_;
            if (cycles_ > maxits + 2) throw std::runtime_error("batched GMRES: no progress");
        }
        int err = 0;
        PILU_HIP(hipMemcpy(&err, ie->err.get() + 1, sizeof(int), hipMemcpyDeviceToHost));
        if (err & kErrTimeout) throw std::runtime_error("batched sweep: dependency wait timed out (watchdog)");
        // results: true residual per instance
        double* pr = part_.get();
        spmv_launch(d_a, -1.0, d_sol, 1.0, d_rhs, res_.get(), nullptr, true, pr);
        k_bresult<<<ceil_div(ni, 256), 256, 0, stream_>>>(ni, K_, ctl_.get(), pr, pb, sys_chunk_.get(), int_out_.get(),
                                                          int_out_.get() + ni, dbl_out_.get());
        PILU_HIP(hipMemcpyAsync(its, int_out_.get(), sizeof(int) * ni, hipMemcpyDeviceToHost, stream_));
        PILU_HIP(hipMemcpyAsync(conv, int_out_.get() + ni, sizeof(int) * ni, hipMemcpyDeviceToHost, stream_));
        PILU_HIP(hipMemcpyAsync(rel, dbl_out_.get(), sizeof(double) * ni, hipMemcpyDeviceToHost, stream_));
        PILU_HIP(hipStreamSynchronize(stream_));
        return cycles_;
    }

    lljbash::IluSolver& ilu() { return ilu_; }
    int cycles() const { return cycles_; }

private:
    using Ext = lljbash::IluSolver::Ext;

    void drop_graph() {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
    }

    // gate: 0 every chunk, 1 chunks of unfinished systems, 2 chunks of systems running the current step
    void spmv_launch(const CsrMatrix* a, double alpha, const double* x, double beta, const double* yin, double* y,
                     const double* dotv, bool nrm, double* part, int gate = 0) {
        const BCtl* ctl = gate ? ctl_.get() : nullptr;
        switch (K_) {
#define PILU_CASE(KK)                                                                                                 \
    case KK:                                                                                                          \
        k_bspmv<KK><<<nchunks_, kBThreads, 0, stream_>>>(chunk_beg_.get(), a->row_ptr, a->col_idx, a->value, alpha, \
                                                          x, beta, yin, y, dotv, nrm, part, chunk_sys_.get(), ctl,    \
                                                          gate == 2);                                                 \
        break;
            PILU_CASE(1) PILU_CASE(2) PILU_CASE(4) PILU_CASE(8)
#undef PILU_CASE
        }
    }

    // partials of ||b||^2 per chunk and RHS
    void norm_launch(const double* b, double* part) {
        switch (K_) {
#define PILU_CASE(KK)                                                                                       \
    case KK:                                                                                                \
        k_bnorm<KK><<<nchunks_, kBThreads, 0, stream_>>>(chunk_beg_.get(), b, part);                        \
        break;
            PILU_CASE(1) PILU_CASE(2) PILU_CASE(4) PILU_CASE(8)
#undef PILU_CASE
        }
    }

    template <int K, typename T>
    void sweeps_t(Ext* e, const T* val, const double* b, double* x) {
        k_ssweep<K, T, false><<<nsys_, kSThreads, 0, stream_>>>(sgrp_l_.get(), soff_l_.get(), sperm_l_.get(),
                                                                e->rp.get(), e->ci.get(), e->dg.get(), val, b,
                                                                y_.get(), ctl_.get());
        k_ssweep<K, T, true><<<nsys_, kSThreads, 0, stream_>>>(sgrp_u_.get(), soff_u_.get(), sperm_u_.get(),
                                                               e->rp.get(), e->ci.get(), e->dg.get(), val, y_.get(),
                                                               x, ctl_.get());
    }

    template <int K>
    void sweeps_k(Ext* e, const double* b, double* x) {
        if (e->precision == 32) sweeps_t<K, float>(e, e->fac32.get(), b, x);
        else sweeps_t<K, double>(e, e->fac.get(), b, x);
    }

    void sweeps(Ext* e, const double* b, double* x) {
        switch (K_) {
            case 1: sweeps_k<1>(e, b, x); break;
            case 2: sweeps_k<2>(e, b, x); break;
            case 4: sweeps_k<4>(e, b, x); break;
            default: sweeps_k<8>(e, b, x); break;
        }
    }

    void mgs_launch(const double* vj, double* w, const double* vnext, const double* pin, double* pout, long hstride,
                    int hpos) {
        switch (K_) {
#define PILU_CASE(KK)                                                                                                \
    case KK:                                                                                                         \
        k_bmgs<KK><<<nchunks_, kBThreads, 0, stream_>>>(chunk_beg_.get(), chunk_sys_.get(), sys_chunk_.get(), vj, w, \
                                                         vnext, pin, pout, hh_.get(), hstride, hpos, ctl_.get());    \
        break;
            PILU_CASE(1) PILU_CASE(2) PILU_CASE(4) PILU_CASE(8)
#undef PILU_CASE
        }
    }

    void scale_launch(double* v, int mode) {
        switch (K_) {
#define PILU_CASE(KK)                                                                                          \
    case KK:                                                                                                   \
        k_bscale<KK><<<nchunks_, kBThreads, 0, stream_>>>(chunk_beg_.get(), chunk_sys_.get(), v, ctl_.get(), mode); \
        break;
            PILU_CASE(1) PILU_CASE(2) PILU_CASE(4) PILU_CASE(8)
#undef PILU_CASE
        }
    }

    void update_launch(double* sol, long nk, long hstride) {
        switch (K_) {
#define PILU_CASE(KK)                                                                                              \
    case KK:                                                                                                       \
        k_bupdate<KK><<<nchunks_, kBThreads, 0, stream_>>>(chunk_beg_.get(), chunk_sys_.get(), z_.get(), nk,       \
                                                            ctl_.get(), hh_.get(), hstride, m_, sol);               \
        break;
            PILU_CASE(1) PILU_CASE(2) PILU_CASE(4) PILU_CASE(8)
#undef PILU_CASE
        }
    }

    void enqueue_cycle(const CsrMatrix* a, const double* rhs, double* sol, Ext* ie, int m, long hstride, long nk) {
        const int ni = nsys_ * K_;
        const int ib = ceil_div(ni, 256);
        double* p0 = part_.get();
        double* p1 = p0 + static_cast<long>(nchunks_) * K_;
        double* v = v_.get();
        double* z = z_.get();
        spmv_launch(a, -1.0, sol, 1.0, rhs, v, nullptr, true, p0, 1);
        k_bcycle_begin<<<ib, 256, 0, stream_>>>(ni, K_, ctl_.get(), p0, sys_chunk_.get(), hh_.get(), hstride, m);
        scale_launch(v, 0);
        for (int i = 0; i < m; ++i) {
            double* vi = v + i * nk;
            double* vi1 = v + (i + 1) * nk;
            double* zi = z + i * nk;
            k_bgate<<<ib, 256, 0, stream_>>>(ni, ctl_.get(), i);
            sweeps(ie, vi, zi);
            spmv_launch(a, 1.0, zi, 0.0, nullptr, vi1, v, false, p0, 2);
            double* pin = p0;
            double* pout = p1;
            for (int j = 0; j <= i; ++j) {
                mgs_launch(v + j * nk, vi1, j < i ? v + (j + 1) * nk : nullptr, pin, pout, hstride,
                           i * (m + 1) + j);
                std::swap(pin, pout);
            }
            k_bfinish<<<ib, 256, 0, stream_>>>(ni, K_, ctl_.get(), pin, sys_chunk_.get(), hh_.get(), hstride, m);
            scale_launch(vi1, 1);
        }
        k_bbacksolve<<<ib, 256, 0, stream_>>>(ni, ctl_.get(), hh_.get(), hstride, m);
        update_launch(sol, nk, hstride);
        k_bend_cycle<<<1, 1024, 0, stream_>>>(ni, ctl_.get(), flags_.get());
    }

    lljbash::IluSolver ilu_;
    hipStream_t stream_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    const void *gv_ = nullptr, *gz_ = nullptr, *grhs_ = nullptr, *gsol_ = nullptr, *gval_ = nullptr,
               *grp_ = nullptr, *gci_ = nullptr;
    int K_ = 8, nsys_ = 0, n_ = 0, nchunks_ = 0, m_ = -1, maxits_ = -1, cycles_ = 0;
    bool ready_ = false;
    DevBuf<int> chunk_beg_, chunk_sys_, sys_chunk_, row_sys_, flags_, int_out_;
    DevBuf<int> sperm_l_, sgrp_l_, soff_l_, sperm_u_, sgrp_u_, soff_u_;  // per-system level schedules
    DevBuf<double> v_, z_, hh_, part_, y_, res_, dbl_out_;
    DevBuf<BCtl> ctl_;
};

}  // namespace pilu

// ---- C ABI (include/par_ilu_ex.h) ----
namespace {
template <typename F>
int bguard(const char* what, F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s: %s\n", what, e.what());
        return std::string(e.what()) == "missing diag" ? -2 : -1;
    }
}
pilu::BatchedGmres* B(BatchGmresHdl h) { return static_cast<pilu::BatchedGmres*>(h); }
}  // namespace

extern "C" {

BatchGmresHdl BatchGmresCreateEx(void) { return new pilu::BatchedGmres; }

void BatchGmresDestroyEx(BatchGmresHdl h) { delete B(h); }

int BatchGmresSetupEx(BatchGmresHdl h, const CsrMatrix* d_a, int lof, int nsys, const int* sys_ptr, int nrhs) {
    return bguard("BatchGmresSetup", [&] { return B(h)->setup(d_a, lof, nsys, sys_ptr, nrhs); });
}

int BatchGmresSetFactorPrecisionEx(BatchGmresHdl h, int bits) {
    return bguard("BatchGmresSetFactorPrecision", [&] { B(h)->set_precision(bits); return 0; });
}

int BatchGmresFactorizeEx(BatchGmresHdl h, const double* d_avalues) {
    return bguard("BatchGmresFactorize", [&] { return B(h)->factorize(d_avalues) ? 0 : -1; });
}

int BatchGmresSolveEx(BatchGmresHdl h, const CsrMatrix* d_a, const double* d_rhs, double* d_sol, double rel_tol,
                      int m, int maxits, int* its, int* converged, double* true_rel_res) {
    return bguard("BatchGmresSolve",
                  [&] { return B(h)->solve(d_a, d_rhs, d_sol, rel_tol, m, maxits, its, converged, true_rel_res); });
}

}  // extern "C"
