// lljbash::PreconditionedGmres on the GPU.
// Interface: include/gmres.hpp (reference include/gmres.hpp:9-47); the
// iteration is src/gmres.cpp:138-288 with all vector work on the device.
//
// Everything inside a restart cycle is enqueued without host round trips:
// the scalar part of each Arnoldi step (step condition, Givens update,
// breakdown and NaN checks) runs in single-block kernels on a small control
// block in device memory, and every vector kernel of a step is gated by the
// control block's step flag.  The host only looks at the control block one
// step behind (to stop enqueuing) and at the end of each cycle.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "gmres.hpp"
#include "ilu_internal.hpp"

namespace pilu {

constexpr int kVecBlocks = 2048;  // fixed grid of the reducing vector kernels (8 waves per CU)
constexpr int kVecThreads = 256;

struct GmresCtl {
    int its;           // Arnoldi steps so far (the reference's `its`)
    int maxits;
    int im;            // restart length m
    int cycle_active;  // inner loop still running
    int step_active;   // the current step runs
    int i;             // last step index of the current cycle
    int status;        // 0 running, 1 converged at cycle start, 2 NaN estimate
    int breakdown;     // ||w|| <= 1e-16 in the last step
    int hist_cap;
    int nhist;         // residual estimates written to the history
    double tol;
    double beta;       // latest residual estimate
    double inv;        // scale for the vector just normalised
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the block in a fixed order; every thread gets the total.
__device__ double block_sum(double v) {
    __shared__ double s_part[kVecThreads / 64];
    __shared__ double s_total;
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) t += s_part[w];
        s_total = t;
    }
    __syncthreads();
    return s_total;
}

// Sum of kVecBlocks partials, identical in every block.
__device__ double sum_partials(const double* part) {
    double v = 0.0;
    for (int k = threadIdx.x; k < kVecBlocks; k += blockDim.x) v += part[k];
    return block_sum(v);
}

// y = alpha A x + beta y (beta == 0: y not read), G lanes per row; with
// dotv, part[block] = partial sum of dotv[r] * y[r]; with nrm, partial sum
// of y[r]^2.
template <int G>
__global__ __launch_bounds__(kVecThreads) void k_spmv(int n, const int* __restrict__ rp,
                                                      const int* __restrict__ ci,
                                                      const double* __restrict__ val, double alpha,
                                                      const double* __restrict__ x, double beta,
                                                      const double* yin, double* y,
                                                      const double* __restrict__ dotv, bool nrm,
                                                      double* part, const int* gate) {
    if (gate && *gate == 0) return;
    constexpr int R = 64 / G;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int slot = lane / G;
    const int t = lane & (G - 1);
    const long step = static_cast<long>(gridDim.x) * (kVecThreads / 64) * R;
    double dacc = 0.0;
    for (long row = (static_cast<long>(blockIdx.x) * (kVecThreads / 64) + wave) * R + slot; row < n; row += step) {
        double s = 0.0;
        for (int e = rp[row] + t; e < rp[row + 1]; e += G) s = fma(val[e], x[ci[e]], s);
        s = group_sum<G>(s);
        if (t == 0) {
            const double yv = beta == 0.0 ? alpha * s : beta * yin[row] + alpha * s;
            y[row] = yv;
            if (dotv) dacc = fma(dotv[row], yv, dacc);
            else if (nrm) dacc = fma(yv, yv, dacc);
        }
    }
    if (part) {
        const double tot = block_sum(dacc);
        if (threadIdx.x == 0) part[blockIdx.x] = tot;
    }
}

// One modified Gram-Schmidt step (src/gmres.cpp:219-224):
//   h = (v_j, w) from the partials of the previous kernel;
//   w -= h v_j;  then partials of (vnext, w) or of (w, w) if vnext is null.
__global__ __launch_bounds__(kVecThreads) void k_mgs(int n, const double* __restrict__ vj, double* w,
                                                     const double* __restrict__ vnext,
                                                     const double* part_in, double* part_out,
                                                     double* hout, const int* gate) {
    if (gate && *gate == 0) return;
    const double h = sum_partials(part_in);
    if (blockIdx.x == 0 && threadIdx.x == 0) *hout = h;
    double acc = 0.0;
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    for (long r = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += stride) {
        const double wr = fma(-h, vj[r], w[r]);
        w[r] = wr;
        acc = fma(vnext ? vnext[r] : wr, wr, acc);
    }
    const double tot = block_sum(acc);
    if (threadIdx.x == 0) part_out[blockIdx.x] = tot;
}

// ---------------------------------------------------------------------
// The whole modified Gram-Schmidt of step i in one launch (src/gmres.cpp:
// 219-233): workgroup g keeps rows [g*chunk, (g+1)*chunk) of w in
// registers (R per thread) and, for j = 0..i,
//   w -= h_j v_j;  partial of (v_{j+1}, w), or of (w, w) after the last j;
// every workgroup publishes its partial as a value-flag and sums all of
// them in index order, so every workgroup holds the same h_{j+1}.  v_{j+1}
// stays in registers for the next term: each basis vector is read once.
// h_0 comes from the SpMV's kVecBlocks partials as in the per-term path.
// Finally w is scaled by 1 / ||w|| unless ||w|| <= 1e-16 (the breakdown
// of src/gmres.cpp:228-233) and ||w|| is left in *tnorm.  The grid must be
// co-resident (cooperative launch); partial slots come in two generations
// (step parity), each step re-arming the other one.  Every wait carries the
// watchdog.
// ---------------------------------------------------------------------
constexpr int kMgsThreads = 512;

__device__ double block_sum512(double v) {
    __shared__ double s_part[kMgsThreads / 64];
    __shared__ double s_total;
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kMgsThreads / 64; ++k) t += s_part[k];
        s_total = t;
    }
    __syncthreads();
    return s_total;
}

template <int R>
__global__ __launch_bounds__(kMgsThreads) void k_mgs_all(int n, int i, int m, int chunk, const double* __restrict__ vv,
                                                         double* w, const double* part0, double* pbuf, double* col,
                                                         double* tnorm, const GmresCtl* ctl, int* err,
                                                         uint64_t wd_limit) {
    if (ctl->step_active == 0) return;
    const int nwg = gridDim.x, wg = blockIdx.x;
    double* pb = pbuf + static_cast<long>(i & 1) * (m + 1) * nwg;
    double* pnext = pbuf + static_cast<long>((i + 1) & 1) * (m + 1) * nwg;
    for (int j = threadIdx.x; j <= m; j += kMgsThreads)
        publish_raw(reinterpret_cast<uint64_t*>(pnext + static_cast<long>(j) * nwg + wg), kPending);
    const Watchdog wd(wd_limit);
    bool timed_out = false;
    const long r0 = static_cast<long>(wg) * chunk;
    const long r1 = r0 + chunk < n ? r0 + chunk : n;
    // vn holds v_j and vq v_{j+1} at term j; v_{j+2} is requested into vn
    // before the wait for h_{j+1}, so the load overlaps the reduction.
    double wr[R], vn[R], vq[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const long r = r0 + threadIdx.x + static_cast<long>(k) * kMgsThreads;
        wr[k] = r < r1 ? w[r] : 0.0;
        vn[k] = r < r1 ? vv[r] : 0.0;
        vq[k] = r < r1 && i > 0 ? vv[static_cast<long>(n) + r] : 0.0;
    }
    double h = 0.0;
    {
        double v = 0.0;
        for (int k = threadIdx.x; k < kVecBlocks; k += kMgsThreads) v += part0[k];
        h = block_sum512(v);
    }
    for (int j = 0; j <= i; ++j) {
        if (wg == 0 && threadIdx.x == 0) col[j] = h;
        const bool more = j < i;
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            wr[k] = fma(-h, vn[k], wr[k]);
            acc = more ? fma(vq[k], wr[k], acc) : fma(wr[k], wr[k], acc);
        }
        const double tot = block_sum512(acc);
        double* slot = pb + static_cast<long>(j) * nwg;
        if (threadIdx.x == 0) publish(slot + wg, tot);
        if (j + 1 < i) {
            const double* v2 = vv + static_cast<long>(j + 2) * n;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const long r = r0 + threadIdx.x + static_cast<long>(k) * kMgsThreads;
                vn[k] = r < r1 ? v2[r] : 0.0;
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const double t = vn[k];
            vn[k] = vq[k];
            vq[k] = t;
        }
        double v = 0.0;
        if (threadIdx.x < nwg) {
            uint64_t bits;
            while ((bits = peek_bits(slot + threadIdx.x)) == kPending) {
                if (timed_out || wd.expired()) { timed_out = true; bits = kQuietNan; break; }
                spin_pause();
            }
            v = __longlong_as_double(bits);
        }
        h = block_sum512(v);
    }
    const double t = std::sqrt(h);
    if (wg == 0 && threadIdx.x == 0) *tnorm = t;
    const double a = 1.0 / t;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const long r = r0 + threadIdx.x + static_cast<long>(k) * kMgsThreads;
        if (r < r1) w[r] = t <= 1e-16 ? wr[k] : wr[k] * a;
    }
    if (timed_out && threadIdx.x == 0) atomicOr(err, kErrTimeout);
}

__global__ void k_fill_pending(long n, uint64_t* p) {
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    for (long k = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; k < n; k += stride) p[k] = kPending;
}

// v *= ctl->inv when the current cycle/step calls for it.
__global__ __launch_bounds__(kVecThreads) void k_scale(int n, double* v, const GmresCtl* ctl, bool step_gate) {
    if (step_gate ? (ctl->step_active == 0 || ctl->breakdown) : ctl->cycle_active == 0) return;
    const double a = ctl->inv;
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    for (long r = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += stride) v[r] *= a;
}

// Start of a cycle (src/gmres.cpp:185-198).
__global__ void k_cycle_begin(GmresCtl* ctl, const double* part, double* rs) {
    const double beta = std::sqrt(sum_partials(part));
    if (threadIdx.x != 0) return;
    ctl->beta = beta;
    ctl->i = -1;
    ctl->breakdown = 0;
    ctl->step_active = 0;
    if (beta <= ctl->tol) {
        ctl->status = 1;
        ctl->cycle_active = 0;
        return;
    }
    ctl->inv = 1.0 / beta;
    rs[0] = beta;
    ctl->cycle_active = 1;
}

// Loop condition of the Krylov loop (src/gmres.cpp:199) for step i.
__global__ void k_step_gate(GmresCtl* ctl, int i) {
    if (ctl->cycle_active && ctl->beta > ctl->tol) {
        const int before = ctl->its++;
        if (before < ctl->maxits) {
            ctl->step_active = 1;
            ctl->i = i;
            return;
        }
    }
    ctl->step_active = 0;
    ctl->cycle_active = 0;
}

// End of an Arnoldi step (src/gmres.cpp:226-262): h_{i+1,i} = ||w||,
// breakdown test, earlier rotations on column i, new rotation, estimate.
__global__ void k_arnoldi_finish(GmresCtl* ctl, const double* part, const double* tnorm, double* hh, double* hist) {
    if (ctl->step_active == 0) return;
    const double t = tnorm ? *tnorm : std::sqrt(sum_partials(part));
    if (threadIdx.x != 0) return;
    const int im = ctl->im, im1 = im + 1;
    const int i = ctl->i, i1 = i + 1;
    double* col = hh + static_cast<long>(i) * im1;
    double* c = hh + static_cast<long>(im1) * im;
    double* s = c + im1;
    double* rs = s + im1;
    col[i1] = t;
    if (t <= 1e-16) {
        ctl->breakdown = 1;
        ctl->cycle_active = 0;
        return;
    }
    ctl->inv = 1.0 / t;
    for (int k = 1; k <= i; ++k) {
        const int k1 = k - 1;
        const double tk = col[k1];
        col[k1] = c[k1] * tk + s[k1] * col[k];
        col[k] = -s[k1] * tk + c[k1] * col[k];
    }
    const double a = col[i], b = col[i1];
    double gam = std::sqrt(a * a + b * b);
    if (gam == 0.0) gam = 1e-16;
    c[i] = col[i] / gam;
    s[i] = col[i1] / gam;
    rs[i1] = -s[i] * rs[i];
    rs[i] = c[i] * rs[i];
    col[i] = c[i] * col[i] + s[i] * col[i1];
    const double beta = std::fabs(rs[i1]);
    ctl->beta = beta;
    if (hist && ctl->nhist < ctl->hist_cap) hist[ctl->nhist++] = beta;
    if (std::isnan(beta)) {
        ctl->status = 2;
        ctl->cycle_active = 0;
    }
}

// Upper-triangular solve for the step coefficients (src/gmres.cpp:265-272).
__global__ void k_backsolve(GmresCtl* ctl, double* hh) {
    if (ctl->status == 2 || ctl->i < 0) return;
    const int im = ctl->im, im1 = im + 1;
    const int i = ctl->i;
    double* rs = hh + static_cast<long>(im1) * im + 2 * im1;
    rs[i] = rs[i] / hh[static_cast<long>(i) * im1 + i];
    for (int ii = i - 1; ii >= 0; --ii) {
        double t = rs[ii];
        for (int j = ii + 1; j <= i; ++j) t -= hh[static_cast<long>(j) * im1 + ii] * rs[j];
        rs[ii] = t / hh[static_cast<long>(ii) * im1 + ii];
    }
}

// sol += sum_j rs_j z_j, j = 0..i, applied in ascending j per entry
// (src/gmres.cpp:274-276).
__global__ __launch_bounds__(kVecThreads) void k_update(int n, const double* __restrict__ z, const GmresCtl* ctl,
                                                        const double* hh, double* sol) {
    if (ctl->status == 2 || ctl->i < 0) return;
    const int im = ctl->im, im1 = im + 1;
    const int i = ctl->i;
    const double* rs = hh + static_cast<long>(im1) * im + 2 * im1;
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    for (long r = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += stride) {
        double x = sol[r];
        for (int j = 0; j <= i; ++j) x = fma(rs[j], z[static_cast<long>(j) * n + r], x);
        sol[r] = x;
    }
}

// A must be a valid CSR before the SpMV kernels read through it:
// row_ptr[0] == 0, non-decreasing, row_ptr[n] == nnz, columns in [0, n).
__global__ void k_check_csr(int n, long nnz, const int* __restrict__ rp, const int* __restrict__ ci, int* bad) {
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i <= n; i += stride) {
        const bool ok = (i == 0) ? rp[0] == 0 : (rp[i] >= rp[i - 1] && (i < n || rp[n] == nnz));
        if (!ok) atomicOr(bad, 1);
    }
}

__global__ void k_check_cols(int n, long nnz, const int* __restrict__ ci, int* bad) {
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    for (long k = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; k < nnz; k += stride) {
        const int c = ci[k];
        if (c < 0 || c >= n) atomicOr(bad, 2);
    }
}

__global__ void k_sqrt_sum(const double* part, double* out) {
    const double v = std::sqrt(sum_partials(part));
    if (threadIdx.x == 0) *out = v;
}

}  // namespace pilu

struct lljbash::PreconditionedGmres::Ext {
    int n = 0, m = 0;
    pilu::DevBuf<double> vv, z, hh, part, hist, scratch, pbuf;
    pilu::DevBuf<int> flags;
    pilu::DevBuf<pilu::GmresCtl> ctl;
    pilu::DevBuf<int> arp, aci;
    pilu::DevBuf<double> aval, rhs, sol;
    pilu::GmresCtl* snap = nullptr;  // pinned host snapshots (2 slots)
    hipEvent_t snap_ev[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> tev;     // timing events (pairs)
    int hist_len = 0;
    double true_res = -1.0;
    std::vector<double> hv, hz;      // host vectors of a host-side preconditioner
    ~Ext() {
        if (snap) (void)hipHostFree(snap);
        for (auto& e : snap_ev) if (e) (void)hipEventDestroy(e);
        for (auto& e : tev) (void)hipEventDestroy(e);
    }
};

namespace lljbash {

namespace {

using GExt = PreconditionedGmres::Ext;

template <int G>
void spmv_g(hipStream_t s, int n, const int* rp, const int* ci, const double* val, double alpha, const double* x,
            double beta, const double* yin, double* y, const double* dotv, bool nrm, double* part, const int* gate) {
    pilu::k_spmv<G><<<pilu::kVecBlocks, pilu::kVecThreads, 0, s>>>(n, rp, ci, val, alpha, x, beta, yin, y, dotv,
                                                                  nrm, part, gate);
}

void spmv(int lanes, hipStream_t s, int n, const int* rp, const int* ci, const double* val, double alpha,
          const double* x, double beta, const double* yin, double* y, const double* dotv, bool nrm, double* part,
          const int* gate) {
    switch (lanes) {
        case 4: spmv_g<4>(s, n, rp, ci, val, alpha, x, beta, yin, y, dotv, nrm, part, gate); break;
        case 8: spmv_g<8>(s, n, rp, ci, val, alpha, x, beta, yin, y, dotv, nrm, part, gate); break;
        case 16: spmv_g<16>(s, n, rp, ci, val, alpha, x, beta, yin, y, dotv, nrm, part, gate); break;
        case 32: spmv_g<32>(s, n, rp, ci, val, alpha, x, beta, yin, y, dotv, nrm, part, gate); break;
        default: spmv_g<64>(s, n, rp, ci, val, alpha, x, beta, yin, y, dotv, nrm, part, gate); break;
    }
}

inline int grid_for_n(long n, int per) { return static_cast<int>(std::max<long>(1, std::min<long>((n + per - 1) / per, 65535))); }

// Launch shape of k_mgs_all: one co-resident workgroup per compute unit,
// rows per thread R a power of two up to 16 (32 would spill registers);
// nwg == 0 when n is too large (the per-term kernels are used then).
struct MgsPlan {
    int nwg = 0, chunk = 0, r = 0;
};

template <int R>
bool mgs_fits(int nwg) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(pilu::k_mgs_all<R>),
                                                     pilu::kMgsThreads, 0) != hipSuccess)
        return false;
    return per_cu >= 1 && nwg <= pilu::kMgsThreads;
}

MgsPlan plan_mgs(int n) {
    MgsPlan p;
    static int cus = -1;
    if (cus < 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 0;
    }
    if (cus <= 0 || n <= 0) return p;
    const int chunk = (n + cus - 1) / cus;
    int r = 1;
    while (r < 16 && r * pilu::kMgsThreads < chunk) r <<= 1;
    if (r * pilu::kMgsThreads < chunk) return p;
    bool ok = false;
    switch (r) {
        case 1: ok = mgs_fits<1>(cus); break;
        case 2: ok = mgs_fits<2>(cus); break;
        case 4: ok = mgs_fits<4>(cus); break;
        case 8: ok = mgs_fits<8>(cus); break;
        default: ok = mgs_fits<16>(cus); break;
    }
    if (!ok) return p;
    p.nwg = cus;
    p.chunk = chunk;
    p.r = r;
    return p;
}

template <int R>
hipError_t launch_mgs_r(const MgsPlan& p, hipStream_t s, int n, int i, int m, const double* vv, double* w,
                        const double* part0, double* pbuf, double* col, double* tnorm, const pilu::GmresCtl* ctl,
                        int* err, uint64_t wd) {
    int chunk = p.chunk;
    void* args[] = {&n, &i, &m, &chunk, &vv, &w, &part0, &pbuf, &col, &tnorm, &ctl, &err, &wd};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(pilu::k_mgs_all<R>), dim3(p.nwg),
                                      dim3(pilu::kMgsThreads), args, 0, s);
}

// Set once the runtime refuses a cooperative launch in this process (it
// does, for instance, while another process on the GPU holds the
// cooperative queue); solves then use the per-term kernels.
std::atomic<bool> g_coop_refused{false};

// The one-launch Gram-Schmidt; false (error cleared) when the runtime
// refuses the cooperative launch, so the caller can use the per-term kernels.
bool launch_mgs_all(const MgsPlan& p, hipStream_t s, int n, int i, int m, const double* vv, double* w,
                    const double* part0, double* pbuf, double* col, double* tnorm, const pilu::GmresCtl* ctl,
                    int* err, uint64_t wd) {
    hipError_t rc;
    switch (p.r) {
        case 1: rc = launch_mgs_r<1>(p, s, n, i, m, vv, w, part0, pbuf, col, tnorm, ctl, err, wd); break;
        case 2: rc = launch_mgs_r<2>(p, s, n, i, m, vv, w, part0, pbuf, col, tnorm, ctl, err, wd); break;
        case 4: rc = launch_mgs_r<4>(p, s, n, i, m, vv, w, part0, pbuf, col, tnorm, ctl, err, wd); break;
        case 8: rc = launch_mgs_r<8>(p, s, n, i, m, vv, w, part0, pbuf, col, tnorm, ctl, err, wd); break;
        default: rc = launch_mgs_r<16>(p, s, n, i, m, vv, w, part0, pbuf, col, tnorm, ctl, err, wd); break;
    }
    if (rc == hipSuccess) return true;
    (void)hipGetLastError();
    g_coop_refused = true;
    return false;
}

struct TimedPair {
    hipEvent_t a, b;
    int kind;  // 0 = mv, 1 = precon
};

// precon (used when ilu is null): M^-1 on host vectors, called once per
// Arnoldi step with the step's basis vector copied to the host.
std::pair<bool, int> solve_core(GExt* e, IluSolver* ilu, const std::function<void(const double*, double*)>* precon,
                                const GmresParameters& prm, GmresStat& st, int n, long nnz, const int* rp,
                                const int* ci, const double* val, const double* rhs, double* sol) {
    const bool host_precon = !ilu && precon && *precon;
    if (!host_precon && (!ilu || !ilu->DeviceState())) return {false, 0};
    IluSolver::Ext* ie = host_precon ? nullptr : ilu->DeviceState();
    if (ie && ie->n != n) throw std::runtime_error("GMRES: matrix size differs from the preconditioner's");
    hipStream_t s = ie ? pilu::ilu_stream(ie) : nullptr;
    if (host_precon) {
        e->hv.resize(static_cast<size_t>(n));
        e->hz.resize(static_cast<size_t>(n));
    }
    const int m = prm.krylov_subspace_dimension;
    const int im1 = m + 1;
    if (m < 1) throw std::runtime_error("GMRES: krylov_subspace_dimension must be >= 1");
    e->vv.reserve(static_cast<size_t>(im1) * n);
    e->z.reserve(static_cast<size_t>(m) * n);
    e->hh.reserve(static_cast<size_t>(im1) * (m + 3));
    e->part.reserve(2 * pilu::kVecBlocks + 8);
    e->hist.reserve(static_cast<size_t>(prm.max_iterations) + 2);
    e->scratch.reserve(n);
    e->ctl.reserve(1);
    e->n = n;
    e->m = m;
    if (!e->snap) {
        PILU_HIP(hipHostMalloc(reinterpret_cast<void**>(&e->snap), 2 * sizeof(pilu::GmresCtl)));
        for (auto& ev : e->snap_ev) PILU_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    const double avg = n > 0 ? static_cast<double>(nnz) / n : 1.0;
    const int lanes = pilu::lanes_for_avg(avg);
    // One-launch Gram-Schmidt when each of the co-resident workgroups can
    // keep its rows of w in registers.
    const MgsPlan mgs = g_coop_refused ? MgsPlan{} : plan_mgs(n);
    double* pbuf = nullptr;
    if (mgs.nwg > 0) {
        e->pbuf.reserve(static_cast<size_t>(2) * (m + 1) * mgs.nwg);
        pbuf = e->pbuf.get();
    }

    // Validate A before any kernel indexes through it.
    e->flags.reserve(4);
    PILU_HIP(hipMemsetAsync(e->flags.get(), 0, sizeof(int), s));
    pilu::k_check_csr<<<256, 256, 0, s>>>(n, nnz, rp, ci, e->flags.get());
    int bad = 0;
    PILU_HIP(hipMemcpyAsync(&bad, e->flags.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    PILU_HIP(hipStreamSynchronize(s));
    if (bad) throw std::runtime_error("GMRES: row_ptr of A is not a valid CSR row pointer");
    pilu::k_check_cols<<<1024, 256, 0, s>>>(n, nnz, ci, e->flags.get());
    PILU_HIP(hipMemcpyAsync(&bad, e->flags.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    PILU_HIP(hipStreamSynchronize(s));
    if (bad) throw std::runtime_error("GMRES: column index of A out of range");

    pilu::GmresCtl h{};
    h.its = 0;
    h.maxits = prm.max_iterations;
    h.im = m;
    h.hist_cap = prm.max_iterations + 2;
    h.tol = prm.tolerance;
    PILU_HIP(hipMemcpyAsync(e->ctl.get(), &h, sizeof(h), hipMemcpyHostToDevice, s));
    PILU_HIP(hipMemsetAsync(e->hh.get(), 0, sizeof(double) * im1 * (m + 3), s));

    pilu::GmresCtl* dctl = e->ctl.get();
    const int* step_gate = &dctl->step_active;
    double* vv = e->vv.get();
    double* z = e->z.get();
    double* hh = e->hh.get();
    double* part0 = e->part.get();
    double* part1 = part0 + pilu::kVecBlocks;
    double* tnorm = part0 + 2 * pilu::kVecBlocks;
    int* mgs_err = e->flags.get() + 1;
    PILU_HIP(hipMemsetAsync(mgs_err, 0, sizeof(int), s));
    const uint64_t wd = pilu::watchdog_ticks(10.0);
    std::vector<TimedPair> timed;
    size_t ev_used = 0;
    auto next_event = [&]() {
        if (ev_used == e->tev.size()) {
            hipEvent_t ev;
            PILU_HIP(hipEventCreate(&ev));
            e->tev.push_back(ev);
        }
        return e->tev[ev_used++];
    };
    auto t_begin = [&](int kind) {
        TimedPair p{next_event(), next_event(), kind};
        PILU_HIP(hipEventRecord(p.a, s));
        timed.push_back(p);
    };
    auto t_end = [&]() { PILU_HIP(hipEventRecord(timed.back().b, s)); };

    const auto wall0 = std::chrono::steady_clock::now();
    bool ok = true;
    int its = 0;
    while (its < prm.max_iterations) {
        // r = rhs - A sol, ||r|| (src/gmres.cpp:181-185)
        t_begin(0);
        spmv(lanes, s, n, rp, ci, val, -1.0, sol, 1.0, rhs, vv, nullptr, true, part0, nullptr);
        t_end();
        pilu::k_cycle_begin<<<1, pilu::kVecThreads, 0, s>>>(dctl, part0, hh + static_cast<long>(im1) * m + 2 * im1);
        PILU_HIP(hipMemcpyAsync(&h, dctl, sizeof(h), hipMemcpyDeviceToHost, s));
        PILU_HIP(hipStreamSynchronize(s));
        if (h.status == 1) break;  // beta <= tol (src/gmres.cpp:186-188)
        // A NaN residual would make the reference's outer loop spin forever
        // (its inner loop never runs); report failure instead.
        if (std::isnan(h.beta)) { ok = false; break; }
        pilu::k_scale<<<pilu::kVecBlocks, pilu::kVecThreads, 0, s>>>(n, vv, dctl, false);
        if (pbuf) {
            const long cnt = static_cast<long>(2) * (m + 1) * mgs.nwg;
            pilu::k_fill_pending<<<grid_for_n(cnt, 256), 256, 0, s>>>(cnt, reinterpret_cast<uint64_t*>(pbuf));
        }
        for (int i = 0; i < m; ++i) {
            double* vi = vv + static_cast<long>(i) * n;
            double* vi1 = vv + static_cast<long>(i + 1) * n;
            double* zi = z + static_cast<long>(i) * n;
            pilu::k_step_gate<<<1, 1, 0, s>>>(dctl, i);
            t_begin(1);
            if (host_precon) {
                // the step runs only while the cycle is active (the gate the
                // device kernels read)
                pilu::GmresCtl hc{};
                PILU_HIP(hipMemcpyAsync(&hc, dctl, sizeof(hc), hipMemcpyDeviceToHost, s));
                PILU_HIP(hipStreamSynchronize(s));
                if (hc.step_active) {
                    PILU_HIP(hipMemcpy(e->hv.data(), vi, sizeof(double) * n, hipMemcpyDeviceToHost));
                    (*precon)(e->hv.data(), e->hz.data());
                    PILU_HIP(hipMemcpy(zi, e->hz.data(), sizeof(double) * n, hipMemcpyHostToDevice));
                }
            } else {
                pilu::ilu_apply(ie, vi, zi, step_gate);
            }
            t_end();
            t_begin(0);
            spmv(lanes, s, n, rp, ci, val, 1.0, zi, 0.0, nullptr, vi1, vv, false, part0, step_gate);
            t_end();
            // a refused cooperative launch switches this and later steps to
            // the per-term kernels
            if (pbuf && !launch_mgs_all(mgs, s, n, i, m, vv, vi1, part0, pbuf, hh + static_cast<long>(i) * im1,
                                        tnorm, dctl, mgs_err, wd))
                pbuf = nullptr;
            if (pbuf) {
                pilu::k_arnoldi_finish<<<1, pilu::kVecThreads, 0, s>>>(dctl, nullptr, tnorm, hh, e->hist.get());
            } else {
                double* pin = part0;
                double* pout = part1;
                for (int j = 0; j <= i; ++j) {
                    const double* vnext = j < i ? vv + static_cast<long>(j + 1) * n : nullptr;
                    pilu::k_mgs<<<pilu::kVecBlocks, pilu::kVecThreads, 0, s>>>(
                        n, vv + static_cast<long>(j) * n, vi1, vnext, pin, pout, hh + static_cast<long>(i) * im1 + j,
                        step_gate);
                    std::swap(pin, pout);
                }
                pilu::k_arnoldi_finish<<<1, pilu::kVecThreads, 0, s>>>(dctl, pin, nullptr, hh, e->hist.get());
                pilu::k_scale<<<pilu::kVecBlocks, pilu::kVecThreads, 0, s>>>(n, vi1, dctl, true);
            }
            PILU_HIP(hipMemcpyAsync(&e->snap[i & 1], dctl, sizeof(h), hipMemcpyDeviceToHost, s));
            PILU_HIP(hipEventRecord(e->snap_ev[i & 1], s));
            if (i >= 1) {
                PILU_HIP(hipEventSynchronize(e->snap_ev[(i - 1) & 1]));
                if (!e->snap[(i - 1) & 1].cycle_active) break;
            }
        }
        pilu::k_backsolve<<<1, 1, 0, s>>>(dctl, hh);
        pilu::k_update<<<pilu::kVecBlocks, pilu::kVecThreads, 0, s>>>(n, z, dctl, hh, sol);
        PILU_HIP(hipGetLastError());
        PILU_HIP(hipMemcpyAsync(&h, dctl, sizeof(h), hipMemcpyDeviceToHost, s));
        PILU_HIP(hipStreamSynchronize(s));
        its = h.its;
        if (h.status == 2) { ok = false; break; }          // src/gmres.cpp:259-262
        if (h.beta <= h.tol) break;                          // :278-280
        if (its > prm.max_iterations) { ok = false; break; } // :281-283
    }
    its = h.its;
    // FP64 true residual of the returned iterate (reported, not used).
    spmv(lanes, s, n, rp, ci, val, -1.0, sol, 1.0, rhs, e->scratch.get(), nullptr, true, part0, nullptr);
    pilu::k_sqrt_sum<<<1, pilu::kVecThreads, 0, s>>>(part0, part1);
    PILU_HIP(hipMemcpyAsync(&e->true_res, part1, sizeof(double), hipMemcpyDeviceToHost, s));
    if (ie) pilu::check_sweep_errors(ie);
    if (mgs.nwg > 0) {  // cooperative launches may have run (before a refusal too)
        int werr = 0;
        PILU_HIP(hipMemcpyAsync(&werr, mgs_err, sizeof(int), hipMemcpyDeviceToHost, s));
        PILU_HIP(hipStreamSynchronize(s));
        if (werr & pilu::kErrTimeout) throw std::runtime_error("GMRES: Gram-Schmidt reduction wait timed out (watchdog)");
    }
    st.total_fgmr_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    for (const auto& p : timed) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            if (p.kind == 0) st.total_mv_time += ms * 1e-3;
            else st.total_precon_time += ms * 1e-3;
        }
    }
    e->hist_len = h.nhist;
    return {ok, its};
}

}  // namespace

PreconditionedGmres::~PreconditionedGmres() { delete intermediate_.ext; }

std::pair<bool, int> PreconditionedGmres::Solve(const CsrMatrix* mat, double* rhs, double* sol) {
    if (!intermediate_.ext) intermediate_.ext = new Ext;
    Ext* e = intermediate_.ext;
    const int n = mat->size;
    const long nnz = mat->row_ptr[n];
    const auto t0 = std::chrono::steady_clock::now();
    e->arp.reserve(static_cast<size_t>(n) + 1);
    e->aci.reserve(std::max<long>(nnz, 1));
    e->aval.reserve(std::max<long>(nnz, 1));
    e->rhs.reserve(n);
    e->sol.reserve(n);
    PILU_HIP(hipMemcpy(e->arp.get(), mat->row_ptr, sizeof(int) * (static_cast<size_t>(n) + 1), hipMemcpyHostToDevice));
    PILU_HIP(hipMemcpy(e->aci.get(), mat->col_idx, sizeof(int) * nnz, hipMemcpyHostToDevice));
    PILU_HIP(hipMemcpy(e->aval.get(), mat->value, sizeof(double) * nnz, hipMemcpyHostToDevice));
    PILU_HIP(hipMemcpy(e->rhs.get(), rhs, sizeof(double) * n, hipMemcpyHostToDevice));
    PILU_HIP(hipMemcpy(e->sol.get(), sol, sizeof(double) * n, hipMemcpyHostToDevice));
    stat_.total_init_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    auto r = solve_core(e, intermediate_.ilu, &precon_, param_, stat_, n, nnz, e->arp.get(), e->aci.get(),
                        e->aval.get(), e->rhs.get(), e->sol.get());
    PILU_HIP(hipMemcpy(sol, e->sol.get(), sizeof(double) * n, hipMemcpyDeviceToHost));
    return r;
}

std::pair<bool, int> PreconditionedGmres::SolveDevice(const CsrMatrix* d_mat, const double* d_rhs, double* d_sol) {
    if (!intermediate_.ext) intermediate_.ext = new Ext;
    const int n = d_mat->size;
    int h_last = 0;
    PILU_HIP(hipMemcpy(&h_last, d_mat->row_ptr + n, sizeof(int), hipMemcpyDeviceToHost));
    const long nnz = h_last;
    return solve_core(intermediate_.ext, intermediate_.ilu, &precon_, param_, stat_, n, nnz, d_mat->row_ptr,
                      d_mat->col_idx, d_mat->value, d_rhs, d_sol);
}

double PreconditionedGmres::LastTrueResidual() const {
    return intermediate_.ext ? intermediate_.ext->true_res : -1.0;
}

int PreconditionedGmres::ResidualHistory(double* out, int cap) const {
    const Ext* e = intermediate_.ext;
    if (!e || e->hist_len <= 0) return 0;
    const int k = std::min(cap, e->hist_len);
    if (k > 0) PILU_HIP(hipMemcpy(out, e->hist.get(), sizeof(double) * k, hipMemcpyDeviceToHost));
    return k;
}

}  // namespace lljbash
