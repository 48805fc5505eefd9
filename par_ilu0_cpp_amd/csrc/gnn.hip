// Learned diagonal compensation: a 3-layer message-passing network over the
// matrix graph that predicts, for every row, the relaxation factor omega_i
// of the compensated factorization (IluSolver::SetDiagonalCompensation).
// No counterpart in the reference.
//
// Graph: rows are nodes, off-diagonal entries a_ij are edges with weight
// w_ij = a_ij / sqrt(|a_ii a_jj|).  Node features (16, eight used): scale-
// free row statistics (diagonal dominance, signed row sum, extreme
// couplings, degree, coupling energy, lower-triangle share).
// Layer l (input width D = 16, 32, 32; hidden 32):
//   M = W_graph H;  Z = [H | M] W_l + b_l;  H' = relu(Z)
// Head: omega = sigmoid(H_3 w_o + b_o).
// The dense products run on BF16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32
// accumulation; gathers, head, loss and Adam are FP32.  Training minimises
// the mean squared error between omega and per-row targets.  All reductions
// over rows are two-level with a fixed order, so gradients are
// deterministic.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "common.hpp"
#include "gmres.hpp"
#include "par_ilu_ex.h"

namespace pilu {
namespace gnn {

constexpr int kHid = 32;
constexpr int kFeat = 16;
constexpr int kLayers = 3;
constexpr int kInDim[kLayers] = {kFeat, kHid, kHid};
// parameter layout: W0 (32x32), b0, W1 (64x32), b1, W2 (64x32), b2, wo (32), bo
constexpr int kOffW[kLayers] = {0, 1024 + 32, 1024 + 32 + 2048 + 32};
constexpr int kOffB[kLayers] = {1024, 1024 + 32 + 2048, 1024 + 32 + 2048 + 32 + 2048};
constexpr int kOffWo = 1024 + 32 + 2 * (2048 + 32);
constexpr int kOffBo = kOffWo + kHid;
constexpr int kParams = kOffBo + 1;  // 5249
constexpr int kRowsPerWaveW = 512;   // rows per wave in the weight-gradient pass

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    return static_cast<unsigned short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);  // round to nearest even
}

__device__ __forceinline__ short bits_bf(float f) { return static_cast<short>(f2bf(f)); }

// ---- features ----
__global__ void k_diag_values(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                              const double* __restrict__ va, double* __restrict__ d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e)
        if (ci[e] == i) v = va[e];
    d[i] = v;
}

__global__ void k_features(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ va, const double* __restrict__ d, float* __restrict__ feat,
                           float* __restrict__ ew) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double di = d[i];
    const double ai = fabs(di) > 0 ? fabs(di) : 1.0;
    double sabs = 0, ssum = 0, mx = 0, mn = 1e300, sw2 = 0, lower = 0;
    int cnt = 0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
        const int j = ci[e];
        if (j == i) { ew[e] = 0.f; continue; }
        const double a = va[e];
        const double dj = fabs(d[j]) > 0 ? fabs(d[j]) : 1.0;
        const double w = a / sqrt(ai * dj);
        ew[e] = static_cast<float>(w);
        sabs += fabs(a);
        ssum += a;
        mx = fabs(a) > mx ? fabs(a) : mx;
        mn = fabs(a) < mn ? fabs(a) : mn;
        sw2 += w * w;
        if (j < i) lower += fabs(a);
        ++cnt;
    }
    if (cnt == 0) mn = 0;
    float* f = feat + static_cast<long>(i) * kFeat;
    f[0] = 1.f;
    f[1] = static_cast<float>(sabs / ai);
    f[2] = static_cast<float>(ssum / (di != 0 ? di : 1.0));
    f[3] = static_cast<float>(mx / ai);
    f[4] = static_cast<float>(mn / ai);
    f[5] = cnt / 8.f;
    f[6] = static_cast<float>(sw2);
    f[7] = static_cast<float>(sabs > 0 ? lower / sabs : 0.0);
    for (int k = 8; k < kFeat; ++k) f[k] = 0.f;
}

// out[i] (+)= sum_e w_e H[col_e]   (D lanes per row, entries in order)
template <int D>
__global__ __launch_bounds__(256) void k_aggregate(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const float* __restrict__ w, const float* __restrict__ h,
                                                   float* __restrict__ out, bool accumulate) {
    constexpr int R = 64 / D;
    const int lane = threadIdx.x & 63;
    const int d = lane % D;
    const long step = static_cast<long>(gridDim.x) * 4 * R;
    // rows strided over a capped grid; a row's edges are read four at a time
    // and summed in edge order
    for (long row = (static_cast<long>(blockIdx.x) * 4 + (threadIdx.x >> 6)) * R + lane / D; row < n; row += step) {
        float acc = 0.f;
        int e = rp[row];
        const int e1 = rp[row + 1];
        for (; e + 4 <= e1; e += 4) {
            float we[4], he[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                we[q] = w[e + q];
                he[q] = h[static_cast<long>(ci[e + q]) * D + d];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = fmaf(we[q], he[q], acc);
        }
        for (; e < e1; ++e) acc = fmaf(w[e], h[static_cast<long>(ci[e]) * D + d], acc);
        float* o = out + row * D + d;
        *o = accumulate ? *o + acc : acc;
    }
}

constexpr int kAggBlocks = 4096;  // grid cap of k_aggregate (rows beyond it are strided)

// Z = [H | M] W + b, Hout = relu(Z); one wave per 16 rows, two 16x16 output tiles.
template <int D>
__global__ __launch_bounds__(256) void k_dense_fwd(int n, const float* __restrict__ h, const float* __restrict__ m,
                                                   const unsigned short* __restrict__ wb, const float* __restrict__ b,
                                                   float* __restrict__ z, float* __restrict__ hout) {
    constexpr int K = 2 * D;
    const int lane = threadIdx.x & 63;
    const long r0 = (static_cast<long>(blockIdx.x) * 4 + (threadIdx.x >> 6)) * 16;
    if (r0 >= n) return;
    const int ar = lane & 15, kq = lane >> 4;
    const long arow = r0 + ar;
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < K / 16; ++ks) {
        const int k0 = ks * 16 + kq * 4;
        bf16x4 a, b0, b1;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int k = k0 + t;
            float x = 0.f;
            if (arow < n) x = k < D ? h[arow * D + k] : m[arow * D + (k - D)];
            a[t] = bits_bf(x);
            b0[t] = static_cast<short>(wb[k * kHid + ar]);
            b1[t] = static_cast<short>(wb[k * kHid + 16 + ar]);
        }
        acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1, acc1, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long row = r0 + kq * 4 + j;
        if (row >= n) continue;
        const float z0 = acc0[j] + b[ar], z1 = acc1[j] + b[16 + ar];
        z[row * kHid + ar] = z0;
        z[row * kHid + 16 + ar] = z1;
        hout[row * kHid + ar] = z0 > 0.f ? z0 : 0.f;
        hout[row * kHid + 16 + ar] = z1 > 0.f ? z1 : 0.f;
    }
}

// dX = (dHout * [Z > 0]) W^T, split into dH (first D columns) and dM.
template <int D>
__global__ __launch_bounds__(256) void k_dense_bwd_x(int n, const float* __restrict__ dhout, const float* __restrict__ z,
                                                     const unsigned short* __restrict__ wb, float* __restrict__ dh,
                                                     float* __restrict__ dm) {
    constexpr int K = 2 * D;
    const int lane = threadIdx.x & 63;
    const long r0 = (static_cast<long>(blockIdx.x) * 4 + (threadIdx.x >> 6)) * 16;
    if (r0 >= n) return;
    const int ar = lane & 15, kq = lane >> 4;
    const long arow = r0 + ar;
    f32x4 acc[K / 16];
#pragma unroll
    for (int t = 0; t < K / 16; ++t) acc[t] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < kHid / 16; ++ks) {
        const int k0 = ks * 16 + kq * 4;
        bf16x4 a;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            float g = 0.f;
            if (arow < n) {
                const long o = arow * kHid + k0 + t;
                g = z[o] > 0.f ? dhout[o] : 0.f;
            }
            a[t] = bits_bf(g);
        }
#pragma unroll
        for (int ct = 0; ct < K / 16; ++ct) {
            bf16x4 bb;
#pragma unroll
            for (int t = 0; t < 4; ++t) bb[t] = static_cast<short>(wb[(ct * 16 + ar) * kHid + k0 + t]);
            acc[ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[ct], 0, 0, 0);
        }
    }
#pragma unroll
    for (int ct = 0; ct < K / 16; ++ct)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long row = r0 + kq * 4 + j;
            if (row >= n) continue;
            const int c = ct * 16 + ar;
            if (c < D) dh[row * D + c] = acc[ct][j];
            else dm[row * D + (c - D)] = acc[ct][j];
        }
}

// Per-wave partials of dW = [H | M]^T dZ (K x 32) and db = sum dZ (32).
template <int D>
__global__ __launch_bounds__(256) void k_dense_bwd_w(int n, const float* __restrict__ h, const float* __restrict__ m,
                                                     const float* __restrict__ dhout, const float* __restrict__ z,
                                                     float* __restrict__ part) {
    constexpr int K = 2 * D;
    constexpr int LEN = K * kHid + kHid;
    const int lane = threadIdx.x & 63;
    const long wv = static_cast<long>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const long rb = wv * kRowsPerWaveW;
    const int ar = lane & 15, kq = lane >> 4;
    f32x4 acc[K / 16][2];
#pragma unroll
    for (int ft = 0; ft < K / 16; ++ft) acc[ft][0] = acc[ft][1] = {0.f, 0.f, 0.f, 0.f};
    float dbacc = 0.f;
    for (long r0 = rb; r0 < rb + kRowsPerWaveW && r0 < n; r0 += 16) {
        // B operand rows: data rows r0 + kq*4 + t; columns c = ct*16 + ar
        bf16x4 bz[2];
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const long row = r0 + kq * 4 + t;
                float g = 0.f;
                if (row < n) {
                    const long o = row * kHid + ct * 16 + ar;
                    g = z[o] > 0.f ? dhout[o] : 0.f;
                }
                bz[ct][t] = bits_bf(g);
            }
#pragma unroll
        for (int ft = 0; ft < K / 16; ++ft) {
            bf16x4 a;  // A[f][k] = X[row k][f], f = ft*16 + ar
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const long row = r0 + kq * 4 + t;
                const int f = ft * 16 + ar;
                float x = 0.f;
                if (row < n) x = f < D ? h[row * D + f] : m[row * D + (f - D)];
                a[t] = bits_bf(x);
            }
            acc[ft][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bz[0], acc[ft][0], 0, 0, 0);
            acc[ft][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bz[1], acc[ft][1], 0, 0, 0);
        }
        if (lane < kHid) {
            for (int t = 0; t < 16; ++t) {
                const long row = r0 + t;
                if (row < n) {
                    const long o = row * kHid + lane;
                    dbacc += z[o] > 0.f ? dhout[o] : 0.f;
                }
            }
        }
    }
    float* out = part + wv * LEN;
#pragma unroll
    for (int ft = 0; ft < K / 16; ++ft)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int j = 0; j < 4; ++j) out[(ft * 16 + kq * 4 + j) * kHid + ct * 16 + ar] = acc[ft][ct][j];
    if (lane < kHid) out[K * kHid + lane] = dbacc;
}

// out[j] = sum over parts in order (one thread per j)
__global__ void k_sum_parts(int nparts, int len, const float* __restrict__ part, float* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    float s = 0.f;
    for (int p = 0; p < nparts; ++p) s += part[static_cast<long>(p) * len + j];
    out[j] = s;
}

// omega = sigmoid(H3 wo + bo)
__global__ void k_head(int n, const float* __restrict__ h3, const float* __restrict__ prm, float* __restrict__ omega) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float u = prm[kOffBo];
    for (int d = 0; d < kHid; ++d) u = fmaf(h3[static_cast<long>(i) * kHid + d], prm[kOffWo + d], u);
    omega[i] = 1.f / (1.f + expf(-u));
}

// MSE loss head backward; per-block partials of (dwo[32], dbo, loss).
__global__ __launch_bounds__(256) void k_head_bwd(int n, const float* __restrict__ h3, const float* __restrict__ omega,
                                                  const float* __restrict__ target, const float* __restrict__ prm,
                                                  float* __restrict__ dh3, float* __restrict__ part) {
    __shared__ float s[256][kHid + 2];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float du = 0.f, l = 0.f;
    if (i < n) {
        const float w = omega[i];
        const float diff = w - target[i];
        du = 2.f / n * diff * w * (1.f - w);
        l = diff * diff / n;
        for (int d = 0; d < kHid; ++d) dh3[static_cast<long>(i) * kHid + d] = du * prm[kOffWo + d];
    }
    for (int d = 0; d < kHid; ++d) s[threadIdx.x][d] = i < n ? du * h3[static_cast<long>(i) * kHid + d] : 0.f;
    s[threadIdx.x][kHid] = du;
    s[threadIdx.x][kHid + 1] = l;
    __syncthreads();
    if (threadIdx.x < kHid + 2) {
        float t = 0.f;
        for (int r = 0; r < 256; ++r) t += s[r][threadIdx.x];
        part[static_cast<long>(blockIdx.x) * (kHid + 2) + threadIdx.x] = t;
    }
}

__global__ void k_adam(int np, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= np) return;
    const float gj = g[j];
    const float a = b1 * m1[j] + (1.f - b1) * gj;
    const float v = b2 * m2[j] + (1.f - b2) * gj * gj;
    m1[j] = a;
    m2[j] = v;
    p[j] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

__global__ void k_to_bf16(int count, const float* __restrict__ src, unsigned short* __restrict__ dst) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < count) dst[j] = f2bf(src[j]);
}

__global__ void k_gather_f(long count, const int* __restrict__ idx, const float* __restrict__ src,
                           float* __restrict__ dst) {
    const long j = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (j < count) dst[j] = src[idx[j]];
}

}  // namespace gnn

class GnnModel {
public:
    GnnModel() {
        params_.resize(gnn::kParams);
        grads_.resize(gnn::kParams);
        m1_.resize(gnn::kParams);
        m2_.resize(gnn::kParams);
        wb_.resize(gnn::kParams);
        PILU_HIP(hipMemset(m1_.get(), 0, sizeof(float) * gnn::kParams));
        PILU_HIP(hipMemset(m2_.get(), 0, sizeof(float) * gnn::kParams));
        // Glorot-style deterministic initialisation (splitmix sequence).
        std::vector<float> h(gnn::kParams, 0.f);
        uint64_t st = 0x9E3779B97F4A7C15ull;
        auto rnd = [&]() {
            st += 0x9E3779B97F4A7C15ull;
            uint64_t x = st;
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
            x ^= x >> 31;
            return static_cast<double>(x >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
        };
        for (int l = 0; l < gnn::kLayers; ++l) {
            const int k = 2 * gnn::kInDim[l];
            const double a = std::sqrt(6.0 / (k + gnn::kHid));
            for (int j = 0; j < k * gnn::kHid; ++j) h[gnn::kOffW[l] + j] = static_cast<float>(a * rnd());
        }
        const double ao = std::sqrt(6.0 / (gnn::kHid + 1));
        for (int j = 0; j < gnn::kHid; ++j) h[gnn::kOffWo + j] = static_cast<float>(ao * rnd());
        set_params(h.data());
    }

    void set_params(const float* host) {
        PILU_HIP(hipMemcpy(params_.get(), host, sizeof(float) * gnn::kParams, hipMemcpyHostToDevice));
        refresh_bf16();
    }
    void get_params(float* host) const {
        PILU_HIP(hipMemcpy(host, params_.get(), sizeof(float) * gnn::kParams, hipMemcpyDeviceToHost));
    }
    void get_grads(float* host) const {
        PILU_HIP(hipMemcpy(host, grads_.get(), sizeof(float) * gnn::kParams, hipMemcpyDeviceToHost));
    }

    // Forward pass; omega (device, n floats) receives the predictions.
    void forward(int n, const int* rp, const int* ci, const double* va, float* omega) {
        prepare(n, rp, ci, va);
        const int wgs16 = ceil_div(n, 64);  // 16 rows per wave, 4 waves per block
        for (int l = 0; l < gnn::kLayers; ++l) {
            const float* hin = l == 0 ? feat_.get() : h_[l - 1].get();
            if (l == 0) {
                gnn::k_aggregate<gnn::kFeat><<<std::min(ceil_div(n, 16), gnn::kAggBlocks), 256>>>(n, rp, ci, ew_.get(), hin, m_[0].get(), false);
                gnn::k_dense_fwd<gnn::kFeat><<<wgs16, 256>>>(n, hin, m_[0].get(), wb_.get() + gnn::kOffW[0],
                                                             params_.get() + gnn::kOffB[0], z_[0].get(), h_[0].get());
            } else {
                gnn::k_aggregate<gnn::kHid><<<std::min(ceil_div(n, 8), gnn::kAggBlocks), 256>>>(n, rp, ci, ew_.get(), hin, m_[l].get(), false);
                gnn::k_dense_fwd<gnn::kHid><<<wgs16, 256>>>(n, hin, m_[l].get(), wb_.get() + gnn::kOffW[l],
                                                            params_.get() + gnn::kOffB[l], z_[l].get(), h_[l].get());
            }
        }
        gnn::k_head<<<ceil_div(n, 256), 256>>>(n, h_[2].get(), params_.get(), omega);
        PILU_HIP(hipGetLastError());
    }

    // Gradients of the MSE loss (into grads_); returns the loss.
    double backward(int n, const int* rp, const int* ci, const double* va, const int* rpt, const int* cit,
                    const int* tmap, const float* target) {
        DevBuf<float> omega(n);
        forward(n, rp, ci, va, omega.get());
        const long nnz = host_nnz(rp, n);
        ewt_.reserve(std::max<long>(nnz, 1));
        gnn::k_gather_f<<<ceil_div(nnz, 256), 256>>>(nnz, tmap, ew_.get(), ewt_.get());
        // head
        const int hb = ceil_div(n, 256);
        part_.reserve(static_cast<size_t>(hb) * (gnn::kHid + 2));
        gnn::k_head_bwd<<<hb, 256>>>(n, h_[2].get(), omega.get(), target, params_.get(), dh_.get(), part_.get());
        small_.reserve(gnn::kHid + 2);
        gnn::k_sum_parts<<<1, 64>>>(hb, gnn::kHid + 2, part_.get(), small_.get());
        PILU_HIP(hipMemcpy(grads_.get() + gnn::kOffWo, small_.get(), sizeof(float) * (gnn::kHid + 1),
                           hipMemcpyDeviceToDevice));
        float loss = 0.f;
        PILU_HIP(hipMemcpy(&loss, small_.get() + gnn::kHid + 1, sizeof(float), hipMemcpyDeviceToHost));
        // layers, last to first; dh_ holds dL/dH_{l+1}
        const int wgs16 = ceil_div(n, 64);
        const long waves_w = ceil_div(n, gnn::kRowsPerWaveW);
        for (int l = gnn::kLayers - 1; l >= 0; --l) {
            const int D = gnn::kInDim[l];
            const float* hin = l == 0 ? feat_.get() : h_[l - 1].get();
            const int len = 2 * D * gnn::kHid + gnn::kHid;
            wpart_.reserve(static_cast<size_t>(ceil_div(waves_w, 4)) * 4 * len);
            if (D == gnn::kFeat) {
                gnn::k_dense_bwd_w<gnn::kFeat><<<ceil_div(waves_w, 4), 256>>>(n, hin, m_[l].get(), dh_.get(),
                                                                              z_[l].get(), wpart_.get());
            } else {
                gnn::k_dense_bwd_w<gnn::kHid><<<ceil_div(waves_w, 4), 256>>>(n, hin, m_[l].get(), dh_.get(),
                                                                             z_[l].get(), wpart_.get());
            }
            // weights then bias are contiguous in the partial: W (2D x 32), b (32)
            gnn::k_sum_parts<<<ceil_div(len, 256), 256>>>(static_cast<int>(waves_w), len, wpart_.get(),
                                                          small_big());
            PILU_HIP(hipMemcpy(grads_.get() + gnn::kOffW[l], small_big(), sizeof(float) * 2 * D * gnn::kHid,
                               hipMemcpyDeviceToDevice));
            PILU_HIP(hipMemcpy(grads_.get() + gnn::kOffB[l], small_big() + 2 * D * gnn::kHid,
                               sizeof(float) * gnn::kHid, hipMemcpyDeviceToDevice));
            if (l == 0) break;
            // dH_l = dX[:, :D] + W_graph^T dX[:, D:]
            if (D == gnn::kFeat) {
                gnn::k_dense_bwd_x<gnn::kFeat><<<wgs16, 256>>>(n, dh_.get(), z_[l].get(), wb_.get() + gnn::kOffW[l],
                                                               dh2_.get(), dm_.get());
                gnn::k_aggregate<gnn::kFeat><<<std::min(ceil_div(n, 16), gnn::kAggBlocks), 256>>>(n, rpt, cit, ewt_.get(), dm_.get(), dh2_.get(),
                                                                       true);
            } else {
                gnn::k_dense_bwd_x<gnn::kHid><<<wgs16, 256>>>(n, dh_.get(), z_[l].get(), wb_.get() + gnn::kOffW[l],
                                                              dh2_.get(), dm_.get());
                gnn::k_aggregate<gnn::kHid><<<std::min(ceil_div(n, 8), gnn::kAggBlocks), 256>>>(n, rpt, cit, ewt_.get(), dm_.get(), dh2_.get(),
                                                                     true);
            }
            std::swap(dh_, dh2_);
        }
        PILU_HIP(hipGetLastError());
        PILU_HIP(hipDeviceSynchronize());
        return loss;
    }

    double train_step(int n, const int* rp, const int* ci, const double* va, const int* rpt, const int* cit,
                      const int* tmap, const float* target, float lr) {
        const double loss = backward(n, rp, ci, va, rpt, cit, tmap, target);
        ++step_;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, static_cast<float>(step_));
        const float c2 = 1.f - std::pow(b2, static_cast<float>(step_));
        gnn::k_adam<<<ceil_div(gnn::kParams, 256), 256>>>(gnn::kParams, params_.get(), grads_.get(), m1_.get(),
                                                         m2_.get(), lr, b1, b2, 1e-8f, c1, c2);
        refresh_bf16();
        PILU_HIP(hipDeviceSynchronize());
        return loss;
    }

private:
    static long host_nnz(const int* d_rp, int n) {
        int v = 0;
        PILU_HIP(hipMemcpy(&v, d_rp + n, sizeof(int), hipMemcpyDeviceToHost));
        return v;
    }

    void refresh_bf16() {
        gnn::k_to_bf16<<<ceil_div(gnn::kParams, 256), 256>>>(gnn::kParams, params_.get(), wb_.get());
        PILU_HIP(hipGetLastError());
    }

    float* small_big() {
        big_.reserve(2 * gnn::kHid * gnn::kHid + gnn::kHid);
        return big_.get();
    }

    void prepare(int n, const int* rp, const int* ci, const double* va) {
        const long nnz = host_nnz(rp, n);
        const long nr = static_cast<long>(ceil_div(n, 64)) * 64;  // room for whole 16-row tiles
        dg_.reserve(n);
        feat_.reserve(nr * gnn::kFeat);
        ew_.reserve(std::max<long>(nnz, 1));
        for (int l = 0; l < gnn::kLayers; ++l) {
            m_[l].reserve(nr * gnn::kInDim[l]);
            z_[l].reserve(nr * gnn::kHid);
            h_[l].reserve(nr * gnn::kHid);
        }
        dh_.reserve(nr * gnn::kHid);
        dh2_.reserve(nr * gnn::kHid);
        dm_.reserve(nr * gnn::kHid);
        gnn::k_diag_values<<<ceil_div(n, 256), 256>>>(n, rp, ci, va, dg_.get());
        gnn::k_features<<<ceil_div(n, 256), 256>>>(n, rp, ci, va, dg_.get(), feat_.get(), ew_.get());
    }

    DevBuf<float> params_, grads_, m1_, m2_, feat_, ew_, ewt_, m_[gnn::kLayers], z_[gnn::kLayers],
        h_[gnn::kLayers], dh_, dh2_, dm_, part_, wpart_, small_, big_;
    DevBuf<unsigned short> wb_;
    DevBuf<double> dg_;
    long step_ = 0;
};

}  // namespace pilu

namespace {
template <typename F>
int gguard(const char* what, F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s: %s\n", what, e.what());
        return -1;
    }
}
pilu::GnnModel* Gm(GnnHdl h) { return static_cast<pilu::GnnModel*>(h); }
}  // namespace

extern "C" {

GnnHdl GnnCreateEx(void) {
    try {
        return new pilu::GnnModel;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "GnnCreate: %s\n", e.what());
        return nullptr;
    }
}

void GnnDestroyEx(GnnHdl h) { delete Gm(h); }

int GnnParameterCountEx(void) { return pilu::gnn::kParams; }

int GnnSetParametersEx(GnnHdl h, const float* host) {
    return gguard("GnnSetParameters", [&] { Gm(h)->set_params(host); return 0; });
}

int GnnGetParametersEx(GnnHdl h, float* host) {
    return gguard("GnnGetParameters", [&] { Gm(h)->get_params(host); return 0; });
}

int GnnPredictEx(GnnHdl h, const CsrMatrix* d_a, float* d_omega) {
    return gguard("GnnPredict", [&] {
        Gm(h)->forward(d_a->size, d_a->row_ptr, d_a->col_idx, d_a->value, d_omega);
        PILU_HIP(hipDeviceSynchronize());
        return 0;
    });
}

int GnnGradientsEx(GnnHdl h, const CsrMatrix* d_a, const CsrMatrix* d_at, const int* d_tmap, const float* d_target,
                   float* host_grads, double* loss) {
    return gguard("GnnGradients", [&] {
        const double l = Gm(h)->backward(d_a->size, d_a->row_ptr, d_a->col_idx, d_a->value, d_at->row_ptr,
                                         d_at->col_idx, d_tmap, d_target);
        if (loss) *loss = l;
        if (host_grads) Gm(h)->get_grads(host_grads);
        return 0;
    });
}

int GnnTrainStepEx(GnnHdl h, const CsrMatrix* d_a, const CsrMatrix* d_at, const int* d_tmap, const float* d_target,
                   float lr, double* loss) {
    return gguard("GnnTrainStep", [&] {
        const double l = Gm(h)->train_step(d_a->size, d_a->row_ptr, d_a->col_idx, d_a->value, d_at->row_ptr,
                                           d_at->col_idx, d_tmap, d_target, lr);
        if (loss) *loss = l;
        return 0;
    });
}

bool GmresSolveLearnedDeviceEx(GmresHdl gmres, IluSolverHdl ilu, GnnHdl gnn, const CsrMatrix* d_a,
                               const double* d_rhs, double* d_sol, int budget, int* iter, int* fallback) {
    auto* g = static_cast<lljbash::PreconditionedGmres*>(gmres);
    auto* s = static_cast<lljbash::IluSolver*>(ilu);
    if (fallback) *fallback = 0;
    if (iter) *iter = 0;
    try {
        const int n = d_a->size;
        pilu::DevBuf<float> omega(n);
        pilu::DevBuf<double> x0(n);
        PILU_HIP(hipMemcpy(x0.get(), d_sol, sizeof(double) * n, hipMemcpyDeviceToDevice));
        Gm(gnn)->forward(n, d_a->row_ptr, d_a->col_idx, d_a->value, omega.get());
        s->SetDiagonalCompensation(omega.get());
        bool fok = s->FactorizeDeviceFromA(d_a->value);
        s->SetDiagonalCompensation(nullptr);
        std::pair<bool, int> r{false, 0};
        if (fok) r = g->SolveDevice(d_a, d_rhs, d_sol);
        if (!fok || !r.first || r.second > budget) {
            PILU_HIP(hipMemcpy(d_sol, x0.get(), sizeof(double) * n, hipMemcpyDeviceToDevice));
            s->FactorizeDeviceFromA(d_a->value);
            r = g->SolveDevice(d_a, d_rhs, d_sol);
            if (fallback) *fallback = 1;
        }
        if (iter) *iter = r.second;
        return r.first;
    } catch (const std::exception& e) {
        s->SetDiagonalCompensation(nullptr);
        std::fprintf(stderr, "GmresSolveLearnedDevice: %s\n", e.what());
        return false;
    }
}

}  // extern "C"
