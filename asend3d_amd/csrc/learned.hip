// Learned factored preconditioner on the GPU (gfx950): node features, BF16
// MFMA message passing, the factor L (node blocks + edge couplings), the
// apply z = L L^T r, the training loss, its backward pass and Adam.
//
// Network (per node i; N(i) = nodes sharing an element with i):
//   h1 = relu(W0 x_i + b0)                               x_i: 8 features
//   a  = mean_{j in N(i)} h_j                            (layers 1..3)
//   h' = relu([Ws | Wn] [h_i ; a] + b)                   BF16 MFMA, FP32 accumulate
//   o_i = Wh h4_i + bh                                   24 outputs
// Factor (C_i = chol of the 3x3 diagonal block of A, Ahat_ij = C_i^-1 A_ij C_j^-T):
//   Lambda_ii = [[e^o0, 0, 0], [o3, e^o1, 0], [o4, o5, e^o2]]
//   Lambda_ij = reshape3x3(o_i[6..14] + o_j[15..23] + Wa vec(Ahat_ij))   j an edge neighbour
//   L_ij = C_i^-T Lambda_ij,   P = L L^T
// Training pair (r = A e, e), z = P r: the energy left by the best damped
// step x = omega z, relative to the initial energy,
//   loss = min_omega ||e - omega z||_A^2 / ||e||_A^2 = 1 - (z.r)^2 / ((z.A z)(e.A e)),
// which does not depend on the scale of P (the Chebyshev smoother that uses P
// sets its own scale from lambda_max(P A)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "asend_gpu.h"
#include "device_common.hpp"
#include "learned.hpp"

namespace asend {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef uint16_t bf16;

__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even; NaN stays NaN
__device__ __forceinline__ bf16 f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)((u >> 16) | 0x40u);
    return (bf16)((u + (((u >> 16) & 1u) + 0x7fffu)) >> 16);
}

constexpr int kNodesPerWave = 32;
constexpr int kWgChunk = 1024;  // nodes per block of the weight-gradient kernel
constexpr int kHeadPad = 32;    // head rows padded to one MFMA tile
constexpr float kExpClamp = 10.f;
constexpr int kMaxPairs = 2 * kMaxUses;  // outer-product pairs of one gradient

inline int blocks_for(int64_t n, int per) { return (int)((n + per - 1) / per); }

// ---------------------------------------------------------------------------
// Node features and operator scaling
// ---------------------------------------------------------------------------
__global__ void k_features(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ n2e_ptr,
                           const int32_t* __restrict__ n2e, const double* __restrict__ E, double mean_log_e,
                           const uint8_t* __restrict__ clamped, const int32_t* __restrict__ sptr,
                           const float* __restrict__ K, const int32_t* __restrict__ kdiag, float* __restrict__ X) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double s = 0.0, mn = 1e300, mx = -1e300;
    int ne = n2e_ptr[r + 1] - n2e_ptr[r];
    for (int k = n2e_ptr[r]; k < n2e_ptr[r + 1]; ++k) {
        double v = log(E[n2e[k]]) - mean_log_e;
        s += v;
        mn = fmin(mn, v);
        mx = fmax(mx, v);
    }
    float* x = X + (int64_t)kFeat * r;
    if (ne > 0) {
        x[0] = (float)(s / ne);
        x[1] = (float)mn;
        x[2] = (float)mx;
    } else {
        x[0] = x[1] = x[2] = 0.f;
    }
    x[3] = (float)(rp[r + 1] - rp[r] - 1) / 26.f;
    x[4] = clamped[r] ? 1.f : 0.f;
    float b[9];
    for (int c = 0; c < 9; ++c) b[c] = K[sell_index(sptr, r, kdiag[r], c)];
    x[5] = tanhf(logf(b[4] / b[0]));
    x[6] = tanhf(logf(b[8] / b[0]));
    x[7] = (fabsf(b[1]) + fabsf(b[2]) + fabsf(b[5])) / (b[0] + b[4] + b[8]);
}

// Ci = C^-1, C = chol(sym(B)), B the 3x3 diagonal block (FP64 arithmetic).
__global__ void k_prep_ci(int n, const int32_t* __restrict__ sptr, const float* __restrict__ K,
                          const int32_t* __restrict__ kdiag, float* __restrict__ Ci) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double B[9];
    for (int c = 0; c < 9; ++c) B[c] = K[sell_index(sptr, r, kdiag[r], c)];
    double b10 = 0.5 * (B[3] + B[1]), b20 = 0.5 * (B[6] + B[2]), b21 = 0.5 * (B[7] + B[5]);
    double c00 = sqrt(B[0]);
    double c10 = b10 / c00, c20 = b20 / c00;
    double c11 = sqrt(B[4] - c10 * c10);
    double c21 = (b21 - c20 * c10) / c11;
    double c22 = sqrt(B[8] - c20 * c20 - c21 * c21);
    double i00 = 1.0 / c00, i11 = 1.0 / c11, i22 = 1.0 / c22;
    double i10 = -c10 * i00 * i11;
    double i21 = -c21 * i11 * i22;
    double i20 = -(c20 * i00 + c21 * i10) * i22;
    float* o = Ci + 9 * (int64_t)r;
    o[0] = (float)i00;
    o[1] = 0.f;
    o[2] = 0.f;
    o[3] = (float)i10;
    o[4] = (float)i11;
    o[5] = 0.f;
    o[6] = (float)i20;
    o[7] = (float)i21;
    o[8] = (float)i22;
}

// Ahat_is = Ci_i A_ij Ci_j^T for each edge slot s (j = enb[s][i]).
__global__ void k_prep_ahat(int n, int W, const int32_t* __restrict__ enb, const int32_t* __restrict__ eslot,
                            const int32_t* __restrict__ sptr, const float* __restrict__ K,
                            const float* __restrict__ Ci, float* __restrict__ Ahat) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float ci[9];
    for (int c = 0; c < 9; ++c) ci[c] = Ci[9 * (int64_t)i + c];
    for (int s = 0; s < W; ++s) {
        const int j = enb[(int64_t)s * n + i];
        float out[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (j >= 0) {
            const int k = eslot[(int64_t)s * n + i];
            float a[9], cj[9], t[9];
            for (int c = 0; c < 9; ++c) {
                a[c] = K[sell_index(sptr, i, k, c)];
                cj[c] = Ci[9 * (int64_t)j + c];
            }
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q)
                    t[3 * p + q] = ci[3 * p] * a[q] + ci[3 * p + 1] * a[3 + q] + ci[3 * p + 2] * a[6 + q];
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q)
                    out[3 * p + q] = t[3 * p] * cj[3 * q] + t[3 * p + 1] * cj[3 * q + 1] + t[3 * p + 2] * cj[3 * q + 2];
        }
        for (int c = 0; c < 9; ++c) Ahat[((int64_t)s * 9 + c) * n + i] = out[c];
    }
}

// ---------------------------------------------------------------------------
// Forward network
// ---------------------------------------------------------------------------
// Layer 0 (FP32 VALU with BF16-valued weights): h1 = relu(W0 x + b0)
__global__ __launch_bounds__(256) void k_layer0(int n, const float* __restrict__ X, const bf16* __restrict__ W0,
                                                const float* __restrict__ b0, bf16* __restrict__ H) {
    int node = blockIdx.x * 4 + threadIdx.x / 64, c = threadIdx.x % 64;
    if (node >= n) return;
    const float* x = X + (int64_t)kFeat * node;
    float s = b0[c];
    for (int k = 0; k < kFeat; ++k) s += bf2f(W0[c * kFeat + k]) * x[k];
    H[(int64_t)kHid * node + c] = f2bf(fmaxf(s, 0.f));
}

// a_i = mean over neighbours (excluding i) of h_j.  16 lanes per node, 4
// features (8 B) per lane, so a wavefront gathers one 128 B row for each of
// its 4 nodes per load instruction; four neighbour rows in flight per step,
// summed in row order.
constexpr int kAggLanes = kHid / 4;  // lanes per node
__global__ __launch_bounds__(256) void k_aggregate(int n, const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ bc, const bf16* __restrict__ H,
                                                   bf16* __restrict__ A) {
    const int node = blockIdx.x * (256 / kAggLanes) + threadIdx.x / kAggLanes;
    const int c = 4 * (threadIdx.x % kAggLanes);
    if (node >= n) return;
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    int deg = 0;
    int k = rp[node];
    const int end = rp[node + 1];
    auto row = [&](int j) { return *reinterpret_cast<const uint2*>(H + (int64_t)kHid * j + c); };
    auto add = [&](uint2 v) {
        s[0] += __uint_as_float(v.x << 16);
        s[1] += __uint_as_float(v.x & 0xffff0000u);
        s[2] += __uint_as_float(v.y << 16);
        s[3] += __uint_as_float(v.y & 0xffff0000u);
    };
    for (; k + 4 <= end; k += 4) {
        int j[4];
        uint2 h[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) j[u] = bc[k + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) h[u] = (j[u] != node) ? row(j[u]) : make_uint2(0u, 0u);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (j[u] != node) {
                add(h[u]);
                ++deg;
            }
    }
    for (; k < end; ++k) {
        const int j = bc[k];
        if (j == node) continue;
        add(row(j));
        ++deg;
    }
    uint2 o;
    o.x = (uint32_t)f2bf(deg ? s[0] / (float)deg : 0.f) | ((uint32_t)f2bf(deg ? s[1] / (float)deg : 0.f) << 16);
    o.y = (uint32_t)f2bf(deg ? s[2] / (float)deg : 0.f) | ((uint32_t)f2bf(deg ? s[3] / (float)deg : 0.f) << 16);
    *reinterpret_cast<uint2*>(A + (int64_t)kHid * node + c) = o;
}

// Node GEMM on MFMA: out[node][col] = sum_k Arow[node][k] W[col][k] (+ epilogue)
//   Arow = [A1[node][0..K1) | A2[node][0..KA-K1)], BF16; W is BF16 [N][KA].
//   MODE 0: out BF16 [n][N] = relu(acc + bias)
//   MODE 1: out FP32 [n][N] = acc
//   MODE 2: out FP32 [n][kHead] = acc + bias, columns < kHead only
// One wavefront owns 32 nodes and all N columns (N/32 accumulator tiles of
// v_mfma_f32_32x32x16_bf16); its A fragments stay in registers across tiles.
// Lane layout (K = 16): A[m = lane%32][k = 8(lane/32) + i], B[k = 8(lane/32) + i][n = lane%32],
// D[m = 8(v/4) + 4(lane/32) + v%4][n = lane%32].
template <int KA, int K1, int N, int MODE>
__global__ __launch_bounds__(256) void k_node_gemm(int n, const bf16* __restrict__ A1, const bf16* __restrict__ A2,
                                                   const bf16* __restrict__ W, const float* __restrict__ bias,
                                                   void* __restrict__ out) {
    const int lane = threadIdx.x % 64;
    const int node0 = (blockIdx.x * 4 + threadIdx.x / 64) * kNodesPerWave;
    if (node0 >= n) return;
    const int arow = node0 + lane % 32;
    const int kq = 8 * (lane / 32);
    bf16x8 af[KA / 16];
#pragma unroll
    for (int s = 0; s < KA / 16; ++s) {
        const int k = 16 * s + kq;
        if (arow < n) {
            if (k < K1) af[s] = *(const bf16x8*)(A1 + (int64_t)K1 * arow + k);
            else af[s] = *(const bf16x8*)(A2 + (int64_t)(KA - K1) * arow + (k - K1));
        } else {
            af[s] = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
        }
    }
#pragma unroll
    for (int t = 0; t < N / 32; ++t) {
        const int col = 32 * t + lane % 32;
        f32x16 acc = {0};
#pragma unroll
        for (int s = 0; s < KA / 16; ++s) {
            bf16x8 bfr = *(const bf16x8*)(W + (int64_t)KA * col + 16 * s + kq);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[s], bfr, acc, 0, 0, 0);
        }
        const float bcol = (MODE == 1 || (MODE == 2 && col >= kHead)) ? 0.f : bias[col];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int row = node0 + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
            if (row >= n) continue;
            if (MODE == 0) {
                ((bf16*)out)[(int64_t)N * row + col] = f2bf(fmaxf(acc[v] + bcol, 0.f));
            } else if (MODE == 1) {
                ((float*)out)[(int64_t)N * row + col] = acc[v];
            } else if (col < kHead) {
                ((float*)out)[(int64_t)kHead * row + col] = acc[v] + bcol;
            }
        }
    }
}

__device__ __forceinline__ float clamp_o(float o) { return fminf(fmaxf(o, -kExpClamp), kExpClamp); }

// L_is = Ci_i^T Lambda_is, stored component-major per slot.
__global__ void k_factor(int n, int W, const float* __restrict__ O, const float* __restrict__ Ci,
                         const float* __restrict__ Ahat, const int32_t* __restrict__ enb, const bf16* __restrict__ Wa,
                         float* __restrict__ Ls) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float ci[9];
    for (int c = 0; c < 9; ++c) ci[c] = Ci[9 * (int64_t)i + c];
    const float* o = O + (int64_t)kHead * i;
    float lam[9] = {expf(clamp_o(o[0])), 0.f, 0.f, o[3], expf(clamp_o(o[1])), 0.f, o[4], o[5], expf(clamp_o(o[2]))};
    for (int s = 0; s <= W; ++s) {
        if (s > 0) {
            const int j = enb[(int64_t)(s - 1) * n + i];
            if (j < 0) {
                for (int c = 0; c < 9; ++c) Ls[((int64_t)s * 9 + c) * n + i] = 0.f;
                continue;
            }
            const float* oj = O + (int64_t)kHead * j;
            float ah[9];
            for (int c = 0; c < 9; ++c) ah[c] = Ahat[((int64_t)(s - 1) * 9 + c) * n + i];
            for (int c = 0; c < 9; ++c) {
                float v = o[6 + c] + oj[15 + c];
                for (int d = 0; d < 9; ++d) v += bf2f(Wa[9 * c + d]) * ah[d];
                lam[c] = v;
            }
        }
        // (Ci^T Lambda)[a][b] = sum_k Ci[k][a] Lambda[k][b]
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                Ls[((int64_t)s * 9 + 3 * a + b) * n + i] =
                    ci[a] * lam[b] + ci[3 + a] * lam[3 + b] + ci[6 + a] * lam[6 + b];
    }
}

// L^T in the same row layout as L: LT block s of row j is (L_ij)^T for the
// node i of slot s (s = 0: i = j; s = 1 + k: i = enb[k][j], whose slot for j
// is 1 + erev[k][j]).  Built once per factor so that both halves of
// z = L (L^T r) are coalesced row gathers.
__global__ __launch_bounds__(256) void k_factor_transpose(int n, int W, const float* __restrict__ Ls,
                                                          const int32_t* __restrict__ enb,
                                                          const int32_t* __restrict__ erev, float* __restrict__ LT) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) LT[(int64_t)(3 * a + b) * n + j] = Ls[(int64_t)(3 * b + a) * n + j];
    for (int s = 0; s < W; ++s) {
        const int i = enb[(int64_t)s * n + j];
        const int slot = i < 0 ? 0 : 1 + erev[(int64_t)s * n + j];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                LT[((int64_t)(s + 1) * 9 + 3 * a + b) * n + j] =
                    i < 0 ? 0.f : Ls[((int64_t)slot * 9 + 3 * b + a) * n + i];
    }
}

// FP16 copy of a factor in the row layout: each node's row is scaled by the
// largest magnitude among its (1 + W) blocks, which keeps every stored entry
// in [-1, 1] whatever the moduli and penalty (the node blocks carry the
// inverse Cholesky factor of the operator's diagonal block); sc restores it.
__global__ __launch_bounds__(256) void k_factor_to16(int n, int W, const float* __restrict__ L,
                                                     __half* __restrict__ L16, float* __restrict__ sc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nb = 9 * (W + 1);
    float m = 0.f;
    for (int k = 0; k < nb; ++k) m = fmaxf(m, fabsf(L[(int64_t)k * n + i]));
    const float s = m > 0.f ? 1.f / m : 1.f;
    for (int k = 0; k < nb; ++k) L16[(int64_t)k * n + i] = __float2half(L[(int64_t)k * n + i] * s);
    sc[i] = m > 0.f ? m : 1.f;
}

// z_i = sc_i sum_s L16_is t_{col(s)} with the FP16 copy.  MODE 0 stores z;
// MODE 1 / 2 fuse the first / a later Chebyshev step: v = c2 z (+ c1 d),
// d = v, x = v (x += v).
template <int MODE>
__global__ __launch_bounds__(256) void k_factor_l16(int n, int W, const __half* __restrict__ L16,
                                                    const float* __restrict__ sc, const int32_t* __restrict__ enb,
                                                    const float* __restrict__ t, float* __restrict__ z,
                                                    const float* __restrict__ coef, float* __restrict__ d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float t0 = t[3 * i], t1 = t[3 * i + 1], t2 = t[3 * i + 2];
    float b[9];
    for (int c = 0; c < 9; ++c) b[c] = __half2float(L16[(int64_t)c * n + i]);
    float z0 = b[0] * t0 + b[1] * t1 + b[2] * t2;
    float z1 = b[3] * t0 + b[4] * t1 + b[5] * t2;
    float z2 = b[6] * t0 + b[7] * t1 + b[8] * t2;
    for (int s = 0; s < W; ++s) {
        const int j = enb[(int64_t)s * n + i];
        if (j < 0) continue;
        for (int c = 0; c < 9; ++c) b[c] = __half2float(L16[((int64_t)(s + 1) * 9 + c) * n + i]);
        t0 = t[3 * j];
        t1 = t[3 * j + 1];
        t2 = t[3 * j + 2];
        z0 += b[0] * t0 + b[1] * t1 + b[2] * t2;
        z1 += b[3] * t0 + b[4] * t1 + b[5] * t2;
        z2 += b[6] * t0 + b[7] * t1 + b[8] * t2;
    }
    const float m = sc[i];
    const float zz[3] = {m * z0, m * z1, m * z2};
    if (MODE == 0) {
        for (int c = 0; c < 3; ++c) z[3 * i + c] = zz[c];
        return;
    }
    const float c1 = coef[0], c2 = coef[1];
    for (int c = 0; c < 3; ++c) {
        const float v = (MODE == 1) ? c2 * zz[c] : c1 * d[3 * i + c] + c2 * zz[c];
        d[3 * i + c] = v;
        if (MODE == 1) z[3 * i + c] = v;
        else z[3 * i + c] += v;
    }
}

// z_i = sum_s L_is t_{col(s)};  SUB: z = e - (L t)
template <bool SUB>
__global__ __launch_bounds__(256) void k_factor_l(int n, int W, const float* __restrict__ Ls,
                                                  const int32_t* __restrict__ enb, const float* __restrict__ t,
                                                  const float* __restrict__ e, float* __restrict__ z) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float t0 = t[3 * i], t1 = t[3 * i + 1], t2 = t[3 * i + 2];
    // (plain loads: the 93 MB factor of the 1M-DOF mesh stays cache-resident
    // between products; non-temporal loads measured 3.7% slower)
    float b[9];
    for (int c = 0; c < 9; ++c) b[c] = Ls[(int64_t)c * n + i];
    float z0 = b[0] * t0 + b[1] * t1 + b[2] * t2;
    float z1 = b[3] * t0 + b[4] * t1 + b[5] * t2;
    float z2 = b[6] * t0 + b[7] * t1 + b[8] * t2;
    for (int s = 0; s < W; ++s) {
        const int j = enb[(int64_t)s * n + i];
        if (j < 0) continue;
        for (int c = 0; c < 9; ++c) b[c] = Ls[((int64_t)(s + 1) * 9 + c) * n + i];
        t0 = t[3 * j];
        t1 = t[3 * j + 1];
        t2 = t[3 * j + 2];
        z0 += b[0] * t0 + b[1] * t1 + b[2] * t2;
        z1 += b[3] * t0 + b[4] * t1 + b[5] * t2;
        z2 += b[6] * t0 + b[7] * t1 + b[8] * t2;
    }
    if (SUB) {
        z[3 * i] = e[3 * i] - z0;
        z[3 * i + 1] = e[3 * i + 1] - z1;
        z[3 * i + 2] = e[3 * i + 2] - z2;
    } else {
        z[3 * i] = z0;
        z[3 * i + 1] = z1;
        z[3 * i + 2] = z2;
    }
}

// ---------------------------------------------------------------------------
// Training sample and loss
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t mix32(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return (uint32_t)x;
}

// Gaussian sample (Box-Muller), zero on penalised nodes
__global__ void k_randn(int n_nodes, uint64_t seed, const uint8_t* __restrict__ clamped, float* __restrict__ e) {
    const int64_t n = 3 * (int64_t)n_nodes;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t a = mix32(seed * 0x9E3779B97F4A7C15ull + 2 * (uint64_t)i);
        uint32_t b = mix32(seed * 0x9E3779B97F4A7C15ull + 2 * (uint64_t)i + 1);
        float u1 = ((float)a + 1.0f) * (1.0f / 4294967296.0f);
        float u2 = (float)b * (1.0f / 4294967296.0f);
        e[i] = clamped[i / 3] ? 0.f : sqrtf(-2.f * logf(u1)) * cosf(6.2831853f * u2);
    }
}

__global__ __launch_bounds__(kBlock) void k_dot3f(int64_t n, const float* __restrict__ e, const float* __restrict__ r,
                                                  const float* __restrict__ z, const float* __restrict__ q,
                                                  double* partials) {
    double t[3] = {0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        t[0] += (double)e[i] * r[i];
        t[1] += (double)z[i] * r[i];
        t[2] += (double)z[i] * q[i];
    }
    block_partials<3>(t, partials);
}

// c = e.Ae, a = z.r, b = z.Az;  sv = {c, a, b, loss, alpha, beta} with
// loss = 1 - a^2/(b c), alpha = dloss/da = -2a/(b c), beta = -2 dloss/db = 2a^2/(b^2 c)
__global__ __launch_bounds__(kBlock) void k_loss_fin(const double* __restrict__ partials, int nparts,
                                                     double* __restrict__ sv) {
    double t[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        t[0] += partials[3 * b];
        t[1] += partials[3 * b + 1];
        t[2] += partials[3 * b + 2];
    }
    block_sum<3>(t);
    if (threadIdx.x == 0) {
        const double c = t[0], a = t[1], b = t[2];
        sv[0] = c;
        sv[1] = a;
        sv[2] = b;
        sv[3] = 1.0 - a * a / (b * c);
        sv[4] = -2.0 * a / (b * c);
        sv[5] = 2.0 * a * a / (b * b * c);
    }
}

// Left vectors of dLoss/dL = g1 t^T + g2 u^T:  g1 = 2 alpha r + beta q,  g2 = beta r
__global__ void k_grad_vectors(int64_t n, const double* __restrict__ sv, const float* __restrict__ r,
                               const float* __restrict__ q, float* __restrict__ g1, float* __restrict__ g2) {
    const float al = (float)sv[4], be = (float)sv[5];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        g1[i] = 2.f * al * r[i] + be * q[i];
        g2[i] = be * r[i];
    }
}

// dLoss/dL_ij = sum_q a_q[i] b_q[j]^T, the outer-product pairs (a_q, b_q)
// describing the loss gradient with respect to L.  Chained through
// L_ij = Ci_i^T Lambda_ij to the head outputs: dO[i][0..5] (node block),
// dO[i][6..14] (sum over i's edges), per-edge gradients DOE (gathered by
// k_grad_to into dO[j][15..23]) and the per-node contribution to dWa (Gea).
struct PairSet {
    int q;
    const float* a[kMaxPairs];
    const float* b[kMaxPairs];
};

__global__ void k_grad_factor(int n, int W, const PairSet ps, const float* __restrict__ O,
                              const float* __restrict__ Ci, const float* __restrict__ Ahat,
                              const int32_t* __restrict__ enb, float* __restrict__ dO, float* __restrict__ DOE,
                              float* __restrict__ Gea) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float ci[9];
    for (int c = 0; c < 9; ++c) ci[c] = Ci[9 * (int64_t)i + c];
    float dlam[9], dl[9];
    auto block_grad = [&](int j) {
        for (int c = 0; c < 9; ++c) dl[c] = 0.f;
        for (int q = 0; q < ps.q; ++q) {
            const float* aq = ps.a[q];
            const float* bq = ps.b[q];
            const float ai[3] = {aq[3 * i], aq[3 * i + 1], aq[3 * i + 2]};
            const float bj[3] = {bq[3 * j], bq[3 * j + 1], bq[3 * j + 2]};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) dl[3 * a + b] += ai[a] * bj[b];
        }
        // dLambda = Ci dL
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                dlam[3 * a + b] = ci[3 * a] * dl[b] + ci[3 * a + 1] * dl[3 + b] + ci[3 * a + 2] * dl[6 + b];
    };
    block_grad(i);
    const float* o = O + (int64_t)kHead * i;
    float* d = dO + (int64_t)kHead * i;
    for (int k = 0; k < 3; ++k) {
        const bool live = o[k] > -kExpClamp && o[k] < kExpClamp;
        d[k] = live ? dlam[4 * k] * expf(o[k]) : 0.f;
    }
    d[3] = dlam[3];
    d[4] = dlam[6];
    d[5] = dlam[7];
    float from[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    float gea[81];
    for (int c = 0; c < 81; ++c) gea[c] = 0.f;
    for (int sl = 0; sl < W; ++sl) {
        const int j = enb[(int64_t)sl * n + i];
        if (j < 0) {
            for (int c = 0; c < 9; ++c) DOE[((int64_t)sl * 9 + c) * n + i] = 0.f;
            continue;
        }
        block_grad(j);
        float ah[9];
        for (int c = 0; c < 9; ++c) ah[c] = Ahat[((int64_t)sl * 9 + c) * n + i];
        for (int c = 0; c < 9; ++c) {
            from[c] += dlam[c];
            DOE[((int64_t)sl * 9 + c) * n + i] = dlam[c];
            for (int q = 0; q < 9; ++q) gea[9 * c + q] += dlam[c] * ah[q];
        }
    }
    for (int c = 0; c < 9; ++c) d[6 + c] = from[c];
    for (int c = 0; c < 81; ++c) Gea[81 * (int64_t)i + c] = gea[c];
}

// dO[j][15..23] = sum over edges (i -> j) of the edge gradient stored at i
__global__ void k_grad_to(int n, int W, const int32_t* __restrict__ enb, const int32_t* __restrict__ erev,
                          const float* __restrict__ DOE, float* __restrict__ dO) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = 0; s < W; ++s) {
        const int i = enb[(int64_t)s * n + j];
        if (i < 0) continue;
        const int sl = erev[(int64_t)s * n + j];
        for (int c = 0; c < 9; ++c) acc[c] += DOE[((int64_t)sl * 9 + c) * n + i];
    }
    for (int c = 0; c < 9; ++c) dO[(int64_t)kHead * j + 15 + c] = acc[c];
}

// ---------------------------------------------------------------------------
// Backward network
// ---------------------------------------------------------------------------
// dH4 = Wh^T dO  (FP32, BF16-valued weights)
__global__ __launch_bounds__(256) void k_head_back(int n, const float* __restrict__ dO, const bf16* __restrict__ Wh,
                                                   float* __restrict__ dH) {
    int node = blockIdx.x * 4 + threadIdx.x / 64, c = threadIdx.x % 64;
    if (node >= n) return;
    float s = 0.f;
    for (int j = 0; j < kHead; ++j) s += bf2f(Wh[j * kHid + c]) * dO[(int64_t)kHead * node + j];
    dH[(int64_t)kHid * node + c] = s;
}

// Per-block partial sums over node chunks of
//   g[row * ncol + col] = sum_node U[node][row] * V[node][col]   and
//   g[nrow * ncol + row] = sum_node U[node][row]                 (bias rows)
// U is FP32 [n][nrow]; V is FP32 (vkind 0) or BF16 (vkind 1) [n][ncol].
__global__ __launch_bounds__(256) void k_small_wgrad(int n, int nrow, int ncol, const float* __restrict__ U,
                                                     const void* __restrict__ V, int vkind,
                                                     float* __restrict__ partials) {
    const int np = nrow * ncol + nrow;
    const int chunk = (n + gridDim.x - 1) / gridDim.x;
    const int n0 = blockIdx.x * chunk, n1 = min(n, n0 + chunk);
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
        float s = 0.f;
        if (p < nrow * ncol) {
            const int row = p / ncol, col = p % ncol;
            for (int i = n0; i < n1; ++i) {
                float v = vkind ? bf2f(((const bf16*)V)[(int64_t)ncol * i + col])
                                : ((const float*)V)[(int64_t)ncol * i + col];
                s += U[(int64_t)nrow * i + row] * v;
            }
        } else {
            const int row = p - nrow * ncol;
            for (int i = n0; i < n1; ++i) s += U[(int64_t)nrow * i + row];
        }
        partials[(int64_t)np * blockIdx.x + p] = s;
    }
}

// g[p] = sum_b partials[b][p]  (fixed order: deterministic)
__global__ void k_reduce_partials(int nparts, int np, const float* __restrict__ partials, float* __restrict__ g) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    float s = 0.f;
    for (int b = 0; b < nparts; ++b) s += partials[(int64_t)np * b + p];
    g[p] = s;
}

// dz = bf16(dH * [h_out > 0])
__global__ void k_relu_back(int64_t m, const float* __restrict__ dH, const bf16* __restrict__ Hout,
                            bf16* __restrict__ dz) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        dz[i] = f2bf(bf2f(Hout[i]) > 0.f ? dH[i] : 0.f);
}

// FP32 variant for layer 0: dz0 = dH * [h1 > 0]
__global__ void k_relu_back_f32(int64_t m, const float* __restrict__ dH, const bf16* __restrict__ Hout,
                                float* __restrict__ dz) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        dz[i] = bf2f(Hout[i]) > 0.f ? dH[i] : 0.f;
}

// Weight gradient on MFMA: dW[o][i] (64 x 128) = sum_node dz[node][o] * X[node][i],
// X = [Hin | Ain]; db[o] = sum_node dz[node][o].  Each block reduces one chunk
// of nodes, staging 16-node slabs of dz and X in LDS; its 8 output tiles are
// split over the 4 wavefronts (wave w: rows 32 (w/2).., columns 64 (w%2)..).
// Per-block partials are summed afterwards in a fixed order.
__global__ __launch_bounds__(256) void k_wgrad_mfma(int n, const bf16* __restrict__ dz, const bf16* __restrict__ Hin,
                                                    const bf16* __restrict__ Ain, float* __restrict__ partials) {
    __shared__ bf16 sdz[16][kHid];
    __shared__ bf16 sx[16][2 * kHid];
    __shared__ float sdb[4][kHid];
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int n0 = blockIdx.x * kWgChunk, n1 = min(n, n0 + kWgChunk);
    const int to = w / 2;
    const int kq = 8 * (lane / 32);
    f32x16 acc0 = {0}, acc1 = {0};
    float dbp = 0.f;
    for (int base = n0; base < n1; base += 16) {
        __syncthreads();
        for (int t = threadIdx.x; t < 16 * kHid; t += 256) {
            int rr = t / kHid, c = t % kHid, node = base + rr;
            sdz[rr][c] = node < n1 ? dz[(int64_t)kHid * node + c] : (bf16)0;
        }
        for (int t = threadIdx.x; t < 16 * 2 * kHid; t += 256) {
            int rr = t / (2 * kHid), c = t % (2 * kHid), node = base + rr;
            bf16 v = 0;
            if (node < n1) v = c < kHid ? Hin[(int64_t)kHid * node + c] : Ain[(int64_t)kHid * node + (c - kHid)];
            sx[rr][c] = v;
        }
        __syncthreads();
        for (int rr = 4 * w; rr < 4 * w + 4; ++rr) dbp += bf2f(sdz[rr][lane]);
        bf16x8 a, b0, b1;
        for (int i = 0; i < 8; ++i) {
            a[i] = (short)sdz[kq + i][32 * to + lane % 32];
            b0[i] = (short)sx[kq + i][64 * (w % 2) + lane % 32];
            b1[i] = (short)sx[kq + i][64 * (w % 2) + 32 + lane % 32];
        }
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b1, acc1, 0, 0, 0);
    }
    float* P = partials + (int64_t)kLayerSize * blockIdx.x;
    const int i0 = 64 * (w % 2) + lane % 32;
    for (int v = 0; v < 16; ++v) {
        const int o = 32 * to + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
        P[(int64_t)2 * kHid * o + i0] = acc0[v];
        P[(int64_t)2 * kHid * o + i0 + 32] = acc1[v];
    }
    sdb[w][lane] = dbp;
    __syncthreads();
    if (threadIdx.x < kHid)
        P[kHid * 2 * kHid + threadIdx.x] =
            sdb[0][threadIdx.x] + sdb[1][threadIdx.x] + sdb[2][threadIdx.x] + sdb[3][threadIdx.x];
}

// dHin = dX[:, :64] + aggregation^T(dX[:, 64:])
__global__ __launch_bounds__(256) void k_agg_back(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ bc,
                                                  const float* __restrict__ dX, float* __restrict__ dH) {
    int node = blockIdx.x * 4 + threadIdx.x / 64, c = threadIdx.x % 64;
    if (node >= n) return;
    float s = dX[(int64_t)2 * kHid * node + c];
    for (int k = rp[node]; k < rp[node + 1]; ++k) {
        int i = bc[k];
        if (i == node) continue;
        int deg = rp[i + 1] - rp[i] - 1;
        s += dX[(int64_t)2 * kHid * i + kHid + c] / (float)deg;
    }
    dH[(int64_t)kHid * node + c] = s;
}

// Adam on the FP32 master parameters
__global__ void k_adam(int np, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1t, float b2t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float gi = g[i];
    if (!(fabsf(gi) < 3.0e38f)) gi = 0.f;  // a non-finite sample gradient does not poison the moments
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / (1.f - b1t)) / (sqrtf(vi / (1.f - b2t)) + eps);
}

// BF16 copies of the parameters: flat (same layout), per-layer transposed
// Wcat^T [128][64] for dX = dz Wcat, and the head padded to 32 rows.
__global__ void k_refresh_bf16(const float* __restrict__ p, bf16* __restrict__ wbf, bf16* __restrict__ wcatT,
                               bf16* __restrict__ whp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kHeadPad * kHid) whp[i] = i < kHead * kHid ? f2bf(p[kOffWh + i]) : (bf16)0;
    if (i >= kNumParams) return;
    wbf[i] = f2bf(p[i]);
    if (i >= kOffL1 && i < kOffWh) {
        int l = (i - kOffL1) / kLayerSize, off = (i - kOffL1) % kLayerSize;
        if (off < kHid * 2 * kHid) {
            int o = off / (2 * kHid), in = off % (2 * kHid);
            wcatT[(int64_t)l * 2 * kHid * kHid + (int64_t)kHid * in + o] = f2bf(p[i]);
        }
    }
}

}  // namespace

void launch_factor_apply16(hipStream_t st, const Factor& F, const float* r, float* t, float* z) {
    const int nb = blocks_for(F.n, 256);
    hipLaunchKernelGGL((k_factor_l16<0>), dim3(nb), dim3(256), 0, st, F.n, F.n_edge, F.LsT16, F.scT16, F.enb, r, t,
                       (const float*)nullptr, (float*)nullptr);
    hipLaunchKernelGGL((k_factor_l16<0>), dim3(nb), dim3(256), 0, st, F.n, F.n_edge, F.Ls16, F.sc16, F.enb, t, z,
                       (const float*)nullptr, (float*)nullptr);
}

void launch_factor_cheb16(hipStream_t st, const Factor& F, const float* r, float* t, const float* coef, float* d,
                          float* x, bool first) {
    const int nb = blocks_for(F.n, 256);
    hipLaunchKernelGGL((k_factor_l16<0>), dim3(nb), dim3(256), 0, st, F.n, F.n_edge, F.LsT16, F.scT16, F.enb, r, t,
                       (const float*)nullptr, (float*)nullptr);
    if (first)
        hipLaunchKernelGGL((k_factor_l16<1>), dim3(nb), dim3(256), 0, st, F.n, F.n_edge, F.Ls16, F.sc16, F.enb, t, x,
                           coef, d);
    else
        hipLaunchKernelGGL((k_factor_l16<2>), dim3(nb), dim3(256), 0, st, F.n, F.n_edge, F.Ls16, F.sc16, F.enb, t, x,
                           coef, d);
}

void launch_factor_apply(hipStream_t st, const Factor& F, const float* r, float* t, float* z) {
    const int nb = blocks_for(F.n, 256);
    hipLaunchKernelGGL((k_factor_l<false>), dim3(nb), dim3(256), 0, st, F.n, F.n_edge, F.LsT, F.enb, r,
                       (const float*)nullptr, t);
    hipLaunchKernelGGL((k_factor_l<false>), dim3(nb), dim3(256), 0, st, F.n, F.n_edge, F.Ls, F.enb, t,
                       (const float*)nullptr, z);
}

// ===========================================================================

struct Learned {
    std::string err;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16 *wbf = nullptr, *wcatT = nullptr, *whp = nullptr;
    int64_t step = 0;
    int cap = 0, cap_w = -1;  // node / edge-slot capacity of the buffers
    float* X = nullptr;
    bf16* H[kLayers + 1] = {nullptr, nullptr, nullptr, nullptr};
    bf16* A[kLayers] = {nullptr, nullptr, nullptr};
    float *dH = nullptr, *dX = nullptr, *O = nullptr, *dO = nullptr, *dz0 = nullptr;
    bf16* dz = nullptr;
    float *Ci = nullptr, *Ahat = nullptr, *Ls = nullptr, *LsT = nullptr, *DOE = nullptr, *Gea = nullptr;
    __half *Ls16 = nullptr, *LsT16 = nullptr;  // FP16 copies for the smoother and their row scales
    float *sc16 = nullptr, *scT16 = nullptr;
    float *e = nullptr, *r = nullptr, *t = nullptr, *y = nullptr, *w = nullptr, *u = nullptr;
    float* wpart = nullptr;
    float* pairbuf = nullptr;  // L^T images of the factor-use vectors (2 kMaxUses x 3n)
    double* partials = nullptr;
    double* sv = nullptr;
};

namespace {

#define LCHK(expr)                                                       \
    do {                                                                 \
        hipError_t e_ = (expr);                                          \
        if (e_ != hipSuccess) {                                          \
            lp->err = std::string(#expr) + ": " + hipGetErrorString(e_); \
            return ASEND_ERR_HIP;                                        \
        }                                                                \
    } while (0)

template <typename T>
void lfree(T*& q) {
    if (q) (void)hipFree((void*)q);
    q = nullptr;
}

void free_work(Learned* lp) {
    lfree(lp->X);
    for (auto& h : lp->H) lfree(h);
    for (auto& a : lp->A) lfree(a);
    lfree(lp->dH);
    lfree(lp->dX);
    lfree(lp->O);
    lfree(lp->dO);
    lfree(lp->dz0);
    lfree(lp->dz);
    lfree(lp->Ci);
    lfree(lp->Ahat);
    lfree(lp->Ls);
    lfree(lp->LsT);
    lfree(lp->Ls16);
    lfree(lp->LsT16);
    lfree(lp->sc16);
    lfree(lp->scT16);
    lfree(lp->DOE);
    lfree(lp->Gea);
    lfree(lp->e);
    lfree(lp->r);
    lfree(lp->t);
    lfree(lp->y);
    lfree(lp->w);
    lfree(lp->u);
    lfree(lp->wpart);
    lfree(lp->pairbuf);
    lp->cap = 0;
    lp->cap_w = -1;
}

int ensure_cap(Learned* lp, int n, int W) {
    if (n <= lp->cap && W <= lp->cap_w) return ASEND_OK;
    free_work(lp);
    const int64_t c = n;
    const int64_t wslots = std::max(W, 1);
    LCHK(hipMalloc((void**)&lp->X, c * kFeat * sizeof(float)));
    for (auto& h : lp->H) LCHK(hipMalloc((void**)&h, c * kHid * sizeof(bf16)));
    for (auto& a : lp->A) LCHK(hipMalloc((void**)&a, c * kHid * sizeof(bf16)));
    LCHK(hipMalloc((void**)&lp->dH, c * kHid * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->dX, c * 2 * kHid * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->O, c * kHead * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->dO, c * kHead * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->dz0, c * kHid * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->dz, c * kHid * sizeof(bf16)));
    LCHK(hipMalloc((void**)&lp->Ci, c * 9 * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->Ahat, wslots * 9 * c * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->Ls, (wslots + 1) * 9 * c * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->LsT, (wslots + 1) * 9 * c * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->Ls16, (wslots + 1) * 9 * c * sizeof(__half)));
    LCHK(hipMalloc((void**)&lp->LsT16, (wslots + 1) * 9 * c * sizeof(__half)));
    LCHK(hipMalloc((void**)&lp->sc16, c * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->scT16, c * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->DOE, wslots * 9 * c * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->Gea, c * 81 * sizeof(float)));
    for (float** q : {&lp->e, &lp->r, &lp->t, &lp->y, &lp->w, &lp->u})
        LCHK(hipMalloc((void**)q, c * 3 * sizeof(float)));
    const int nwg = std::max(blocks_for(n, kWgChunk), 64);
    LCHK(hipMalloc((void**)&lp->wpart, (size_t)nwg * kLayerSize * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->pairbuf, (size_t)2 * kMaxUses * c * 3 * sizeof(float)));
    lp->cap = n;
    lp->cap_w = W;
    return ASEND_OK;
}

Factor factor_of(const Learned* lp, const LearnedGraph& g) {
    Factor F;
    F.n = g.n_nodes;
    F.n_edge = g.n_edge;
    F.Ls = lp->Ls;
    F.LsT = lp->LsT;
    F.Ls16 = lp->Ls16;
    F.LsT16 = lp->LsT16;
    F.sc16 = lp->sc16;
    F.scT16 = lp->scT16;
    F.enb = g.enb;
    F.erev = g.erev;
    return F;
}

void forward(Learned* lp, hipStream_t st, const LearnedGraph& g) {
    const int n = g.n_nodes, W = g.n_edge;
    const int nb = blocks_for(n, 256);
    hipLaunchKernelGGL(k_features, dim3(nb), dim3(256), 0, st, n, g.rp, g.n2e_ptr, g.n2e, g.E, g.mean_log_e,
                       g.clamped, g.op.sptr, g.K32, g.kdiag, lp->X);
    hipLaunchKernelGGL(k_prep_ci, dim3(nb), dim3(256), 0, st, n, g.op.sptr, g.K32, g.kdiag, lp->Ci);
    hipLaunchKernelGGL(k_prep_ahat, dim3(nb), dim3(256), 0, st, n, W, g.enb, g.eslot, g.op.sptr, g.K32, lp->Ci,
                       lp->Ahat);
    hipLaunchKernelGGL(k_layer0, dim3(blocks_for(n, 4)), dim3(256), 0, st, n, lp->X, lp->wbf + kOffW0,
                       lp->p + kOffB0, lp->H[0]);
    const int gb = blocks_for(n, 4 * kNodesPerWave);
    for (int l = 0; l < kLayers; ++l) {
        hipLaunchKernelGGL(k_aggregate, dim3(blocks_for(n, 256 / kAggLanes)), dim3(256), 0, st, n, g.rp, g.bc,
                           lp->H[l], lp->A[l]);
        const int off = kOffL1 + l * kLayerSize;
        hipLaunchKernelGGL((k_node_gemm<2 * kHid, kHid, kHid, 0>), dim3(gb), dim3(256), 0, st, n, lp->H[l], lp->A[l],
                           lp->wbf + off, lp->p + off + kHid * 2 * kHid, (void*)lp->H[l + 1]);
    }
    hipLaunchKernelGGL((k_node_gemm<kHid, kHid, kHeadPad, 2>), dim3(gb), dim3(256), 0, st, n, lp->H[kLayers],
                       lp->H[kLayers], lp->whp, lp->p + kOffBh, (void*)lp->O);
    hipLaunchKernelGGL(k_factor, dim3(nb), dim3(256), 0, st, n, W, lp->O, lp->Ci, lp->Ahat, g.enb, lp->wbf + kOffWa,
                       lp->Ls);
    hipLaunchKernelGGL(k_factor_transpose, dim3(nb), dim3(256), 0, st, n, W, lp->Ls, g.enb, g.erev, lp->LsT);
    hipLaunchKernelGGL(k_factor_to16, dim3(nb), dim3(256), 0, st, n, W, lp->Ls, lp->Ls16, lp->sc16);
    hipLaunchKernelGGL(k_factor_to16, dim3(nb), dim3(256), 0, st, n, W, lp->LsT, lp->LsT16, lp->scT16);
}

void backward_from_pairs(Learned* lp, hipStream_t st, const LearnedGraph& g, const PairSet& ps);

// loss for the sample in lp->e and the parameter gradient into lp->g
void loss_and_backward(Learned* lp, hipStream_t st, const LearnedGraph& g) {
    const int n = g.n_nodes, W = g.n_edge;
    const int64_t nd = 3 * (int64_t)n;
    const int nb = blocks_for(n, 256);
    // r = A e, t = L^T r, z = L t, q = A z
    launch_sell3<float>(st, 0, g.op, g.K32, lp->e, lp->r, nullptr, nullptr, nullptr);
    hipLaunchKernelGGL((k_factor_l<false>), dim3(nb), dim3(256), 0, st, n, W, lp->LsT, g.enb, lp->r,
                       (const float*)nullptr, lp->t);
    hipLaunchKernelGGL((k_factor_l<false>), dim3(nb), dim3(256), 0, st, n, W, lp->Ls, g.enb, lp->t,
                       (const float*)nullptr, lp->y);
    launch_sell3<float>(st, 0, g.op, g.K32, lp->y, lp->w, nullptr, nullptr, nullptr);
    hipLaunchKernelGGL(k_dot3f, dim3(kVecBlocks), dim3(kBlock), 0, st, nd, lp->e, lp->r, lp->y, lp->w,
                       lp->partials);
    hipLaunchKernelGGL(k_loss_fin, dim3(1), dim3(kBlock), 0, st, lp->partials, kVecBlocks, lp->sv);
    // u = L^T q ; left vectors (reusing e and y, no longer needed) ; factor gradient
    hipLaunchKernelGGL((k_factor_l<false>), dim3(nb), dim3(256), 0, st, n, W, lp->LsT, g.enb, lp->w,
                       (const float*)nullptr, lp->u);
    hipLaunchKernelGGL(k_grad_vectors, dim3(kVecBlocks), dim3(256), 0, st, nd, lp->sv, lp->r, lp->w, lp->e, lp->y);
    PairSet ps;
    ps.q = 2;
    ps.a[0] = lp->e;
    ps.b[0] = lp->t;
    ps.a[1] = lp->y;
    ps.b[1] = lp->u;
    backward_from_pairs(lp, st, g, ps);
}

// Parameter gradient from dLoss/dL = sum_q a_q b_q^T (forward state of the
// last forward() call).
void backward_from_pairs(Learned* lp, hipStream_t st, const LearnedGraph& g, const PairSet& ps) {
    const int n = g.n_nodes, W = g.n_edge;
    const int nb = blocks_for(n, 256);
    hipLaunchKernelGGL(k_grad_factor, dim3(nb), dim3(256), 0, st, n, W, ps, lp->O, lp->Ci, lp->Ahat, g.enb, lp->dO,
                       lp->DOE, lp->Gea);
    hipLaunchKernelGGL(k_grad_to, dim3(nb), dim3(256), 0, st, n, W, g.enb, g.erev, lp->DOE, lp->dO);
    // edge map Wa: column sums of Gea
    const int nsb = 64;
    hipLaunchKernelGGL(k_small_wgrad, dim3(nsb), dim3(256), 0, st, n, 81, 0, lp->Gea, (const void*)nullptr, 0,
                       lp->wpart);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(128), 0, st, nsb, 81, lp->wpart, lp->g + kOffWa);
    // head
    hipLaunchKernelGGL(k_small_wgrad, dim3(nsb), dim3(256), 0, st, n, kHead, kHid, lp->dO,
                       (const void*)lp->H[kLayers], 1, lp->wpart);
    hipLaunchKernelGGL(k_reduce_partials, dim3(blocks_for(kHead * kHid + kHead, 256)), dim3(256), 0, st, nsb,
                       kHead * kHid + kHead, lp->wpart, lp->g + kOffWh);
    hipLaunchKernelGGL(k_head_back, dim3(blocks_for(n, 4)), dim3(256), 0, st, n, lp->dO, lp->wbf + kOffWh, lp->dH);
    // message-passing layers
    const int gb = blocks_for(n, 4 * kNodesPerWave);
    const int nwg = blocks_for(n, kWgChunk);
    for (int l = kLayers - 1; l >= 0; --l) {
        hipLaunchKernelGGL(k_relu_back, dim3(blocks_for((int64_t)n * kHid, 256)), dim3(256), 0, st,
                           (int64_t)n * kHid, lp->dH, lp->H[l + 1], lp->dz);
        hipLaunchKernelGGL(k_wgrad_mfma, dim3(nwg), dim3(256), 0, st, n, lp->dz, lp->H[l], lp->A[l], lp->wpart);
        const int off = kOffL1 + l * kLayerSize;
        hipLaunchKernelGGL(k_reduce_partials, dim3(blocks_for(kLayerSize, 256)), dim3(256), 0, st, nwg, kLayerSize,
                           lp->wpart, lp->g + off);
        hipLaunchKernelGGL((k_node_gemm<kHid, kHid, 2 * kHid, 1>), dim3(gb), dim3(256), 0, st, n, lp->dz, lp->dz,
                           lp->wcatT + (int64_t)l * 2 * kHid * kHid, (const float*)nullptr, (void*)lp->dX);
        hipLaunchKernelGGL(k_agg_back, dim3(blocks_for(n, 4)), dim3(256), 0, st, n, g.rp, g.bc, lp->dX, lp->dH);
    }
    // layer 0
    hipLaunchKernelGGL(k_relu_back_f32, dim3(blocks_for((int64_t)n * kHid, 256)), dim3(256), 0, st,
                       (int64_t)n * kHid, lp->dH, lp->H[0], lp->dz0);
    hipLaunchKernelGGL(k_small_wgrad, dim3(nsb), dim3(256), 0, st, n, kHid, kFeat, lp->dz0, (const void*)lp->X, 0,
                       lp->wpart);
    hipLaunchKernelGGL(k_reduce_partials, dim3(blocks_for(kHid * kFeat + kHid, 256)), dim3(256), 0, st, nsb,
                       kHid * kFeat + kHid, lp->wpart, lp->g + kOffW0);
}

void adam_step(Learned* lp, hipStream_t st, float lr) {
    lp->step += 1;
    const float b1t = (float)std::pow(0.9, (double)lp->step), b2t = (float)std::pow(0.999, (double)lp->step);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(kNumParams, 256)), dim3(256), 0, st, kNumParams, lp->p, lp->g, lp->m,
                       lp->v, lr, b1t, b2t);
    hipLaunchKernelGGL(k_refresh_bf16, dim3(blocks_for(kNumParams, 256)), dim3(256), 0, st, lp->p, lp->wbf,
                       lp->wcatT, lp->whp);
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int check_graph(Learned* lp, const LearnedGraph& g) {
    if (g.n_nodes <= 0 || g.n_edge < 0 || g.n_edge > kMaxEdges || !g.rp || !g.bc || !g.n2e_ptr || !g.n2e || !g.E ||
        !g.clamped || (g.n_edge > 0 && (!g.enb || !g.erev || !g.eslot)) || !g.K32 || !g.kdiag || !g.op.sptr) {
        lp->err = "learned preconditioner: incomplete graph description";
        return ASEND_ERR_ARG;
    }
    return ASEND_OK;
}

}  // namespace

int learned_create(Learned** out, uint64_t seed) {
    Learned* lp = new Learned();
    *out = lp;
    LCHK(hipMalloc((void**)&lp->p, kNumParams * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->g, kNumParams * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->m, kNumParams * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->v, kNumParams * sizeof(float)));
    LCHK(hipMalloc((void**)&lp->wbf, kNumParams * sizeof(bf16)));
    LCHK(hipMalloc((void**)&lp->wcatT, kLayers * 2 * kHid * kHid * sizeof(bf16)));
    LCHK(hipMalloc((void**)&lp->whp, kHeadPad * kHid * sizeof(bf16)));
    LCHK(hipMalloc((void**)&lp->partials, 3 * kVecBlocks * sizeof(double)));
    LCHK(hipMalloc((void**)&lp->sv, 8 * sizeof(double)));
    // deterministic uniform (Glorot) init of the message-passing layers; the
    // head and the edge map start at zero, so the untrained factor is exactly
    // block Jacobi (L = C^-T)
    std::vector<float> h(kNumParams, 0.f);
    uint64_t s = seed;
    auto uni = [&](float a) { return a * ((float)(splitmix(s) >> 11) * (2.0f / 9007199254740992.0f) - 1.0f); };
    for (int i = 0; i < kHid * kFeat; ++i) h[kOffW0 + i] = uni(std::sqrt(6.0f / (kFeat + kHid)));
    for (int l = 0; l < kLayers; ++l)
        for (int i = 0; i < kHid * 2 * kHid; ++i) h[kOffL1 + l * kLayerSize + i] = uni(std::sqrt(6.0f / (3 * kHid)));
    LCHK(hipMemcpy(lp->p, h.data(), kNumParams * sizeof(float), hipMemcpyHostToDevice));
    LCHK(hipMemset(lp->m, 0, kNumParams * sizeof(float)));
    LCHK(hipMemset(lp->v, 0, kNumParams * sizeof(float)));
    LCHK(hipMemset(lp->g, 0, kNumParams * sizeof(float)));
    hipLaunchKernelGGL(k_refresh_bf16, dim3(blocks_for(kNumParams, 256)), dim3(256), 0, 0, lp->p, lp->wbf, lp->wcatT,
                       lp->whp);
    LCHK(hipGetLastError());
    LCHK(hipDeviceSynchronize());
    return ASEND_OK;
}

void learned_destroy(Learned* lp) {
    if (!lp) return;
    free_work(lp);
    lfree(lp->p);
    lfree(lp->g);
    lfree(lp->m);
    lfree(lp->v);
    lfree(lp->wbf);
    lfree(lp->wcatT);
    lfree(lp->whp);
    lfree(lp->partials);
    lfree(lp->sv);
    delete lp;
}

const char* learned_error(const Learned* lp) { return lp ? lp->err.c_str() : ""; }

int learned_get_params(Learned* lp, float* host, int n) {
    if (n != kNumParams) return ASEND_ERR_ARG;
    LCHK(hipMemcpy(host, lp->p, kNumParams * sizeof(float), hipMemcpyDeviceToHost));
    return ASEND_OK;
}

int learned_set_params(Learned* lp, const float* host, int n) {
    if (n != kNumParams) return ASEND_ERR_ARG;
    LCHK(hipMemcpy(lp->p, host, kNumParams * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_refresh_bf16, dim3(blocks_for(kNumParams, 256)), dim3(256), 0, 0, lp->p, lp->wbf, lp->wcatT,
                       lp->whp);
    LCHK(hipGetLastError());
    LCHK(hipDeviceSynchronize());
    return ASEND_OK;
}

int learned_infer(Learned* lp, hipStream_t st, const LearnedGraph& g, Factor* F) {
    int rc = check_graph(lp, g);
    if (rc || (rc = ensure_cap(lp, g.n_nodes, g.n_edge))) return rc;
    forward(lp, st, g);
    LCHK(hipGetLastError());
    *F = factor_of(lp, g);
    return ASEND_OK;
}

void learned_sample(hipStream_t st, const LearnedGraph& g, uint64_t seed, float* d_e) {
    hipLaunchKernelGGL(k_randn, dim3(kVecBlocks), dim3(256), 0, st, g.n_nodes, seed, g.clamped, d_e);
}

int learned_train_step(Learned* lp, hipStream_t st, const LearnedGraph& g, const float* d_e, uint64_t sample_seed,
                       float lr, double* loss_out) {
    int rc = check_graph(lp, g);
    if (rc || (rc = ensure_cap(lp, g.n_nodes, g.n_edge))) return rc;
    forward(lp, st, g);
    if (d_e) LCHK(hipMemcpyAsync(lp->e, d_e, 3 * (size_t)g.n_nodes * sizeof(float), hipMemcpyDeviceToDevice, st));
    else learned_sample(st, g, sample_seed, lp->e);
    loss_and_backward(lp, st, g);
    adam_step(lp, st, lr);
    LCHK(hipGetLastError());
    if (loss_out) {
        double h[8];
        LCHK(hipMemcpyAsync(h, lp->sv, 8 * sizeof(double), hipMemcpyDeviceToHost, st));
        LCHK(hipStreamSynchronize(st));
        *loss_out = h[3];
    }
    return ASEND_OK;
}

int learned_grad_from_uses(Learned* lp, hipStream_t st, const LearnedGraph& g, const FactorUses& uses,
                           float* h_grad) {
    if (uses.m < 1 || uses.m > kMaxUses) {
        lp->err = "learned preconditioner: bad number of factor uses";
        return ASEND_ERR_ARG;
    }
    const int n = g.n_nodes, W = g.n_edge;
    const int nb = blocks_for(n, 256);
    // dLoss/dP = sum_k zbar_k v_k^T with P = L L^T:
    //   dLoss/dL = sum_k zbar_k (L^T v_k)^T + v_k (L^T zbar_k)^T
    PairSet ps;
    ps.q = 2 * uses.m;
    for (int k = 0; k < uses.m; ++k) {
        float* tv = lp->pairbuf + (int64_t)(2 * k) * 3 * n;
        float* tz = lp->pairbuf + (int64_t)(2 * k + 1) * 3 * n;
        hipLaunchKernelGGL((k_factor_l<false>), dim3(nb), dim3(256), 0, st, n, W, lp->LsT, g.enb, uses.v[k],
                           (const float*)nullptr, tv);
        hipLaunchKernelGGL((k_factor_l<false>), dim3(nb), dim3(256), 0, st, n, W, lp->LsT, g.enb, uses.zbar[k],
                           (const float*)nullptr, tz);
        ps.a[2 * k] = uses.zbar[k];
        ps.b[2 * k] = tv;
        ps.a[2 * k + 1] = uses.v[k];
        ps.b[2 * k + 1] = tz;
    }
    backward_from_pairs(lp, st, g, ps);
    LCHK(hipGetLastError());
    if (h_grad) {
        LCHK(hipStreamSynchronize(st));
        LCHK(hipMemcpy(h_grad, lp->g, kNumParams * sizeof(float), hipMemcpyDeviceToHost));
    }
    return ASEND_OK;
}

int learned_step_from_uses(Learned* lp, hipStream_t st, const LearnedGraph& g, const FactorUses& uses, float lr) {
    int rc = learned_grad_from_uses(lp, st, g, uses, nullptr);
    if (rc) return rc;
    adam_step(lp, st, lr);
    LCHK(hipGetLastError());
    return ASEND_OK;
}

int learned_eval(Learned* lp, hipStream_t st, const LearnedGraph& g, const float* d_e, double* loss, float* h_grad,
                 float* h_feat, float* h_out) {
    int rc = check_graph(lp, g);
    if (rc || (rc = ensure_cap(lp, g.n_nodes, g.n_edge))) return rc;
    forward(lp, st, g);
    LCHK(hipMemcpyAsync(lp->e, d_e, 3 * (size_t)g.n_nodes * sizeof(float), hipMemcpyDeviceToDevice, st));
    loss_and_backward(lp, st, g);
    LCHK(hipGetLastError());
    LCHK(hipStreamSynchronize(st));
    double h[8];
    LCHK(hipMemcpy(h, lp->sv, 8 * sizeof(double), hipMemcpyDeviceToHost));
    if (loss) *loss = h[3];
    if (h_grad) LCHK(hipMemcpy(h_grad, lp->g, kNumParams * sizeof(float), hipMemcpyDeviceToHost));
    if (h_feat) LCHK(hipMemcpy(h_feat, lp->X, (size_t)g.n_nodes * kFeat * sizeof(float), hipMemcpyDeviceToHost));
    if (h_out) LCHK(hipMemcpy(h_out, lp->O, (size_t)g.n_nodes * kHead * sizeof(float), hipMemcpyDeviceToHost));
    return ASEND_OK;
}

}  // namespace asend
