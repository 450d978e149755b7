// Batched multi-right-hand-side PCG over many design variants.
//
// The operators of all variants (each assembled by its own context) are laid
// side by side as one block-diagonal sliced-ELL operator; every variant owns a
// 64-row-aligned range of block rows.  Up to kNrhs right-hand sides per variant
// advance together: one SpMM launch reads each 3x3 block once and applies it
// to all of them, and every reduction, coefficient and update is per
// (variant, rhs) system, so the whole batch costs the same number of launches
// per iteration as a single solve.  Inner solves are Jacobi PCG in FP32 inside
// FP64 iterative refinement (FP64 residual and update).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "asend_gpu.h"
#include "batched.hpp"
#include "device_common.hpp"
#include "kernels.hpp"

namespace asend {

namespace {

// block rows per reduction chunk (never spans two variants); 512 rows keep
// about 4 workgroups per CU busy on a batch of 8 variants of ~100k rows
constexpr int kChunkRows = 512;

// Y[r][d][j0 + j] = (F -) sum_k K_rk X[c_k][.][j0 + j] for j < NR; vectors are
// [row][3][kNrhs].  LPR wavefronts per 64-row slice: a wavefront holds
// 64 / LPR rows, each row's NR right-hand sides split over LPR lanes (NR / LPR
// each).  The LPR lanes of a row load the same K block addresses in one
// instruction, so every block is fetched once per slice; splitting the right-
// hand sides cuts the accumulators and X registers (NR = 16: 140 VGPRs at
// LPR = 1, 78 at 2, 53 at 4), so more wavefronts stay resident per SIMD.
// Slices of variants whose systems have all converged are skipped when
// `active` is given.
template <typename T, int NR, int LPR, bool SUB>
__global__ __launch_bounds__(kBlock) void k_sell3_mrhs(int n_slices, const int32_t* __restrict__ sptr,
                                                       const int32_t* __restrict__ scol, const T* __restrict__ val,
                                                       const T* __restrict__ X, T* __restrict__ Y,
                                                       const T* __restrict__ F, int j0,
                                                       const int32_t* __restrict__ slice_var,
                                                       const uint8_t* __restrict__ active) {
    constexpr int RPW = kWave / LPR;  // rows per wavefront
    constexpr int NRL = NR / LPR;     // right-hand sides per lane
    const int gw = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    const int slice = gw / LPR;
    const int lane = threadIdx.x % kWave;
    const int rl = (gw % LPR) * RPW + lane % RPW;  // row within the slice
    const int jq = j0 + (lane / RPW) * NRL;        // first right-hand side of the lane
    if (slice >= n_slices) return;
    if (active && !active[slice_var[slice]]) return;
    const int64_t r = (int64_t)slice * kWave + rl;
    T acc[3][NRL];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int j = 0; j < NRL; ++j) acc[d][j] = 0;
    const int k0 = sptr[slice], k1 = sptr[slice + 1];
    const int32_t* cp = scol + (int64_t)k0 * kWave + rl;
    const T* vp = val + (int64_t)k0 * 9 * kWave + rl;
    for (int k = k0; k < k1; ++k) {
        // values and columns streamed once: non-temporal, keeping X in L2
        const int64_t c = __builtin_nontemporal_load(cp);
        T a[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) a[q] = __builtin_nontemporal_load(vp + q * kWave);
        const T* xc = X + c * 3 * kNrhs + jq;
#pragma unroll
        for (int j = 0; j < NRL; ++j) {
            const T x0 = xc[j], x1 = xc[kNrhs + j], x2 = xc[2 * kNrhs + j];
            acc[0][j] += a[0] * x0 + a[1] * x1 + a[2] * x2;
            acc[1][j] += a[3] * x0 + a[4] * x1 + a[5] * x2;
            acc[2][j] += a[6] * x0 + a[7] * x1 + a[8] * x2;
        }
        cp += kWave;
        vp += 9 * kWave;
    }
    T* yr = Y + r * 3 * kNrhs + jq;
    const T* fr = F ? F + r * 3 * kNrhs + jq : nullptr;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int j = 0; j < NRL; ++j) yr[d * kNrhs + j] = SUB ? fr[d * kNrhs + j] - acc[d][j] : acc[d][j];
}

// Per-system scalars, index s = variant * kNrhs + j
enum { B_RZ = 0, B_PQ, B_ALPHA, B_BETA, B_RR, B_RR0, B_IT, B_DONE, B_MAXIT, B_TOL2, B_N };

// Chunk reduction of the FP64 residual: each block owns rows [row0, row1) of
// one variant and accumulates, per rhs j, |r|^2 into partials[chunk][0][j].
__global__ __launch_bounds__(256) void k_chunk_rr64(const int32_t* __restrict__ chunk_row,
                                                    const double* __restrict__ a, double* __restrict__ partials) {
    __shared__ double sh[256];
    const int ch = blockIdx.x;
    const int j = threadIdx.x % kNrhs;
    const int64_t e0 = (int64_t)chunk_row[ch] * 3, e1 = (int64_t)chunk_row[ch + 1] * 3;
    double t = 0.0;
#pragma unroll 4
    for (int64_t e = e0 + threadIdx.x / kNrhs; e < e1; e += 256 / kNrhs) {
        const int64_t i = e * kNrhs + j;
        t += a[i] * a[i];
    }
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int s = 128; s >= kNrhs; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < kNrhs) partials[(int64_t)ch * 2 * kNrhs + threadIdx.x] = sh[threadIdx.x];
}

// partials of p.q and p.r per system (the flexible step length p.r / p.q)
__global__ __launch_bounds__(256) void k_chunk_pqr(const int32_t* __restrict__ chunk_row,
                                                   const int32_t* __restrict__ chunk_seg, const float* __restrict__ P,
                                                   const float* __restrict__ Q, const float* __restrict__ R,
                                                   const double* __restrict__ sv, double* __restrict__ partials) {
    __shared__ double sh[2][256];
    const int ch = blockIdx.x;
    const int j = threadIdx.x % kNrhs;
    const int64_t e0 = (int64_t)chunk_row[ch] * 3, e1 = (int64_t)chunk_row[ch + 1] * 3;
    const int s = chunk_seg[ch] * kNrhs + j;
    double t0 = 0.0, t1 = 0.0;
    if (sv[s * B_N + B_DONE] == 0.0) {
#pragma unroll 4
        for (int64_t e = e0 + threadIdx.x / kNrhs; e < e1; e += 256 / kNrhs) {
            const int64_t i = e * kNrhs + j;
            const double pi = P[i];
            t0 += pi * Q[i];
            t1 += pi * R[i];
        }
    }
    sh[0][threadIdx.x] = t0;
    sh[1][threadIdx.x] = t1;
    __syncthreads();
    for (int k = 128; k >= kNrhs; k >>= 1) {
        if (threadIdx.x < k) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + k];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x < kNrhs) {
        partials[(int64_t)ch * 2 * kNrhs + threadIdx.x] = sh[0][threadIdx.x];
        partials[(int64_t)ch * 2 * kNrhs + kNrhs + threadIdx.x] = sh[1][threadIdx.x];
    }
}

// x += alpha p ; r -= alpha q ; partials of (Dinv r).q and r.r  (FP32 inner)
__global__ __launch_bounds__(256) void k_chunk_update(const int32_t* __restrict__ chunk_row,
                                                      const int32_t* __restrict__ chunk_seg,
                                                      const float* __restrict__ dinv, float* __restrict__ X,
                                                      float* __restrict__ R, const float* __restrict__ P,
                                                      const float* __restrict__ Q, const double* __restrict__ sv,
                                                      double* __restrict__ partials) {
    __shared__ double sh[2][256];
    const int ch = blockIdx.x;
    const int j = threadIdx.x % kNrhs;
    const int64_t e0 = (int64_t)chunk_row[ch] * 3, e1 = (int64_t)chunk_row[ch + 1] * 3;
    const int s = chunk_seg[ch] * kNrhs + j;
    double t0 = 0.0, t1 = 0.0;
    if (sv[s * B_N + B_DONE] == 0.0) {
        const float alpha = (float)sv[s * B_N + B_ALPHA];
#pragma unroll 4
        for (int64_t e = e0 + threadIdx.x / kNrhs; e < e1; e += 256 / kNrhs) {
            const int64_t i = e * kNrhs + j;
            X[i] += alpha * P[i];
            const float ri = R[i] - alpha * Q[i];
            R[i] = ri;
            t0 += (double)(dinv[e] * ri) * Q[i];
            t1 += (double)ri * ri;
        }
    }
    sh[0][threadIdx.x] = t0;
    sh[1][threadIdx.x] = t1;
    __syncthreads();
    for (int k = 128; k >= kNrhs; k >>= 1) {
        if (threadIdx.x < k) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + k];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x < kNrhs) {
        partials[(int64_t)ch * 2 * kNrhs + threadIdx.x] = sh[0][threadIdx.x];
        partials[(int64_t)ch * 2 * kNrhs + kNrhs + threadIdx.x] = sh[1][threadIdx.x];
    }
}

// r = b, x = 0, p = Dinv r; partials of r.z and r.r
__global__ __launch_bounds__(256) void k_chunk_init(const int32_t* __restrict__ chunk_row,
                                                    const float* __restrict__ dinv, const float* __restrict__ B,
                                                    float* __restrict__ X, float* __restrict__ R,
                                                    float* __restrict__ P, double* __restrict__ partials) {
    __shared__ double sh[2][256];
    const int ch = blockIdx.x;
    const int j = threadIdx.x % kNrhs;
    const int64_t e0 = (int64_t)chunk_row[ch] * 3, e1 = (int64_t)chunk_row[ch + 1] * 3;
    double t0 = 0.0, t1 = 0.0;
#pragma unroll 4
    for (int64_t e = e0 + threadIdx.x / kNrhs; e < e1; e += 256 / kNrhs) {
        const int64_t i = e * kNrhs + j;
        const float bi = B[i], zi = dinv[e] * bi;
        X[i] = 0.f;
        R[i] = bi;
        P[i] = zi;
        t0 += (double)bi * zi;
        t1 += (double)bi * bi;
    }
    sh[0][threadIdx.x] = t0;
    sh[1][threadIdx.x] = t1;
    __syncthreads();
    for (int k = 128; k >= kNrhs; k >>= 1) {
        if (threadIdx.x < k) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + k];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x < kNrhs) {
        partials[(int64_t)ch * 2 * kNrhs + threadIdx.x] = sh[0][threadIdx.x];
        partials[(int64_t)ch * 2 * kNrhs + kNrhs + threadIdx.x] = sh[1][threadIdx.x];
    }
}

// partials of r.(Dinv r) and r.r (after a reliable update reloaded r)
__global__ __launch_bounds__(256) void k_chunk_rz(const int32_t* __restrict__ chunk_row,
                                                  const float* __restrict__ dinv, const float* __restrict__ R,
                                                  double* __restrict__ partials) {
    __shared__ double sh[2][256];
    const int ch = blockIdx.x;
    const int j = threadIdx.x % kNrhs;
    const int64_t e0 = (int64_t)chunk_row[ch] * 3, e1 = (int64_t)chunk_row[ch + 1] * 3;
    double t0 = 0.0, t1 = 0.0;
#pragma unroll 4
    for (int64_t e = e0 + threadIdx.x / kNrhs; e < e1; e += 256 / kNrhs) {
        const float ri = R[e * kNrhs + j];
        t0 += (double)ri * (dinv[e] * ri);
        t1 += (double)ri * ri;
    }
    sh[0][threadIdx.x] = t0;
    sh[1][threadIdx.x] = t1;
    __syncthreads();
    for (int k = 128; k >= kNrhs; k >>= 1) {
        if (threadIdx.x < k) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + k];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x < kNrhs) {
        partials[(int64_t)ch * 2 * kNrhs + threadIdx.x] = sh[0][threadIdx.x];
        partials[(int64_t)ch * 2 * kNrhs + kNrhs + threadIdx.x] = sh[1][threadIdx.x];
    }
}

// p = Dinv r + beta p
__global__ __launch_bounds__(256) void k_chunk_dir(const int32_t* __restrict__ chunk_row,
                                                   const int32_t* __restrict__ chunk_seg,
                                                   const float* __restrict__ dinv, const float* __restrict__ R,
                                                   float* __restrict__ P, const double* __restrict__ sv) {
    const int ch = blockIdx.x;
    const int j = threadIdx.x % kNrhs;
    const int s = chunk_seg[ch] * kNrhs + j;
    if (sv[s * B_N + B_DONE] != 0.0) return;
    const float beta = (float)sv[s * B_N + B_BETA];
    const int64_t e0 = (int64_t)chunk_row[ch] * 3, e1 = (int64_t)chunk_row[ch + 1] * 3;
#pragma unroll 4
    for (int64_t e = e0 + threadIdx.x / kNrhs; e < e1; e += 256 / kNrhs) {
        const int64_t i = e * kNrhs + j;
        P[i] = dinv[e] * R[i] + beta * P[i];
    }
}

// Per-system finalize over the chunks of its variant: one workgroup per
// variant; the 256 / kNrhs lanes of each right-hand side add a strided share
// of the chunks, then a fixed-order tree over those lanes (deterministic).
// The inner solves are flexible CG (one truncation vector): alpha = p.r / p.q
// and beta = -z.q / p.q from fresh inner products, so a residual replaced by
// a reliable update needs no repair of the recurrence.
//   OP 0 (init):   rz, rr0 = rr, it = 0, done when rr0 = 0 or the system is frozen
//   OP 1 (p.q, p.r): pq, alpha = pr / pq
//   OP 2 (update): beta = -zq / pq, rr, it += 1, done on rr <= tol2 rr0 or it >= max
//   OP 3 (|r|^2):  rr (FP64 residual norm) into out[s]
//   OP 4 (reliable update): rz, rr = rr0 of the reloaded residual, tol2 =
//                  out[s], the iteration count kept; done when frozen or r = 0
__global__ __launch_bounds__(256) void k_sys_fin(const int32_t* __restrict__ seg_chunk,
                                                 const double* __restrict__ partials, int op,
                                                 double* __restrict__ sv, const uint8_t* __restrict__ frozen,
                                                 double* __restrict__ out) {
    __shared__ double sh[2][256];
    const int seg = blockIdx.x;
    const int j = threadIdx.x % kNrhs, q = threadIdx.x / kNrhs;
    const int s = seg * kNrhs + j;
    double* v = sv + (int64_t)s * B_N;
    const bool skip = (op == 1 || op == 2) && v[B_DONE] != 0.0;
    double t0 = 0.0, t1 = 0.0;
    if (!skip)
        for (int ch = seg_chunk[seg] + q; ch < seg_chunk[seg + 1]; ch += 256 / kNrhs) {
            t0 += partials[(int64_t)ch * 2 * kNrhs + j];
            t1 += partials[(int64_t)ch * 2 * kNrhs + kNrhs + j];
        }
    sh[0][threadIdx.x] = t0;
    sh[1][threadIdx.x] = t1;
    __syncthreads();
    for (int w = 128; w >= kNrhs; w >>= 1) {
        if (threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (q != 0 || skip) return;
    t0 = sh[0][j];
    t1 = sh[1][j];
    if (op == 0) {
        v[B_RZ] = t0;
        v[B_RR] = t1;
        v[B_RR0] = t1;
        v[B_IT] = 0.0;
        v[B_DONE] = (frozen[s] || t1 == 0.0) ? 1.0 : 0.0;
    } else if (op == 1) {
        v[B_PQ] = t0;
        v[B_ALPHA] = t1 / t0;
    } else if (op == 2) {
        v[B_BETA] = -t0 / v[B_PQ];
        v[B_RR] = t1;
        v[B_IT] += 1.0;
        if (t1 <= v[B_TOL2] * v[B_RR0] || v[B_IT] >= v[B_MAXIT] || !(t1 == t1)) v[B_DONE] = 1.0;
    } else if (op == 4) {
        v[B_RZ] = t0;
        v[B_RR] = t1;
        v[B_RR0] = t1;
        v[B_TOL2] = out[s];
        v[B_DONE] = (frozen[s] || t1 == 0.0 || v[B_IT] >= v[B_MAXIT]) ? 1.0 : 0.0;
    } else {
        out[s] = t0;
    }
}

// active[v] = some system of variant v still iterating
__global__ void k_var_active(int n_var, const double* __restrict__ sv, uint8_t* __restrict__ active) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_var) return;
    uint8_t a = 0;
    for (int j = 0; j < kNrhs; ++j) a |= (sv[(int64_t)(v * kNrhs + j) * B_N + B_DONE] == 0.0) ? 1 : 0;
    active[v] = a;
}

// FP64 residual -> FP32 inner right-hand side, scaled per system by 1/||r||
__global__ void k_mr_to_f32(int64_t n_elem, const double* __restrict__ R, const double* __restrict__ inv_norm,
                            const int32_t* __restrict__ seg_of_elem_row, float* __restrict__ R32) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_elem;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / (3 * kNrhs);
        const int s = seg_of_elem_row[row] * kNrhs + (int)(i % kNrhs);
        R32[i] = (float)(R[i] * inv_norm[s]);
    }
}

// Reliable update of the search direction per system: P *= ratio[s] (the
// same direction in the new normalisation of its residual), or, where
// ratio[s] = 0, a restart P = Dinv R
__global__ void k_mr_scale(int64_t n_elem, const double* __restrict__ ratio, const int32_t* __restrict__ seg_of_elem_row,
                           const float* __restrict__ dinv, const float* __restrict__ R, float* __restrict__ P) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_elem;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / (3 * kNrhs);
        const int s = seg_of_elem_row[row] * kNrhs + (int)(i % kNrhs);
        const double q = ratio[s];
        P[i] = q != 0.0 ? (float)(P[i] * q) : dinv[i / kNrhs] * R[i];
    }
}

// U += X / inv_norm (FP64 update of every system)
__global__ void k_mr_axpy(int64_t n_elem, const float* __restrict__ X, const double* __restrict__ inv_norm,
                          const int32_t* __restrict__ seg_of_elem_row, double* __restrict__ U) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_elem;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / (3 * kNrhs);
        const int s = seg_of_elem_row[row] * kNrhs + (int)(i % kNrhs);
        const double w = inv_norm[s];
        if (w > 0.0) U[i] += (double)X[i] / w;
    }
}

__global__ void k_add_offset(int64_t n, const int32_t* __restrict__ a, int32_t off, int32_t* __restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i] + off;
}

int nbk(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 8192); }

}  // namespace

struct Batch {
    std::string err;
    hipStream_t st = nullptr;
    int n_var = 0;
    std::vector<int64_t> n_rows, row_off;  // block rows of each variant and their offset
    int64_t rows_total = 0;                // padded to whole slices
    int n_slices = 0;
    int32_t *sptr = nullptr, *scol = nullptr;
    float* K32 = nullptr;
    double* K64 = nullptr;
    float* dinv = nullptr;
    // reduction chunks
    int n_chunks = 0;
    int32_t *chunk_row = nullptr, *chunk_seg = nullptr, *seg_chunk = nullptr, *seg_of_row = nullptr;
    int32_t* slice_var = nullptr;  // variant of each slice
    uint8_t* active = nullptr;     // per variant: some system still iterating
    // vectors [rows_total][3][kNrhs]
    double *F = nullptr, *U = nullptr, *R64 = nullptr;
    float *B32 = nullptr, *X = nullptr, *R = nullptr, *P = nullptr, *Q = nullptr;
    double *sv = nullptr, *partials = nullptr, *rr = nullptr, *inv_norm = nullptr, *ratio = nullptr, *tol2 = nullptr;
    uint8_t* frozen = nullptr;
    double* h_sv = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

#define BCHK(expr)                                                       \
    do {                                                                 \
        hipError_t e_ = (expr);                                          \
        if (e_ != hipSuccess) {                                          \
            b->err = std::string(#expr) + ": " + hipGetErrorString(e_);  \
            return ASEND_ERR_HIP;                                        \
        }                                                                \
    } while (0)

template <typename T>
void bfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// FP64 residual R64 = F - K64 U for every system and its squared norm per system
void residual64(Batch* b) {
    const int nblk = (b->n_slices + kBlock / kWave - 1) / (kBlock / kWave);
    for (int j0 = 0; j0 < kNrhs; j0 += 8)
        hipLaunchKernelGGL((k_sell3_mrhs<double, 8, 1, true>), dim3(nblk), dim3(kBlock), 0, b->st, b->n_slices, b->sptr,
                           b->scol, b->K64, b->U, b->R64, b->F, j0, (const int32_t*)nullptr,
                           (const uint8_t*)nullptr);
    hipLaunchKernelGGL(k_chunk_rr64, dim3(b->n_chunks), dim3(256), 0, b->st, b->chunk_row, b->R64, b->partials);
    hipLaunchKernelGGL(k_sys_fin, dim3(b->n_var), dim3(256), 0, b->st,
                       b->seg_chunk, b->partials, 3, b->sv, b->frozen, b->rr);
}

void spmm32(Batch* b, const float* X, float* Y) {
    constexpr int lpr = 4;  // measured on 8 variants: 1 -> 12.7 s, 2 -> 11.4 s, 4 -> 10.8 s, 8 -> 13.5 s
    const int nblk = (lpr * b->n_slices + kBlock / kWave - 1) / (kBlock / kWave);
    hipLaunchKernelGGL((k_sell3_mrhs<float, kNrhs, lpr, false>), dim3(nblk), dim3(kBlock), 0, b->st, b->n_slices, b->sptr,
                       b->scol, b->K32, X, Y, (const float*)nullptr, 0, b->slice_var, b->active);
}

}  // namespace

int batch_create(Batch** out, const std::vector<OperatorView>& ops) {
    Batch* b = new Batch();
    *out = b;
    b->n_var = (int)ops.size();
    if (b->n_var < 1) {
        b->err = "empty batch";
        return ASEND_ERR_ARG;
    }
    BCHK(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
    BCHK(hipEventCreate(&b->ev0));
    BCHK(hipEventCreate(&b->ev1));
    // layout: variant v owns slices [slice_off, slice_off + n_slices_v)
    std::vector<int32_t> slice_off(b->n_var + 1, 0);
    std::vector<int64_t> slot_off(b->n_var + 1, 0);
    for (int v = 0; v < b->n_var; ++v) {
        slice_off[v + 1] = slice_off[v] + ops[v].n_slices;
        slot_off[v + 1] = slot_off[v] + ops[v].n_slots;
        b->n_rows.push_back(ops[v].n_rows);
        b->row_off.push_back((int64_t)kWave * slice_off[v]);
    }
    b->n_slices = slice_off[b->n_var];
    b->rows_total = (int64_t)kWave * b->n_slices;
    const int64_t n_slots = slot_off[b->n_var];
    if (b->rows_total * 3 * kNrhs >= (1ll << 31) * 4 || b->rows_total >= (1ll << 31)) {
        b->err = "batch too large";
        return ASEND_ERR_UNSUPPORTED;
    }
    BCHK(hipMalloc((void**)&b->sptr, (b->n_slices + 1) * sizeof(int32_t)));
    BCHK(hipMalloc((void**)&b->scol, n_slots * kWave * sizeof(int32_t)));
    BCHK(hipMalloc((void**)&b->K32, n_slots * 9 * kWave * sizeof(float)));
    BCHK(hipMalloc((void**)&b->K64, n_slots * 9 * kWave * sizeof(double)));
    BCHK(hipMalloc((void**)&b->dinv, b->rows_total * 3 * sizeof(float)));
    BCHK(hipMemset(b->dinv, 0, b->rows_total * 3 * sizeof(float)));
    std::vector<int32_t> sptr(b->n_slices + 1, 0);
    for (int v = 0; v < b->n_var; ++v) {
        const OperatorView& o = ops[v];
        for (int s = 0; s < o.n_slices; ++s) sptr[slice_off[v] + s] = o.h_sptr[s] + (int32_t)slot_off[v];
        const int64_t nsl = o.n_slots;
        hipLaunchKernelGGL(k_add_offset, dim3(nbk(nsl * kWave)), dim3(256), 0, b->st, nsl * kWave, o.scol,
                           (int32_t)b->row_off[v], b->scol + slot_off[v] * kWave);
        BCHK(hipMemcpyAsync(b->K32 + slot_off[v] * 9 * kWave, o.K32, nsl * 9 * kWave * sizeof(float),
                            hipMemcpyDeviceToDevice, b->st));
        BCHK(hipMemcpyAsync(b->K64 + slot_off[v] * 9 * kWave, o.K64, nsl * 9 * kWave * sizeof(double),
                            hipMemcpyDeviceToDevice, b->st));
        BCHK(hipMemcpyAsync(b->dinv + b->row_off[v] * 3, o.dinv32, o.n_rows * 3 * sizeof(float),
                            hipMemcpyDeviceToDevice, b->st));
    }
    sptr[b->n_slices] = (int32_t)n_slots;
    {
        std::vector<int32_t> sv(b->n_slices);
        for (int v = 0; v < b->n_var; ++v)
            for (int s = slice_off[v]; s < slice_off[v + 1]; ++s) sv[s] = v;
        BCHK(hipMalloc((void**)&b->slice_var, sv.size() * sizeof(int32_t)));
        BCHK(hipMalloc((void**)&b->active, b->n_var));
        BCHK(hipMemcpyAsync(b->slice_var, sv.data(), sv.size() * 4, hipMemcpyHostToDevice, b->st));
        BCHK(hipStreamSynchronize(b->st));
    }
    BCHK(hipMemcpyAsync(b->sptr, sptr.data(), sptr.size() * sizeof(int32_t), hipMemcpyHostToDevice, b->st));
    // reduction chunks never span two variants
    std::vector<int32_t> crow, cseg, segc(b->n_var + 1, 0), seg_of_row(b->rows_total, 0);
    for (int v = 0; v < b->n_var; ++v) {
        const int64_t r0 = b->row_off[v], r1 = (v + 1 < b->n_var) ? b->row_off[v + 1] : b->rows_total;
        segc[v] = (int32_t)cseg.size();
        for (int64_t r = r0; r < r1; r += kChunkRows) {
            crow.push_back((int32_t)r);
            cseg.push_back(v);
        }
        for (int64_t r = r0; r < r1; ++r) seg_of_row[r] = v;
    }
    segc[b->n_var] = (int32_t)cseg.size();
    crow.push_back((int32_t)b->rows_total);
    b->n_chunks = (int)cseg.size();
    BCHK(hipMalloc((void**)&b->chunk_row, crow.size() * sizeof(int32_t)));
    BCHK(hipMalloc((void**)&b->chunk_seg, cseg.size() * sizeof(int32_t)));
    BCHK(hipMalloc((void**)&b->seg_chunk, segc.size() * sizeof(int32_t)));
    BCHK(hipMalloc((void**)&b->seg_of_row, seg_of_row.size() * sizeof(int32_t)));
    BCHK(hipMemcpyAsync(b->chunk_row, crow.data(), crow.size() * 4, hipMemcpyHostToDevice, b->st));
    BCHK(hipMemcpyAsync(b->chunk_seg, cseg.data(), cseg.size() * 4, hipMemcpyHostToDevice, b->st));
    BCHK(hipMemcpyAsync(b->seg_chunk, segc.data(), segc.size() * 4, hipMemcpyHostToDevice, b->st));
    BCHK(hipMemcpyAsync(b->seg_of_row, seg_of_row.data(), seg_of_row.size() * 4, hipMemcpyHostToDevice, b->st));
    const int64_t ne = b->rows_total * 3 * kNrhs;
    for (double** q : {&b->F, &b->U, &b->R64}) BCHK(hipMalloc((void**)q, ne * sizeof(double)));
    for (float** q : {&b->B32, &b->X, &b->R, &b->P, &b->Q}) BCHK(hipMalloc((void**)q, ne * sizeof(float)));
    const int n_sys = b->n_var * kNrhs;
    BCHK(hipMalloc((void**)&b->sv, (size_t)n_sys * B_N * sizeof(double)));
    BCHK(hipMalloc((void**)&b->partials, (size_t)b->n_chunks * 2 * kNrhs * sizeof(double)));
    BCHK(hipMalloc((void**)&b->rr, n_sys * sizeof(double)));
    BCHK(hipMalloc((void**)&b->inv_norm, n_sys * sizeof(double)));
    BCHK(hipMalloc((void**)&b->ratio, n_sys * sizeof(double)));
    BCHK(hipMalloc((void**)&b->tol2, n_sys * sizeof(double)));
    BCHK(hipMalloc((void**)&b->frozen, n_sys));
    BCHK(hipHostMalloc((void**)&b->h_sv, (size_t)n_sys * B_N * sizeof(double), hipHostMallocDefault));
    BCHK(hipStreamSynchronize(b->st));
    return ASEND_OK;
}

void batch_destroy(Batch* b) {
    if (!b) return;
    if (b->st) (void)hipStreamSynchronize(b->st);
    bfree(b->sptr);
    bfree(b->scol);
    bfree(b->K32);
    bfree(b->K64);
    bfree(b->dinv);
    bfree(b->chunk_row);
    bfree(b->chunk_seg);
    bfree(b->seg_chunk);
    bfree(b->seg_of_row);
    bfree(b->slice_var);
    bfree(b->active);
    for (double** q : {&b->F, &b->U, &b->R64, &b->sv, &b->partials, &b->rr, &b->inv_norm, &b->ratio, &b->tol2})
        bfree(*q);
    for (float** q : {&b->B32, &b->X, &b->R, &b->P, &b->Q}) bfree(*q);
    bfree(b->frozen);
    if (b->h_sv) (void)hipHostFree(b->h_sv);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

const char* batch_error(const Batch* b) { return b ? b->err.c_str() : ""; }

int64_t batch_rows_total(const Batch* b) { return b->rows_total; }
int64_t batch_row_offset(const Batch* b, int v) { return b->row_off[v]; }

int batch_solve(Batch* b, const double* h_F, double* h_U, int nrhs, double tol, double inner_tol, int64_t max_it,
                int check_every, int verbose, BatchStats* stats) {
    if (nrhs < 1 || nrhs > kNrhs || !(tol > 0) || !(inner_tol > 0) || check_every < 1) {
        b->err = "bad batch solve arguments";
        return ASEND_ERR_ARG;
    }
    const int n_sys = b->n_var * kNrhs;
    const int64_t ne = b->rows_total * 3 * kNrhs;
    BCHK(hipMemcpyAsync(b->F, h_F, ne * sizeof(double), hipMemcpyHostToDevice, b->st));
    BCHK(hipMemsetAsync(b->U, 0, ne * sizeof(double), b->st));
    BCHK(hipEventRecord(b->ev0, b->st));
    // ||f|| per system
    residual64(b);
    std::vector<double> ff(n_sys), rr(n_sys), invn(n_sys);
    BCHK(hipMemcpyAsync(ff.data(), b->rr, n_sys * sizeof(double), hipMemcpyDeviceToHost, b->st));
    BCHK(hipStreamSynchronize(b->st));
    rr = ff;
    // FP32 lockstep block-PCG with FP64 reliable updates.  Every system runs
    // one PCG whose residual is normalised at its last refresh; a refresh
    // (U += |r| X in FP64, the FP64 residual of every system, the converged
    // systems frozen, R = R64 / |R64|, P rescaled by |r_old| / |r_new|, rz and
    // |r|^2 recomputed) keeps every search direction.  It comes once every
    // active system has cut its recursive residual by inner_tol or claims
    // convergence; the FP64 residual alone decides convergence.
    std::vector<uint8_t> frozen(n_sys);
    std::vector<double> init((size_t)n_sys * B_N, 0.0), ratio(n_sys, 0.0), tol2(n_sys, 0.0), rr_prev(n_sys, 0.0);
    int64_t total_inner = 0, outer = 0;
    std::vector<int64_t> sys_it(n_sys, 0);
    const double ru2 = inner_tol * inner_tol;
    // next refresh targets and the refresh itself (host part)
    auto classify = [&]() {
        int active = 0;
        for (int s = 0; s < n_sys; ++s) {
            const bool conv = (s % kNrhs) >= nrhs || ff[s] == 0.0 || std::sqrt(rr[s] / ff[s]) <= tol;
            const double inv_old = invn[s];
            frozen[s] = conv ? 1 : 0;
            invn[s] = conv ? 0.0 : 1.0 / std::sqrt(rr[s]);
            // |r_old| / |r_new|: the factor that keeps P's direction in the new
            // normalisation; 0 (restart the direction) where the FP32
            // recursion drifted, i.e. the FP64 residual is more than twice the
            // recursive one: the old directions no longer fit the residual
            ratio[s] = (conv || inv_old == 0.0) ? 0.0 : 1.0 / (inv_old * std::sqrt(rr[s]));
            if (ratio[s] != 0.0) {
                const double rec = b->h_sv[(size_t)s * B_N + B_RR] * rr_prev[s];
                if (!(rec > 0.0) || rr[s] > 4.0 * rec) ratio[s] = 0.0;
            }
            // recursive convergence target of the renormalised system
            tol2[s] = conv ? 0.0 : tol * tol * ff[s] / rr[s];
            active += conv ? 0 : 1;
        }
        return active;
    };
    std::fill(invn.begin(), invn.end(), 0.0);
    int active = classify();
    for (int s = 0; s < n_sys; ++s) {
        init[(size_t)s * B_N + B_TOL2] = tol2[s];
        init[(size_t)s * B_N + B_MAXIT] = (double)max_it;
    }
    if (active) {
        BCHK(hipMemcpyAsync(b->frozen, frozen.data(), n_sys, hipMemcpyHostToDevice, b->st));
        BCHK(hipMemcpyAsync(b->inv_norm, invn.data(), n_sys * sizeof(double), hipMemcpyHostToDevice, b->st));
        std::memcpy(b->h_sv, init.data(), init.size() * sizeof(double));
        BCHK(hipMemcpyAsync(b->sv, b->h_sv, init.size() * sizeof(double), hipMemcpyHostToDevice, b->st));
        hipLaunchKernelGGL(k_mr_to_f32, dim3(nbk(ne)), dim3(256), 0, b->st, ne, b->R64, b->inv_norm, b->seg_of_row,
                           b->B32);
        hipLaunchKernelGGL(k_chunk_init, dim3(b->n_chunks), dim3(256), 0, b->st, b->chunk_row, b->dinv, b->B32, b->X,
                           b->R, b->P, b->partials);
        hipLaunchKernelGGL(k_sys_fin, dim3(b->n_var), dim3(256), 0, b->st, b->seg_chunk, b->partials, 0, b->sv,
                           b->frozen, (double*)nullptr);
        hipLaunchKernelGGL(k_var_active, dim3((b->n_var + 255) / 256), dim3(256), 0, b->st, b->n_var, b->sv,
                           b->active);
    }
    while (active && outer < 200) {
        for (int k = 0; k < check_every; ++k) {
            spmm32(b, b->P, b->Q);
            hipLaunchKernelGGL(k_chunk_pqr, dim3(b->n_chunks), dim3(256), 0, b->st, b->chunk_row, b->chunk_seg, b->P,
                               b->Q, b->R, b->sv, b->partials);
            hipLaunchKernelGGL(k_sys_fin, dim3(b->n_var), dim3(256), 0, b->st, b->seg_chunk, b->partials, 1, b->sv,
                               b->frozen, (double*)nullptr);
            hipLaunchKernelGGL(k_chunk_update, dim3(b->n_chunks), dim3(256), 0, b->st, b->chunk_row, b->chunk_seg,
                               b->dinv, b->X, b->R, b->P, b->Q, b->sv, b->partials);
            hipLaunchKernelGGL(k_sys_fin, dim3(b->n_var), dim3(256), 0, b->st, b->seg_chunk, b->partials, 2, b->sv,
                               b->frozen, (double*)nullptr);
            hipLaunchKernelGGL(k_var_active, dim3((b->n_var + 255) / 256), dim3(256), 0, b->st, b->n_var, b->sv,
                               b->active);
            hipLaunchKernelGGL(k_chunk_dir, dim3(b->n_chunks), dim3(256), 0, b->st, b->chunk_row, b->chunk_seg,
                               b->dinv, b->R, b->P, b->sv);
        }
        total_inner += check_every;
        BCHK(hipGetLastError());
        BCHK(hipMemcpyAsync(b->h_sv, b->sv, (size_t)n_sys * B_N * sizeof(double), hipMemcpyDeviceToHost, b->st));
        BCHK(hipStreamSynchronize(b->st));
        // refresh once every active system is done or has cut its residual by inner_tol
        bool refresh = true, stalled = true;
        for (int s = 0; s < n_sys && refresh; ++s) {
            if (frozen[s]) continue;
            const double* v = b->h_sv + (size_t)s * B_N;
            const bool done = v[B_DONE] != 0.0;
            refresh = done || v[B_RR] <= ru2 * v[B_RR0];
            stalled = stalled && done && v[B_IT] >= v[B_MAXIT];
        }
        if (!refresh) continue;
        ++outer;
        for (int s = 0; s < n_sys; ++s) sys_it[s] = (int64_t)b->h_sv[(size_t)s * B_N + B_IT];
        hipLaunchKernelGGL(k_mr_axpy, dim3(nbk(ne)), dim3(256), 0, b->st, ne, b->X, b->inv_norm, b->seg_of_row, b->U);
        BCHK(hipMemsetAsync(b->X, 0, ne * sizeof(float), b->st));
        residual64(b);
        rr_prev = rr;
        BCHK(hipMemcpyAsync(rr.data(), b->rr, n_sys * sizeof(double), hipMemcpyDeviceToHost, b->st));
        BCHK(hipStreamSynchronize(b->st));
        active = classify();
        if (verbose) {
            double worst = 0.0;
            for (int s = 0; s < n_sys; ++s)
                if ((s % kNrhs) < nrhs && ff[s] > 0.0) worst = std::max(worst, std::sqrt(rr[s] / ff[s]));
            std::fprintf(stderr, "[batch] refresh %lld: %d active systems, %lld lockstep iterations, worst residual %.2e\n",
                         (long long)outer, active, (long long)total_inner, worst);
        }
        if (!active || stalled) break;
        BCHK(hipMemcpyAsync(b->frozen, frozen.data(), n_sys, hipMemcpyHostToDevice, b->st));
        BCHK(hipMemcpyAsync(b->inv_norm, invn.data(), n_sys * sizeof(double), hipMemcpyHostToDevice, b->st));
        BCHK(hipMemcpyAsync(b->ratio, ratio.data(), n_sys * sizeof(double), hipMemcpyHostToDevice, b->st));
        BCHK(hipMemcpyAsync(b->tol2, tol2.data(), n_sys * sizeof(double), hipMemcpyHostToDevice, b->st));
        hipLaunchKernelGGL(k_mr_to_f32, dim3(nbk(ne)), dim3(256), 0, b->st, ne, b->R64, b->inv_norm, b->seg_of_row,
                           b->R);
        hipLaunchKernelGGL(k_mr_scale, dim3(nbk(ne)), dim3(256), 0, b->st, ne, b->ratio, b->seg_of_row, b->dinv, b->R,
                           b->P);
        hipLaunchKernelGGL(k_chunk_rz, dim3(b->n_chunks), dim3(256), 0, b->st, b->chunk_row, b->dinv, b->R,
                           b->partials);
        hipLaunchKernelGGL(k_sys_fin, dim3(b->n_var), dim3(256), 0, b->st, b->seg_chunk, b->partials, 4, b->sv,
                           b->frozen, b->tol2);
        hipLaunchKernelGGL(k_var_active, dim3((b->n_var + 255) / 256), dim3(256), 0, b->st, b->n_var, b->sv,
                           b->active);
    }
    BCHK(hipEventRecord(b->ev1, b->st));
    BCHK(hipEventSynchronize(b->ev1));
    float ms = 0;
    BCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    BCHK(hipMemcpy(h_U, b->U, ne * sizeof(double), hipMemcpyDeviceToHost));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->outer_iterations = outer;
        stats->lockstep_iterations = total_inner;
        double worst = 0.0;
        int64_t nc = 0, sum_it = 0;
        for (int s = 0; s < n_sys; ++s) {
            if ((s % kNrhs) >= nrhs || ff[s] == 0.0) continue;
            const double rel = std::sqrt(rr[s] / ff[s]);
            worst = std::max(worst, rel);
            nc += rel <= tol ? 0 : 1;
            sum_it += sys_it[s];
        }
        stats->max_rel_residual = worst;
        stats->unconverged = nc;
        stats->system_iterations = sum_it;
        stats->device_ms = ms;
    }
    return ASEND_OK;
}

}  // namespace asend
