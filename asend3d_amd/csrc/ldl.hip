// Device LDL^T of the operator over diagonal DOF blocks: the reference's
// LowerTriMat (S/lower_tri_mat_meth.rs:15-227) built for the GPU.
//
// The reference stores each row from its first column inside its block (the
// envelope) and factors column by column with scalar loops.  Inside the
// envelope the factor fills completely and outside it stays exactly zero, so
// the same factor is obtained from a uniform band of 64 x 64 tiles that
// covers the widest row: the factorization is then a blocked right-looking
// LDL^T whose steps are one diagonal-tile factorization, a panel of
// tile products with D^-1 L_kk^-1 and a trailing update of tile products,
// all blocks of the numbering at once.  The solve streams a block's tiles
// with one workgroup, the last wt + 1 solution tiles held in LDS, or, for
// few long blocks, runs one launch per tile row across workgroups.
#include "ldl.hpp"

#include <algorithm>
#include <cmath>

#include "device_common.hpp"

namespace asend {

namespace {

constexpr int kLd = kLdlTile + 1;  // LDS row stride of a staged tile (doubles)

// Operator values inside each block's envelope into the banded tiles: one
// lane per block row (3 DOF rows) of a 64-row slice of the sliced-ELL FP64
// operator, every slot (3x3 block) of that row.
__global__ __launch_bounds__(kWave) void k_ldl_fill(LdlDesc d, int n_rows, const int32_t* __restrict__ sptr,
                                                    const int32_t* __restrict__ scol, const double* __restrict__ K,
                                                    double* __restrict__ L) {
    const int slice = blockIdx.x, lane = threadIdx.x;
    const int R = slice * kWave + lane;
    if (R >= n_rows) return;
    const int s0 = sptr[slice], ns = sptr[slice + 1] - s0;
    for (int k = 0; k < ns; ++k) {
        const int C = scol[(int64_t)kWave * (s0 + k) + lane];
        if (C < 0 || C > R) continue;  // upper blocks and padding
        for (int a = 0; a < 3; ++a) {
            const int64_t i = 3 * (int64_t)R + a;
            const int blk = (int)(i / d.bs);
            const int64_t b0 = (int64_t)blk * d.bs, li = i - b0;
            for (int c = 0; c < 3; ++c) {
                const int64_t j = 3 * (int64_t)C + c;
                if (j > i || j < b0) continue;
                const int64_t lj = j - b0;
                const int I = (int)(li / kLdlTile), J = (int)(lj / kLdlTile);
                if (I - J > d.wt) continue;
                const double v = K[((int64_t)9 * (s0 + k) + 3 * a + c) * kWave + lane];
                L[ldl_tile(d, blk, I, J) + (li % kLdlTile) * kLdlTile + lj % kLdlTile] += v;
            }
        }
    }
}

// Identity rows past the end of each block's last tile.
__global__ void k_ldl_pad(LdlDesc d, double* __restrict__ L) {
    const int blk = blockIdx.x;
    const int64_t s = (int64_t)blk * d.bs;
    const int64_t m = std::min(d.bs, d.n - s);
    const int nt = ldl_block_nt(d, blk);
    const int r = threadIdx.x;
    if ((int64_t)(nt - 1) * kLdlTile + r >= m) L[ldl_tile(d, blk, nt - 1, nt - 1) + r * kLdlTile + r] = 1.0;
}

__device__ __forceinline__ void stage_tile(double* __restrict__ s, const double* __restrict__ g) {
    for (int e = threadIdx.x; e < kLdlTileElems; e += blockDim.x) s[(e >> 6) * kLd + (e & 63)] = g[e];
}

// acc = A B^T over one staged 64 x 64 pair, 256 threads, 4 x 4 outputs each
// at rows ti + 16 x, columns tj + 16 y.
__device__ __forceinline__ void tile_abt(const double* __restrict__ sA, const double* __restrict__ sB,
                                         double (&acc)[4][4]) {
    const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
#pragma unroll 4
    for (int m = 0; m < kLdlTile; ++m) {
        double a[4], b[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) a[x] = sA[(ti + 16 * x) * kLd + m];
#pragma unroll
        for (int y = 0; y < 4; ++y) b[y] = sB[(tj + 16 * y) * kLd + m];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[x][y] = fma(a[x], b[y], acc[x][y]);
    }
}

// Step k, diagonal tile of every block: LDL^T in LDS (column j: pivot
// d_j, rank-1 update of the trailing lower triangle, then column j scaled by
// 1 / d_j), written back with zeros above the diagonal; then the panel
// operator P = D^-1 L_kk^-1 (row j of L_kk^-1 over d_j) for the next kernel.
__global__ __launch_bounds__(kBlock) void k_ldl_diag(LdlDesc d, int k, double* __restrict__ L,
                                                     double* __restrict__ P, int32_t* __restrict__ neg) {
    const int blk = blockIdx.x;
    if (k >= ldl_block_nt(d, blk)) return;
    __shared__ double A[kLdlTile * kLd];
    __shared__ double Li[kLdlTile * kLd];
    double* T = L + ldl_tile(d, blk, k, k);
    stage_tile(A, T);
    const int t = threadIdx.x;
    for (int j = 0; j < kLdlTile; ++j) {
        __syncthreads();
        const double dj = A[j * kLd + j];
        const double inv = 1.0 / dj;
        if (t == 0 && dj < 0.0) atomicAdd(neg, 1);
        for (int e = t; e < kLdlTileElems; e += kBlock) {
            const int i = e >> 6, m = e & 63;
            if (m > j && m <= i) A[i * kLd + m] -= A[i * kLd + j] * A[m * kLd + j] * inv;
        }
        __syncthreads();
        if (t > j && t < kLdlTile) A[t * kLd + j] *= inv;
    }
    __syncthreads();
    for (int e = t; e < kLdlTileElems; e += kBlock) {
        const int i = e >> 6, m = e & 63;
        T[e] = m <= i ? A[i * kLd + m] : 0.0;
    }
    // unit lower inverse, one column per thread (forward substitution)
    if (t < kLdlTile) {
        for (int r = 0; r < kLdlTile; ++r) {
            double v;
            if (r < t) {
                v = 0.0;
            } else if (r == t) {
                v = 1.0;
            } else {
                double s = 0.0;
                for (int m = t; m < r; ++m) s = fma(A[r * kLd + m], Li[m * kLd + t], s);
                v = -s;
            }
            Li[r * kLd + t] = v;
        }
    }
    __syncthreads();
    double* Pb = P + (int64_t)blk * kLdlTileElems;
    for (int e = t; e < kLdlTileElems; e += kBlock) {
        const int j = e >> 6, m = e & 63;
        Pb[e] = m <= j ? Li[j * kLd + m] / A[j * kLd + j] : 0.0;
    }
}

// Step k, panel: L_Ik = A_Ik L_kk^-T D_k^-1 = A_Ik P^T for the tiles below
// the diagonal (blockIdx.x = I - k - 1).
__global__ __launch_bounds__(kBlock) void k_ldl_panel(LdlDesc d, int k, double* __restrict__ L,
                                                      const double* __restrict__ P) {
    const int blk = blockIdx.y;
    const int I = k + 1 + blockIdx.x;
    if (I >= ldl_block_nt(d, blk)) return;
    __shared__ double sA[kLdlTile * kLd];
    __shared__ double sB[kLdlTile * kLd];
    double* X = L + ldl_tile(d, blk, I, k);
    stage_tile(sA, X);
    stage_tile(sB, P + (int64_t)blk * kLdlTileElems);
    __syncthreads();
    double acc[4][4] = {};
    tile_abt(sA, sB, acc);
    const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) X[(ti + 16 * x) * kLdlTile + tj + 16 * y] = acc[x][y];
}

// Step k, trailing update of the band: A_IJ -= L_Ik D_k L_Jk^T for
// k < J <= I <= k + wt (blockIdx.x enumerates the pairs row by row).
__global__ __launch_bounds__(kBlock) void k_ldl_update(LdlDesc d, int k, double* __restrict__ L) {
    const int blk = blockIdx.y;
    const int q = blockIdx.x;
    int pi = (int)((std::sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
    while ((pi + 1) * (pi + 2) / 2 <= q) ++pi;
    while (pi * (pi + 1) / 2 > q) --pi;
    const int pj = q - pi * (pi + 1) / 2;
    const int I = k + 1 + pi, J = k + 1 + pj;
    if (I >= ldl_block_nt(d, blk)) return;
    __shared__ double sA[kLdlTile * kLd];
    __shared__ double sB[kLdlTile * kLd];
    __shared__ double sd[kLdlTile];
    const double* Dk = L + ldl_tile(d, blk, k, k);
    if (threadIdx.x < kLdlTile) sd[threadIdx.x] = Dk[threadIdx.x * (kLdlTile + 1)];
    stage_tile(sA, L + ldl_tile(d, blk, I, k));
    __syncthreads();
    const double* XJ = L + ldl_tile(d, blk, J, k);
    for (int e = threadIdx.x; e < kLdlTileElems; e += kBlock) sB[(e >> 6) * kLd + (e & 63)] = XJ[e] * sd[e & 63];
    __syncthreads();
    double acc[4][4] = {};
    tile_abt(sA, sB, acc);
    double* C = L + ldl_tile(d, blk, I, J);
    const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) C[(ti + 16 * x) * kLdlTile + tj + 16 * y] -= acc[x][y];
}

// x = L^-T D^-1 L^-1 b, one workgroup per block (ldl_solve,
// S/lower_tri_mat_meth.rs:194-227).  Forward: tile row I of L times the
// solution tiles J < I (lane c of wave w reads column c of rows w + 4p, so
// every load is one 512 B row segment), reduced per row, then the unit lower
// diagonal tile solved by wave 0.  Backward: column tile I of L^T times the
// solution tiles J > I, then z / d minus it, solved with L_II^T.  The
// solution tiles J of the current window live in an LDS ring of wt + 1 tiles.
__global__ __launch_bounds__(kBlock) void k_ldl_solve(LdlDesc d, const double* __restrict__ L,
                                                      const double* b, double* x, const int* __restrict__ done) {
    if (done && *done) return;
    extern __shared__ double smem[];
    const int ring_n = d.wt + 1;
    double* ring = smem;                          // ring_n * 64
    double* sD = ring + (int64_t)ring_n * kLdlTile;  // 64 x kLd
    double* red = sD + kLdlTile * kLd;            // 4 x 64
    const int blk = blockIdx.x;
    const int nt = ldl_block_nt(d, blk);
    const int64_t s = (int64_t)blk * d.bs;
    const int64_t m = std::min(d.bs, d.n - s);
    const int t = threadIdx.x, w = t >> 6, c = t & 63;

    for (int I = 0; I < nt; ++I) {
        double acc[16];
#pragma unroll
        for (int p = 0; p < 16; ++p) acc[p] = 0.0;
        const int j0 = I - d.wt > 0 ? I - d.wt : 0;
        for (int J = j0; J < I; ++J) {
            const double* T = L + ldl_tile(d, blk, I, J);
            const double zc = ring[(J % ring_n) * kLdlTile + c];
#pragma unroll
            for (int p = 0; p < 16; ++p) acc[p] = fma(T[(w + 4 * p) * kLdlTile + c], zc, acc[p]);
        }
        stage_tile(sD, L + ldl_tile(d, blk, I, I));
#pragma unroll
        for (int p = 0; p < 16; ++p) {
            const double v = wave_sum(acc[p]);
            if (c == 0) red[w + 4 * p] = v;
        }
        __syncthreads();
        if (w == 0) {
            const int64_t gi = (int64_t)I * kLdlTile + c;
            double y = (gi < m ? b[s + gi] : 0.0) - red[c];
            for (int j = 0; j < kLdlTile; ++j) {
                const double zj = __shfl(y, j, kWave);
                if (c > j) y -= sD[c * kLd + j] * zj;
            }
            ring[(I % ring_n) * kLdlTile + c] = y;
            if (gi < m) x[s + gi] = y;
        }
        __syncthreads();
    }
    for (int I = nt - 1; I >= 0; --I) {
        double acc = 0.0;
        const int j1 = I + d.wt < nt - 1 ? I + d.wt : nt - 1;
        for (int J = I + 1; J <= j1; ++J) {
            const double* T = L + ldl_tile(d, blk, J, I);
            const double* xr = ring + (J % ring_n) * kLdlTile;
#pragma unroll
            for (int p = 0; p < 16; ++p) acc = fma(T[(w + 4 * p) * kLdlTile + c], xr[w + 4 * p], acc);
        }
        red[w * kLdlTile + c] = acc;
        stage_tile(sD, L + ldl_tile(d, blk, I, I));
        __syncthreads();
        if (w == 0) {
            const int64_t gi = (int64_t)I * kLdlTile + c;
            const double sum = red[c] + red[kLdlTile + c] + red[2 * kLdlTile + c] + red[3 * kLdlTile + c];
            const double z = gi < m ? x[s + gi] : 0.0;
            double y = z / sD[c * kLd + c] - sum;
            for (int j = kLdlTile - 1; j >= 0; --j) {
                const double xj = __shfl(y, j, kWave);
                if (c < j) y -= sD[j * kLd + c] * xj;
            }
            ring[(I % ring_n) * kLdlTile + c] = y;
            if (gi < m) x[s + gi] = y;
        }
        __syncthreads();
    }
}

// The same solve for few, long blocks (the direct factor is one block): one
// launch per tile row, right-looking, with the work of a step spread over the
// tiles below the diagonal; y (work vector) starts as b.  Forward step I:
// every workgroup k solves the unit lower diagonal tile for z_I = L_II^-1 y_I
// itself (y_I is final: its updates came from earlier steps), workgroup 0
// stores x_I = D_I^-1 z_I, and workgroup k subtracts L_KI z_I from y_K,
// K = I + 1 + k.  Backward step J: y_J = L_JJ^-T x_J (x_J final), then
// x_K -= L_JK^T y_J for K = J - 1 - k; the solution ends in y and is copied
// to x.  No step writes what another workgroup of the same step reads, and
// each vector entry is updated by one workgroup per step in step order, so
// the result is deterministic.
__device__ __forceinline__ void tri_solve_lower(const double* __restrict__ sD, double& y, int c) {
    for (int j = 0; j < kLdlTile; ++j) {
        const double zj = __shfl(y, j, kWave);
        if (c > j) y -= sD[c * kLd + j] * zj;
    }
}

__device__ __forceinline__ void tri_solve_upper(const double* __restrict__ sD, double& y, int c) {
    for (int j = kLdlTile - 1; j >= 0; --j) {
        const double xj = __shfl(y, j, kWave);
        if (c < j) y -= sD[j * kLd + c] * xj;
    }
}

__global__ __launch_bounds__(kBlock) void k_ldl_copy(int64_t n, const double* b, double* x, const int* done) {
    if (done && *done) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = b[i];
}

__global__ __launch_bounds__(kBlock) void k_ldl_fwd_step(LdlDesc d, const double* __restrict__ L,
                                                         double* __restrict__ y, double* __restrict__ x, int I,
                                                         const int* __restrict__ done) {
    if (done && *done) return;
    const int blk = blockIdx.y;
    const int nt = ldl_block_nt(d, blk);
    if (I >= nt) return;
    const int r = min(d.wt, nt - 1 - I);
    const int k = blockIdx.x;
    if (k > 0 && k >= r) return;
    __shared__ double sD[kLdlTile * kLd];
    __shared__ double sz[kLdlTile];
    __shared__ double red[kLdlTile];
    const int64_t s = (int64_t)blk * d.bs;
    const int64_t m = std::min(d.bs, d.n - s);
    const int t = threadIdx.x, w = t >> 6, c = t & 63;
    stage_tile(sD, L + ldl_tile(d, blk, I, I));
    __syncthreads();
    if (w == 0) {
        const int64_t gi = (int64_t)I * kLdlTile + c;
        double v = gi < m ? y[s + gi] : 0.0;
        tri_solve_lower(sD, v, c);
        sz[c] = v;
        if (k == 0 && gi < m) x[s + gi] = v / sD[c * kLd + c];
    }
    __syncthreads();
    if (k >= r) return;
    const int K = I + 1 + k;
    const double* T = L + ldl_tile(d, blk, K, I);
    const double zc = sz[c];
    double acc[16];
#pragma unroll
    for (int p = 0; p < 16; ++p) acc[p] = T[(w + 4 * p) * kLdlTile + c] * zc;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const double v = wave_sum(acc[p]);
        if (c == 0) red[w + 4 * p] = v;
    }
    __syncthreads();
    if (t < kLdlTile) {
        const int64_t gk = (int64_t)K * kLdlTile + t;
        if (gk < m) y[s + gk] -= red[t];
    }
}

__global__ __launch_bounds__(kBlock) void k_ldl_bwd_step(LdlDesc d, const double* __restrict__ L,
                                                         double* __restrict__ x, double* __restrict__ y, int J,
                                                         const int* __restrict__ done) {
    if (done && *done) return;
    const int blk = blockIdx.y;
    const int nt = ldl_block_nt(d, blk);
    if (J >= nt) return;
    const int r = min(d.wt, J);
    const int k = blockIdx.x;
    if (k > 0 && k >= r) return;
    __shared__ double sD[kLdlTile * kLd];
    __shared__ double sx[kLdlTile];
    __shared__ double red[4 * kLdlTile];
    const int64_t s = (int64_t)blk * d.bs;
    const int64_t m = std::min(d.bs, d.n - s);
    const int t = threadIdx.x, w = t >> 6, c = t & 63;
    stage_tile(sD, L + ldl_tile(d, blk, J, J));
    __syncthreads();
    if (w == 0) {
        const int64_t gj = (int64_t)J * kLdlTile + c;
        double v = gj < m ? x[s + gj] : 0.0;
        tri_solve_upper(sD, v, c);
        sx[c] = v;
        if (k == 0 && gj < m) y[s + gj] = v;
    }
    __syncthreads();
    if (k >= r) return;
    const int K = J - 1 - k;
    const double* T = L + ldl_tile(d, blk, J, K);
    double acc = 0.0;
#pragma unroll
    for (int p = 0; p < 16; ++p) acc = fma(T[(w + 4 * p) * kLdlTile + c], sx[w + 4 * p], acc);
    red[w * kLdlTile + c] = acc;
    __syncthreads();
    if (t < kLdlTile) {
        const double v = red[t] + red[kLdlTile + t] + red[2 * kLdlTile + t] + red[3 * kLdlTile + t];
        x[s + (int64_t)K * kLdlTile + t] -= v;  // K < J: a full tile
    }
}

size_t solve_lds_bytes(const LdlDesc& d) {
    return ((size_t)(d.wt + 1) * kLdlTile + (size_t)kLdlTile * kLd + 4 * kLdlTile) * sizeof(double);
}

}  // namespace

const char* ldl_plan(LdlDesc* d, int64_t* max_bw, int64_t n, int64_t block_dofs, const std::vector<int32_t>& h_rp,
                     const std::vector<int32_t>& h_bc) {
    if (n <= 0) return "empty operator";
    const int64_t n_rows = (int64_t)h_rp.size() - 1;
    if (3 * n_rows != n) return "operator rows are not 3-DOF block rows";
    const int64_t bs = (block_dofs <= 0 || block_dofs >= n) ? n : block_dofs;
    // envelope: first column of each row inside its block
    // (allocate_from_sparse_mat, S/lower_tri_mat_meth.rs:15-42)
    int64_t W = 1;
    for (int64_t R = 0; R < n_rows; ++R) {
        for (int a = 0; a < 3; ++a) {
            const int64_t i = 3 * R + a;
            const int64_t b0 = (i / bs) * bs;
            int64_t mc = i;
            for (int32_t k = h_rp[R]; k < h_rp[R + 1]; ++k) {
                const int64_t lo = 3 * (int64_t)h_bc[k], hi = lo + 2;
                if (hi < b0) continue;
                const int64_t cand = std::max(lo, b0);
                if (cand < mc) mc = cand;
            }
            W = std::max(W, i - mc + 1);
        }
    }
    d->n = n;
    d->bs = bs;
    d->n_blocks = (int)((n + bs - 1) / bs);
    d->nt_full = (int)((bs + kLdlTile - 1) / kLdlTile);
    d->wt = (int)std::min<int64_t>((W - 1 + kLdlTile - 1) / kLdlTile, d->nt_full - 1);
    if (d->wt > kLdlMaxWt) return "bandwidth exceeds the device solve's window (kLdlMaxWt tiles)";
    d->stride = ldl_row_off(d->nt_full, d->wt) * kLdlTileElems;
    *max_bw = W;
    return nullptr;
}

void ldl_fill(hipStream_t st, const LdlDesc& d, int n_rows, int n_slices, const int32_t* sptr, const int32_t* scol,
              const double* K, double* L) {
    (void)hipMemsetAsync(L, 0, (size_t)d.n_blocks * d.stride * sizeof(double), st);
    hipLaunchKernelGGL(k_ldl_fill, dim3(n_slices), dim3(kWave), 0, st, d, n_rows, sptr, scol, K, L);
    hipLaunchKernelGGL(k_ldl_pad, dim3(d.n_blocks), dim3(kLdlTile), 0, st, d, L);
}

void ldl_factor_launch(hipStream_t st, const LdlDesc& d, double* L, double* P, int32_t* neg) {
    (void)hipMemsetAsync(neg, 0, sizeof(int32_t), st);
    for (int k = 0; k < d.nt_full; ++k) {
        hipLaunchKernelGGL(k_ldl_diag, dim3(d.n_blocks), dim3(kBlock), 0, st, d, k, L, P, neg);
        const int r = std::min(d.wt, d.nt_full - 1 - k);
        if (r <= 0) continue;
        hipLaunchKernelGGL(k_ldl_panel, dim3(r, d.n_blocks), dim3(kBlock), 0, st, d, k, L, P);
        hipLaunchKernelGGL(k_ldl_update, dim3(r * (r + 1) / 2, d.n_blocks), dim3(kBlock), 0, st, d, k, L);
    }
}

void ldl_solve_launch(hipStream_t st, const LdlDesc& d, const double* L, const double* b, double* x, double* work,
                      const int* done) {
    if (d.n_blocks < kLdlStepBlocks) {
        hipLaunchKernelGGL(k_ldl_copy, dim3(1024), dim3(kBlock), 0, st, d.n, b, work, done);
        const int gx = std::max(1, d.wt);
        for (int I = 0; I < d.nt_full; ++I)
            hipLaunchKernelGGL(k_ldl_fwd_step, dim3(gx, d.n_blocks), dim3(kBlock), 0, st, d, L, work, x, I, done);
        for (int J = d.nt_full - 1; J >= 0; --J)
            hipLaunchKernelGGL(k_ldl_bwd_step, dim3(gx, d.n_blocks), dim3(kBlock), 0, st, d, L, x, work, J, done);
        hipLaunchKernelGGL(k_ldl_copy, dim3(1024), dim3(kBlock), 0, st, d.n, work, x, done);
        return;
    }
    static bool attr_set = false;
    const size_t lds = solve_lds_bytes(d);
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void*)k_ldl_solve, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)solve_lds_bytes(LdlDesc{0, 0, 0, 0, kLdlMaxWt, 0}));
        attr_set = true;
    }
    hipLaunchKernelGGL(k_ldl_solve, dim3(d.n_blocks), dim3(kBlock), lds, st, d, L, b, x, done);
}

}  // namespace asend
