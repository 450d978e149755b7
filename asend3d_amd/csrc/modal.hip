// Vector kernels of the shift-invert Lanczos eigen-solve (natural
// frequencies): diagonal shifts of the operator, diagonal scalings, and the
// block projections c = V^T w, w += alpha V c against the stored Krylov basis
// V (m_cur vectors of length n, vector j at V + j n).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.hpp"
#include "kernels.hpp"

namespace asend {

namespace {

constexpr int kModalBlock = 256;
constexpr int kModalMaxVecs = 256;

// K_rr,ii += alpha d[3r + i] on the diagonal of every diagonal block
__global__ void k_add_diag(int n_rows, const int32_t* __restrict__ sptr, const int32_t* __restrict__ kdiag,
                           double* __restrict__ K, const double* __restrict__ d, double alpha) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    for (int i = 0; i < 3; ++i) K[sell_index(sptr, r, kdiag[r], 4 * i)] += alpha * d[3 * (int64_t)r + i];
}

// y = a .* x
__global__ void k_vmul(int64_t n, const double* __restrict__ a, const double* __restrict__ x, double* __restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = a[i] * x[i];
}

// partials[b * m + j] = sum over block b's elements of V_j . w
__global__ __launch_bounds__(kModalBlock) void k_vtw(int64_t n, int m, const double* __restrict__ V,
                                                     const double* __restrict__ w, double* __restrict__ partials) {
    __shared__ double red[kModalBlock / 64];
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t i0 = chunk * blockIdx.x, i1 = (i0 + chunk < n) ? i0 + chunk : n;
    for (int j = 0; j < m; ++j) {
        const double* v = V + (int64_t)j * n;
        double s = 0.0;
        for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) s += v[i] * w[i];
        s = wave_sum(s);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int k = 0; k < kModalBlock / 64; ++k) t += red[k];
            partials[(int64_t)blockIdx.x * m + j] = t;
        }
        __syncthreads();
    }
}

// c[j] = sum over blocks of partials[b * m + j] (fixed order)
__global__ void k_sum_cols(int nb, int m, const double* __restrict__ partials, double* __restrict__ c) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partials[(int64_t)b * m + j];
    c[j] = s;
}

// w += alpha sum_j V_j c[j]
__global__ void k_add_vc(int64_t n, int m, const double* __restrict__ V, const double* __restrict__ c, double alpha,
                         double* __restrict__ w) {
    __shared__ double cs[kModalMaxVecs];
    for (int j = threadIdx.x; j < m; j += blockDim.x) cs[j] = c[j];
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < m; ++j) s += V[(int64_t)j * n + i] * cs[j];
        w[i] += alpha * s;
    }
}

int nblk(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 2048); }

}  // namespace

int modal_max_vectors() { return kModalMaxVecs; }

void launch_add_diag(hipStream_t st, int n_rows, const int32_t* sptr, const int32_t* kdiag, double* K, const double* d,
                     double alpha) {
    hipLaunchKernelGGL(k_add_diag, dim3((n_rows + 255) / 256), dim3(256), 0, st, n_rows, sptr, kdiag, K, d, alpha);
}

void launch_vmul(hipStream_t st, int64_t n, const double* a, const double* x, double* y) {
    hipLaunchKernelGGL(k_vmul, dim3(nblk(n)), dim3(256), 0, st, n, a, x, y);
}

// c = V^T w (m <= modal_max_vectors()); partials holds nb * m doubles
void launch_vtw(hipStream_t st, int64_t n, int m, const double* V, const double* w, double* partials, int nb,
                double* c) {
    hipLaunchKernelGGL(k_vtw, dim3(nb), dim3(kModalBlock), 0, st, n, m, V, w, partials);
    hipLaunchKernelGGL(k_sum_cols, dim3((m + 63) / 64), dim3(64), 0, st, nb, m, partials, c);
}

void launch_add_vc(hipStream_t st, int64_t n, int m, const double* V, const double* c, double alpha, double* w) {
    hipLaunchKernelGGL(k_add_vc, dim3(nblk(n)), dim3(256), 0, st, n, m, V, c, alpha, w);
}

}  // namespace asend
