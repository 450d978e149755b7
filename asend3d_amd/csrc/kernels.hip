// HIP kernels of the solve path for gfx950 (MI355X).
//
// Operator storage: node-block CSR ("BSR") with b x b blocks (b = DOFs per
// node).  Block row r holds the blocks of the nodes coupled to node r (sorted
// by node rank); within a block row, values are component-major:
//   value(block k of row r, component c = b*i + j) = val[b*b*rp[r] + c*nb_r + k]
// with nb_r = rp[r+1] - rp[r].  A group of G lanes sweeps one block row, so
// every value load is a unit-stride burst, and the column index is read once
// per b*b values (4 B per 9 values for b = 3 instead of 4 B per value).
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernels.hpp"

namespace asend {

// ---------------------------------------------------------------------------
// brick8 element integration + colored assembly (FP64).
// One 64-lane wavefront per element; lane (a, b) owns the 3x3 block K_ab.
// Gauss-point geometry (J^-1, det J, physical gradients) is computed once
// per element and shared through LDS.  Elements of one colour share no node,
// so the read-modify-write of each block is race-free without atomics.
//   K_ab,ij = sum_gp w detJ [lam g_a,i g_b,j + mu g_a,j g_b,i + mu (g_a.g_b) d_ij]
// is B_a^T C B_b for isotropic C (Lame lam, mu), the same operator as
// get_ruk_dfd0's B^T C B (S/element_equations.rs:398-411).
// ---------------------------------------------------------------------------
__constant__ double kHexNdSign[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                        {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

__global__ __launch_bounds__(64) void k_hex8_assemble(const int32_t* __restrict__ elems, int n_elems,
                                                      const int32_t* __restrict__ conn_r,
                                                      const double* __restrict__ xyz_r,
                                                      const double* __restrict__ E, const double* __restrict__ nu,
                                                      const double* __restrict__ rho, double c_mass,
                                                      const int32_t* __restrict__ sptr,
                                                      const uint8_t* __restrict__ slot, double* __restrict__ K) {
    const int lane = threadIdx.x;
    const int ei = blockIdx.x;
    if (ei >= n_elems) return;
    const int e = elems[ei];
    __shared__ double X[8][3];
    __shared__ double Ji[8][9];
    __shared__ double W[8];
    __shared__ double Gs[8][8][3];
    __shared__ double Ns[8][8];
    if (lane < 24) {
        int a = lane / 3, c = lane % 3;
        X[a][c] = xyz_r[3 * (int64_t)conn_r[8 * (int64_t)e + a] + c];
    }
    __syncthreads();
    const double g = 0.5773502691896257;
    if (lane < 8) {
        const int gp = lane;
        double s0 = (gp & 1) ? g : -g, s1 = (gp & 2) ? g : -g, s2 = (gp & 4) ? g : -g;
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int a = 0; a < 8; ++a) {
            double fx = 1.0 + kHexNdSign[a][0] * s0, fy = 1.0 + kHexNdSign[a][1] * s1,
                   fz = 1.0 + kHexNdSign[a][2] * s2;
            double d0 = 0.125 * kHexNdSign[a][0] * fy * fz, d1 = 0.125 * fx * kHexNdSign[a][1] * fz,
                   d2 = 0.125 * fx * fy * kHexNdSign[a][2];
            Ns[gp][a] = 0.125 * fx * fy * fz;
            for (int i = 0; i < 3; ++i) {
                J[i][0] += X[a][i] * d0;
                J[i][1] += X[a][i] * d1;
                J[i][2] += X[a][i] * d2;
            }
        }
        double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        double id = 1.0 / det;
        Ji[gp][0] = c00 * id;
        Ji[gp][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
        Ji[gp][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
        Ji[gp][3] = c01 * id;
        Ji[gp][4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
        Ji[gp][5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
        Ji[gp][6] = c02 * id;
        Ji[gp][7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
        Ji[gp][8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
        W[gp] = det;  // Gauss weight 1
    }
    __syncthreads();
    {
        const int gp = lane / 8, a = lane % 8;
        double s0 = (gp & 1) ? g : -g, s1 = (gp & 2) ? g : -g, s2 = (gp & 4) ? g : -g;
        double fx = 1.0 + kHexNdSign[a][0] * s0, fy = 1.0 + kHexNdSign[a][1] * s1, fz = 1.0 + kHexNdSign[a][2] * s2;
        double d0 = 0.125 * kHexNdSign[a][0] * fy * fz, d1 = 0.125 * fx * kHexNdSign[a][1] * fz,
               d2 = 0.125 * fx * fy * kHexNdSign[a][2];
        // dN/dx_k = sum_j dN/ds_j (J^-1)_jk
        for (int k = 0; k < 3; ++k) Gs[gp][a][k] = d0 * Ji[gp][0 + k] + d1 * Ji[gp][3 + k] + d2 * Ji[gp][6 + k];
    }
    __syncthreads();
    const int a = lane / 8, b = lane % 8;
    const double Ee = E[e], ne = nu[e];
    const double lam = Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne));
    const double mu = Ee / (2.0 * (1.0 + ne));
    double k[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double m = 0.0;
    for (int gp = 0; gp < 8; ++gp) {
        const double w = W[gp];
        double ga[3] = {Gs[gp][a][0], Gs[gp][a][1], Gs[gp][a][2]};
        double gb[3] = {Gs[gp][b][0], Gs[gp][b][1], Gs[gp][b][2]};
        double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                k[i][j] += w * (lam * ga[i] * gb[j] + mu * ga[j] * gb[i] + (i == j ? mu * dot : 0.0));
        m += w * Ns[gp][a] * Ns[gp][b];
    }
    if (c_mass != 0.0) {
        double cm = c_mass * rho[e] * m;
        k[0][0] += cm;
        k[1][1] += cm;
        k[2][2] += cm;
    }
    const int ra = conn_r[8 * (int64_t)e + a];
    const int kk = slot[64 * (int64_t)e + lane];
    double* base = K + sell_index(sptr, ra, kk, 0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) base[64 * (3 * i + j)] += k[i][j];
}

void launch_hex8_assemble(hipStream_t st, const int32_t* elems, int n_elems, const int32_t* conn_r,
                          const double* xyz_r, const double* E, const double* nu, const double* rho, double c_mass,
                          const int32_t* sptr, const uint8_t* slot, double* K) {
    if (n_elems <= 0) return;
    hipLaunchKernelGGL(k_hex8_assemble, dim3(n_elems), dim3(64), 0, st, elems, n_elems, conn_r, xyz_r, E, nu, rho,
                       c_mass, sptr, slot, K);
}

// ---------------------------------------------------------------------------
// Grid reductions: every block writes its partial sums with plain stores;
// a single-block finalize kernel (stream-ordered after the producer, so the
// partials are visible without device-scope fences) adds them in a fixed
// order and updates the loop scalars.  Deterministic run to run.
// ---------------------------------------------------------------------------
enum FinOp { FIN_PQ = 0, FIN_RR = 1, FIN_INIT = 2, FIN_UPDATE = 3, FIN_DOT = 4, FIN_MAX = 5 };

__global__ __launch_bounds__(kBlock) void k_finalize(const double* __restrict__ partials, int nparts, int op,
                                                     Scalars* sc) {
    if ((op == FIN_PQ || op == FIN_UPDATE) && sc->done) return;
    if (op == FIN_MAX) {
        double m = 0.0;
        for (int b = threadIdx.x; b < nparts; b += blockDim.x) m = fmax(m, partials[b]);
        m = wave_max(m);
        __shared__ double sh[kBlock / kWave];
        if (threadIdx.x % kWave == 0) sh[threadIdx.x / kWave] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            double mm = 0.0;
            for (int k = 0; k < (int)(blockDim.x / kWave); ++k) mm = fmax(mm, sh[k]);
            sc->maxabs = mm;
        }
        return;
    }
    const int nv = (op == FIN_INIT || op == FIN_UPDATE) ? 2 : 1;
    double t[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        t[0] += partials[(size_t)nv * b];
        if (nv == 2) t[1] += partials[(size_t)nv * b + 1];
    }
    block_sum<2>(t);
    if (threadIdx.x != 0) return;
    switch (op) {
        case FIN_PQ:
            sc->pq = t[0];
            sc->alpha = sc->rz / t[0];
            break;
        case FIN_RR:
            sc->rr = t[0];
            break;
        case FIN_DOT:
            sc->dot = t[0];
            break;
        case FIN_INIT:
            sc->rz = t[0];
            sc->rr = t[1];
            sc->ff = t[1];
            sc->it = 0;
            sc->done = (sc->stop_abs ? (t[0] <= sc->tol) : (t[1] <= sc->tol2 * t[1])) || t[1] == 0.0;
            break;
        case FIN_UPDATE: {
            sc->beta = t[0] / sc->rz;
            sc->rz = t[0];
            sc->rr = t[1];
            sc->it += 1;
            bool conv = sc->stop_abs ? (t[0] <= sc->tol) : (t[1] <= sc->tol2 * sc->ff);
            if (conv || sc->it >= sc->max_it) sc->done = 1;
            break;
        }
        default:
            break;
    }
}

static void finalize(hipStream_t st, const double* partials, int nparts, int op, Scalars* sc) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, st, partials, nparts, op, sc);
}

// ---------------------------------------------------------------------------
// Post-assembly: max |K| (penalty scale), penalty diagonal, FP32 copy,
// Jacobi diagonal.
// ---------------------------------------------------------------------------
constexpr int kMaxBlocks = 1024;

__global__ __launch_bounds__(kBlock) void k_maxabs(const double* __restrict__ v, int64_t n, double* partials) {
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(v[i]));
    block_partial_max(m, partials);
}

void launch_maxabs(hipStream_t st, const double* v, int64_t n, double* partials, Scalars* sc) {
    hipLaunchKernelGGL(k_maxabs, dim3(kMaxBlocks), dim3(kBlock), 0, st, v, n, partials);
    finalize(st, partials, kMaxBlocks, FIN_MAX, sc);
}

// Adds mult[i] * sf to the diagonal entry of DOF pen[i] (unique DOFs).
__global__ void k_penalty_diag(const int64_t* __restrict__ pen, const int32_t* __restrict__ mult, int n_pen,
                               const int32_t* __restrict__ sptr, const int32_t* __restrict__ kdiag,
                               double* __restrict__ K, const Scalars* __restrict__ sc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pen) return;
    int64_t d = pen[i];
    int r = (int)(d / 3), c = (int)(d % 3);
    K[sell_index(sptr, r, kdiag[r], 4 * c)] += mult[i] * sc->sf;
}

void launch_penalty_diag(hipStream_t st, const int64_t* pen, const int32_t* mult, int n_pen, const int32_t* sptr,
                         const int32_t* kdiag, double* K, const Scalars* sc) {
    if (n_pen <= 0) return;
    hipLaunchKernelGGL(k_penalty_diag, dim3((n_pen + 255) / 256), dim3(256), 0, st, pen, mult, n_pen, sptr, kdiag, K,
                       sc);
}

__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

void launch_to_f32(hipStream_t st, const double* a, float* b, int64_t n) {
    hipLaunchKernelGGL(k_to_f32, dim3(2048), dim3(kBlock), 0, st, a, b, n);
}

// Jacobi diagonal: dinv[d] = 1 / K_dd (FP64 and FP32 copies).
__global__ void k_diag_inv(int n_rows, const int32_t* __restrict__ sptr, const int32_t* __restrict__ kdiag,
                           const double* __restrict__ K, double* __restrict__ d64, float* __restrict__ d32) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    for (int c = 0; c < 3; ++c) {
        double iv = 1.0 / K[sell_index(sptr, r, kdiag[r], 4 * c)];
        d64[3 * (int64_t)r + c] = iv;
        d32[3 * (int64_t)r + c] = (float)iv;
    }
}

void launch_diag_inv(hipStream_t st, int n_rows, const int32_t* sptr, const int32_t* kdiag, const double* K,
                     double* d64, float* d32) {
    hipLaunchKernelGGL(k_diag_inv, dim3((n_rows + 255) / 256), dim3(256), 0, st, n_rows, sptr, kdiag, K, d64, d32);
}

// ---------------------------------------------------------------------------
// Sliced-ELL SpMV of the fine operator (3x3 blocks, 64-row slices): one
// wavefront per slice, lane = row.  Slot k of a slice holds one block per
// row (zero blocks pad rows shorter than the slice width), its column index
// at scol[64*k + lane] and its 9 values at val[64*(9k + c) + lane], so every
// load of the wavefront is one contiguous 256 B (FP32) / 512 B (FP64) burst
// and no cross-lane reduction is needed.  Modes as the block-row SpMV.
// ---------------------------------------------------------------------------
template <typename T, int MODE>
__global__ __launch_bounds__(kSellMaxWaves * kWave) void k_sell3(int n_rows, int n_slices,
                                                                 const int32_t* __restrict__ sptr,
                                                  const int32_t* __restrict__ scol, const T* __restrict__ val,
                                                  const T* __restrict__ x, T* __restrict__ y,
                                                  const T* __restrict__ f, double* partials, const Scalars* sc) {
    if (MODE == 1 && sc->done) return;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int slice = lb * (int)(blockDim.x / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int r = slice * kWave + lane;
    double part = 0.0;
    if (slice < n_slices) {
        const int k0 = sptr[slice], k1 = sptr[slice + 1];
        const int32_t* cp = scol + (int64_t)k0 * kWave + lane;
        const T* vp = val + (int64_t)k0 * 9 * kWave + lane;
        T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll 2
        for (int k = k0; k < k1; ++k) {
            const int c = *cp;
            const T x0 = x[3 * (int64_t)c], x1 = x[3 * (int64_t)c + 1], x2 = x[3 * (int64_t)c + 2];
            a0 += vp[0 * kWave] * x0 + vp[1 * kWave] * x1 + vp[2 * kWave] * x2;
            a1 += vp[3 * kWave] * x0 + vp[4 * kWave] * x1 + vp[5 * kWave] * x2;
            a2 += vp[6 * kWave] * x0 + vp[7 * kWave] * x1 + vp[8 * kWave] * x2;
            cp += kWave;
            vp += 9 * kWave;
        }
        if (r < n_rows) {
            if (MODE == 2 || MODE == 3) {
                T y0 = f[3 * (int64_t)r] - a0, y1 = f[3 * (int64_t)r + 1] - a1, y2 = f[3 * (int64_t)r + 2] - a2;
                y[3 * (int64_t)r] = y0;
                y[3 * (int64_t)r + 1] = y1;
                y[3 * (int64_t)r + 2] = y2;
                part = (double)y0 * y0 + (double)y1 * y1 + (double)y2 * y2;
            } else {
                y[3 * (int64_t)r] = a0;
                y[3 * (int64_t)r + 1] = a1;
                y[3 * (int64_t)r + 2] = a2;
                if (MODE == 1)
                    part = (double)a0 * x[3 * (int64_t)r] + (double)a1 * x[3 * (int64_t)r + 1] +
                           (double)a2 * x[3 * (int64_t)r + 2];
            }
        }
    }
    if (MODE == 0 || MODE == 3) return;
    double t[1] = {part};
    block_partials<1>(t, partials);
}

template <typename T>
void launch_sell3(hipStream_t st, int mode, const Sell3& A, const T* val, const T* x, T* y, const T* f,
                  double* partials, Scalars* sc) {
    const int per = A.waves_per_block;
    const int nblk = (A.n_slices + per - 1) / per;
    const int nthr = per * kWave;
    switch (mode) {
        case 0:
            hipLaunchKernelGGL((k_sell3<T, 0>), dim3(nblk), dim3(nthr), 0, st, A.n_rows, A.n_slices, A.sptr, A.scol,
                               val, x, y, f, partials, sc);
            break;
        case 1:
            hipLaunchKernelGGL((k_sell3<T, 1>), dim3(nblk), dim3(nthr), 0, st, A.n_rows, A.n_slices, A.sptr, A.scol,
                               val, x, y, f, partials, sc);
            finalize(st, partials, nblk, FIN_PQ, sc);
            break;
        case 2:
            hipLaunchKernelGGL((k_sell3<T, 2>), dim3(nblk), dim3(nthr), 0, st, A.n_rows, A.n_slices, A.sptr, A.scol,
                               val, x, y, f, partials, sc);
            finalize(st, partials, nblk, FIN_RR, sc);
            break;
        default:
            hipLaunchKernelGGL((k_sell3<T, 3>), dim3(nblk), dim3(nthr), 0, st, A.n_rows, A.n_slices, A.sptr, A.scol,
                               val, x, y, f, partials, sc);
            break;
    }
}

template void launch_sell3<double>(hipStream_t, int, const Sell3&, const double*, const double*, double*,
                                   const double*, double*, Scalars*);
template void launch_sell3<float>(hipStream_t, int, const Sell3&, const float*, const float*, float*, const float*,
                                  double*, Scalars*);

// ---------------------------------------------------------------------------
// Block-row SpMV, b = 3, G lanes per block row, component-major values.
// MODE 0: y = K x
// MODE 1: y = K x and (x . y) -> finalize: sc->pq, sc->alpha = sc->rz / pq
// MODE 2: y = f - K x (x = current solution, f = load) -> finalize: sc->rr = |y|^2
// MODE 3: y = f - K x without the norm
// Kernels with a scalars pointer return at once when sc->done is set.
// ---------------------------------------------------------------------------
template <typename T, int G, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv3(int n_rows, int rows_per_block, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ bc, const T* __restrict__ val,
                                                  const T* __restrict__ x, T* __restrict__ y,
                                                  const T* __restrict__ f, double* partials, const Scalars* sc) {
    if (MODE == 1 && sc->done) return;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int sub = threadIdx.x % G;
    const int grp = threadIdx.x / G;
    constexpr int kGroups = kBlock / G;
    const int r0 = lb * rows_per_block;
    int r1 = r0 + rows_per_block;
    if (r1 > n_rows) r1 = n_rows;
    double part = 0.0;
    for (int r = r0 + grp; r < r1; r += kGroups) {
        const int s = rp[r], nb = rp[r + 1] - s;
        const T* v = val + 9 * (int64_t)s;
        T a0 = 0, a1 = 0, a2 = 0;
        for (int k = sub; k < nb; k += G) {
            const int c = bc[s + k];
            const T x0 = x[3 * (int64_t)c], x1 = x[3 * (int64_t)c + 1], x2 = x[3 * (int64_t)c + 2];
            a0 += v[0 * nb + k] * x0 + v[1 * nb + k] * x1 + v[2 * nb + k] * x2;
            a1 += v[3 * nb + k] * x0 + v[4 * nb + k] * x1 + v[5 * nb + k] * x2;
            a2 += v[6 * nb + k] * x0 + v[7 * nb + k] * x1 + v[8 * nb + k] * x2;
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) {
            a0 += __shfl_xor(a0, o, G);
            a1 += __shfl_xor(a1, o, G);
            a2 += __shfl_xor(a2, o, G);
        }
        if (sub == 0) {
            if (MODE == 2 || MODE == 3) {
                T y0 = f[3 * (int64_t)r] - a0, y1 = f[3 * (int64_t)r + 1] - a1, y2 = f[3 * (int64_t)r + 2] - a2;
                y[3 * (int64_t)r] = y0;
                y[3 * (int64_t)r + 1] = y1;
                y[3 * (int64_t)r + 2] = y2;
                part += (double)y0 * y0 + (double)y1 * y1 + (double)y2 * y2;
            } else {
                y[3 * (int64_t)r] = a0;
                y[3 * (int64_t)r + 1] = a1;
                y[3 * (int64_t)r + 2] = a2;
                if (MODE == 1)
                    part += (double)a0 * x[3 * (int64_t)r] + (double)a1 * x[3 * (int64_t)r + 1] +
                            (double)a2 * x[3 * (int64_t)r + 2];
            }
        }
    }
    if (MODE == 0 || MODE == 3) return;
    double t[1] = {part};
    block_partials<1>(t, partials);
}

template <typename T>
void launch_spmv3(hipStream_t st, int mode, int n_rows, const int32_t* rp, const int32_t* bc, const T* val,
                  const T* x, T* y, const T* f, double* partials, Scalars* sc) {
    constexpr int G = 32;
    const int nblk = kSpmvBlocks;
    int rpb = (n_rows + nblk - 1) / nblk;
    if (mode == 0) {
        hipLaunchKernelGGL((k_spmv3<T, G, 0>), dim3(nblk), dim3(kBlock), 0, st, n_rows, rpb, rp, bc, val, x, y, f,
                           partials, sc);
    } else if (mode == 1) {
        hipLaunchKernelGGL((k_spmv3<T, G, 1>), dim3(nblk), dim3(kBlock), 0, st, n_rows, rpb, rp, bc, val, x, y, f,
                           partials, sc);
        finalize(st, partials, nblk, FIN_PQ, sc);
    } else if (mode == 2) {
        hipLaunchKernelGGL((k_spmv3<T, G, 2>), dim3(nblk), dim3(kBlock), 0, st, n_rows, rpb, rp, bc, val, x, y, f,
                           partials, sc);
        finalize(st, partials, nblk, FIN_RR, sc);
    } else {
        hipLaunchKernelGGL((k_spmv3<T, G, 3>), dim3(nblk), dim3(kBlock), 0, st, n_rows, rpb, rp, bc, val, x, y, f,
                           partials, sc);
    }
}

template void launch_spmv3<double>(hipStream_t, int, int, const int32_t*, const int32_t*, const double*,
                                   const double*, double*, const double*, double*, Scalars*);
template void launch_spmv3<float>(hipStream_t, int, int, const int32_t*, const int32_t*, const float*, const float*,
                                  float*, const float*, double*, Scalars*);

// ---------------------------------------------------------------------------
// Jacobi PCG vector kernels (T = double for the FP64 solver, float for the
// inner FP32 solver).  Reductions are always combined in FP64.
// ---------------------------------------------------------------------------

// r = b, x = 0, p = z = Dinv r; partials of r.z and r.r
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int64_t n, const T* __restrict__ b, const T* __restrict__ dinv,
                                                     T* __restrict__ x, T* __restrict__ r, T* __restrict__ p,
                                                     double* partials) {
    double t[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        T ri = b[i];
        T zi = dinv[i] * ri;
        x[i] = 0;
        r[i] = ri;
        p[i] = zi;
        t[0] += (double)ri * zi;
        t[1] += (double)ri * ri;
    }
    block_partials<2>(t, partials);
}

// x += alpha p; r -= alpha q; partials of r.(Dinv r) and r.r
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int64_t n, const T* __restrict__ dinv, T* __restrict__ x,
                                                       T* __restrict__ r, const T* __restrict__ p,
                                                       const T* __restrict__ q, double* partials,
                                                       const Scalars* sc) {
    if (sc->done) return;
    const T alpha = (T)sc->alpha;
    double t[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        T ri = r[i] - alpha * q[i];
        r[i] = ri;
        t[0] += (double)ri * (dinv[i] * ri);
        t[1] += (double)ri * ri;
    }
    block_partials<2>(t, partials);
}

// p = Dinv r + beta p
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(int64_t n, const T* __restrict__ dinv, const T* __restrict__ r,
                                                    T* __restrict__ p, const Scalars* __restrict__ sc) {
    if (sc->done) return;
    const T beta = (T)sc->beta;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = dinv[i] * r[i] + beta * p[i];
}

template <typename T>
void launch_pcg_init(hipStream_t st, int64_t n, const T* b, const T* dinv, T* x, T* r, T* p, double* partials,
                     Scalars* sc) {
    hipLaunchKernelGGL(k_pcg_init<T>, dim3(kVecBlocks), dim3(kBlock), 0, st, n, b, dinv, x, r, p, partials);
    finalize(st, partials, kVecBlocks, FIN_INIT, sc);
}
template <typename T>
void launch_pcg_update(hipStream_t st, int64_t n, const T* dinv, T* x, T* r, const T* p, const T* q,
                       double* partials, Scalars* sc) {
    hipLaunchKernelGGL(k_pcg_update<T>, dim3(kVecBlocks), dim3(kBlock), 0, st, n, dinv, x, r, p, q, partials, sc);
    finalize(st, partials, kVecBlocks, FIN_UPDATE, sc);
}
template <typename T>
void launch_pcg_dir(hipStream_t st, int64_t n, const T* dinv, const T* r, T* p, const Scalars* sc) {
    hipLaunchKernelGGL(k_pcg_dir<T>, dim3(kVecBlocks), dim3(kBlock), 0, st, n, dinv, r, p, sc);
}

#define INST_PCG(T)                                                                                               \
    template void launch_pcg_init<T>(hipStream_t, int64_t, const T*, const T*, T*, T*, T*, double*, Scalars*);    \
    template void launch_pcg_update<T>(hipStream_t, int64_t, const T*, T*, T*, const T*, const T*, double*,       \
                                       Scalars*);                                                                 \
    template void launch_pcg_dir<T>(hipStream_t, int64_t, const T*, const T*, T*, const Scalars*);
INST_PCG(double)
INST_PCG(float)

// ---------------------------------------------------------------------------
// Mixed precision glue.
// ---------------------------------------------------------------------------
// r32 = r64 / ||r64||  (norm from sc->rr)
__global__ void k_scale_to_f32(int64_t n, const double* __restrict__ r, float* __restrict__ r32,
                               const Scalars* __restrict__ sc) {
    const double s = 1.0 / sqrt(sc->rr);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        r32[i] = (float)(r[i] * s);
}
// u += ||r64|| * d32
__global__ void k_axpy_from_f32(int64_t n, double* __restrict__ u, const float* __restrict__ d,
                                const Scalars* __restrict__ sc_outer) {
    const double s = sqrt(sc_outer->rr);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        u[i] += s * (double)d[i];
}

void launch_scale_to_f32(hipStream_t st, int64_t n, const double* r, float* r32, const Scalars* sc) {
    hipLaunchKernelGGL(k_scale_to_f32, dim3(kVecBlocks), dim3(kBlock), 0, st, n, r, r32, sc);
}
void launch_axpy_from_f32(hipStream_t st, int64_t n, double* u, const float* d, const Scalars* sc_outer) {
    hipLaunchKernelGGL(k_axpy_from_f32, dim3(kVecBlocks), dim3(kBlock), 0, st, n, u, d, sc_outer);
}

// dot(a, b) in FP64 -> sc->dot (used for the compliance f.u)
__global__ __launch_bounds__(kBlock) void k_dot64(int64_t n, const double* __restrict__ a,
                                                  const double* __restrict__ b, double* partials) {
    double t[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        t[0] += a[i] * b[i];
    block_partials<1>(t, partials);
}
void launch_dot64(hipStream_t st, int64_t n, const double* a, const double* b, double* partials, Scalars* sc) {
    hipLaunchKernelGGL(k_dot64, dim3(kVecBlocks), dim3(kBlock), 0, st, n, a, b, partials);
    finalize(st, partials, kVecBlocks, FIN_DOT, sc);
}

__global__ void k_axpy64(int64_t n, double alpha, const double* __restrict__ x, double* __restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] += alpha * x[i];
}
void launch_axpy64(hipStream_t st, int64_t n, double alpha, const double* x, double* y) {
    hipLaunchKernelGGL(k_axpy64, dim3(kVecBlocks), dim3(kBlock), 0, st, n, alpha, x, y);
}

// ---------------------------------------------------------------------------
// Compliance sensitivity, brick8: dJ/dE_e = -(1/E_e) lam_e^T K_e u_e
//   lam_e^T K_e u_e = sum_gp w detJ [lame (div u)(div lam) + 2 mu eps(u):eps(lam)]
// One thread per element, FP64 throughout (the adjoint reduction stays FP64).
// Restates the product lambda^T dR/dd of the DV loop
// (S/model_analysis.rs:1694-1696) for per-element modulus variables.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_hex8_sens(int n_el, const int32_t* __restrict__ conn_r,
                                                   const double* __restrict__ xyz_r, const double* __restrict__ E,
                                                   const double* __restrict__ nu, const double* __restrict__ u,
                                                   const double* __restrict__ lam, double* __restrict__ out) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    double X[8][3], U[8][3], L[8][3];
    for (int a = 0; a < 8; ++a) {
        int64_t r = conn_r[8 * (int64_t)e + a];
        for (int c = 0; c < 3; ++c) {
            X[a][c] = xyz_r[3 * r + c];
            U[a][c] = u[3 * r + c];
            L[a][c] = lam[3 * r + c];
        }
    }
    const double Ee = E[e], ne = nu[e];
    const double lamL = ne / ((1.0 + ne) * (1.0 - 2.0 * ne));  // lame / E
    const double mu = 1.0 / (2.0 * (1.0 + ne));                 // mu / E
    const double g = 0.5773502691896257;
    double acc = 0.0;
    for (int gp = 0; gp < 8; ++gp) {
        double s0 = (gp & 1) ? g : -g, s1 = (gp & 2) ? g : -g, s2 = (gp & 4) ? g : -g;
        double dN[8][3];
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int a = 0; a < 8; ++a) {
            double fx = 1.0 + kHexNdSign[a][0] * s0, fy = 1.0 + kHexNdSign[a][1] * s1,
                   fz = 1.0 + kHexNdSign[a][2] * s2;
            dN[a][0] = 0.125 * kHexNdSign[a][0] * fy * fz;
            dN[a][1] = 0.125 * fx * kHexNdSign[a][1] * fz;
            dN[a][2] = 0.125 * fx * fy * kHexNdSign[a][2];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) J[i][j] += X[a][i] * dN[a][j];
        }
        double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        double id = 1.0 / det;
        double Ji[3][3] = {{c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id,
                            (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
                           {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id,
                            (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
                           {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id,
                            (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
        double gu[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, gl[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int a = 0; a < 8; ++a) {
            double gx[3];
            for (int k = 0; k < 3; ++k) gx[k] = dN[a][0] * Ji[0][k] + dN[a][1] * Ji[1][k] + dN[a][2] * Ji[2][k];
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k) {
                    gu[i][k] += U[a][i] * gx[k];
                    gl[i][k] += L[a][i] * gx[k];
                }
        }
        double divu = gu[0][0] + gu[1][1] + gu[2][2], divl = gl[0][0] + gl[1][1] + gl[2][2];
        double ee = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) ee += 0.5 * (gu[i][k] + gu[k][i]) * 0.5 * (gl[i][k] + gl[k][i]);
        acc += det * (lamL * divu * divl + 2.0 * mu * ee);
    }
    (void)Ee;
    out[e] = -acc;  // lamL and mu above are already divided by E_e
}

void launch_hex8_sens(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                      const double* nu, const double* u, const double* lam, double* out) {
    hipLaunchKernelGGL(k_hex8_sens, dim3((n_el + 127) / 128), dim3(128), 0, st, n_el, conn_r, xyz_r, E, nu, u, lam,
                       out);
}

}  // namespace asend

namespace asend {

// Streaming copy used to calibrate HBM traffic counters and to measure the
// achievable copy bandwidth (read n*4 + write n*4 bytes per launch).
__global__ __launch_bounds__(kBlock) void k_copy_f32(int64_t n, const float4* __restrict__ a, float4* __restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

void launch_copy_f32(hipStream_t st, int64_t n4, const float* a, float* b) {
    hipLaunchKernelGGL(k_copy_f32, dim3(4096), dim3(kBlock), 0, st, n4, (const float4*)a, (float4*)b);
}

}  // namespace asend
