// Geometrically nonlinear brick8 statics (total Lagrangian, St. Venant-
// Kirchhoff): element residual and tangent at a displacement state (gathered
// into the sliced-ELL operator by the shared row gather) and the penalty part
// of the Newton right-hand side.
//
// Strain: Green-Lagrange E = (H + H^T + H^T H) / 2 with H the displacement
// gradient (S/element_soln_fields.rs:1345-1461, n_lgeom terms); stress
// S = C E for isotropic C.  Per Gauss point (8 points, weight 1):
//   r_t   += detJ  F S g_t                                  (first Piola on g_t)
//   K_ts  += detJ [lam (F g_t)(F g_s)^T + mu ((F F^T)(g_t.g_s) + (F g_s)(F g_t)^T)
//                  + (g_t^T S g_s) I]
// which is get_ruk_dfd0's B(u)^T C B(u) plus its geometric term
// sigma . d2E/du du (S/element_equations.rs:206-418, :319-331, :398-411).
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernels.hpp"

namespace asend {

namespace {

constexpr int kNlWaves = 4;  // elements per workgroup (one wavefront each)

__device__ __forceinline__ double group8_sum(double v) {
    v += __shfl_xor(v, 1, kWave);
    v += __shfl_xor(v, 2, kWave);
    v += __shfl_xor(v, 4, kWave);
    return v;
}

__device__ __forceinline__ double stride8_sum(double v) {
    v += __shfl_xor(v, 8, kWave);
    v += __shfl_xor(v, 16, kWave);
    v += __shfl_xor(v, 32, kWave);
    return v;
}

// One wavefront per element.  Phase 1, lane = 8 gp + t: the Jacobian and the
// displacement gradient at Gauss point gp are sums over the 8 nodes t, formed
// by 8-lane butterflies; each lane then holds the kinematics of its point,
// adds detJ F S g_t to the residual of node t (summed over gp by a second
// butterfly) and the group leader stages F, F F^T and S in LDS.  Phase 2,
// lane = 8 a + b: the 3x3 tangent block of nodes (a, b) summed over the 8
// points, stored as row records: Ke[kRecStride (8 e + a) + 9 b + 3 i + k],
// Re[24 e + 3 t + i].
__global__ __launch_bounds__(256) void k_hex8_nl_element(int n_el, const int32_t* __restrict__ conn_r,
                                                         const double* __restrict__ xyz_r,
                                                         const double* __restrict__ E,
                                                         const double* __restrict__ nu,
                                                         const double* __restrict__ u, double* __restrict__ Ke,
                                                         double* __restrict__ Re) {
    __shared__ double s_g[kNlWaves][8][8][3];
    __shared__ double s_det[kNlWaves][8];
    __shared__ double s_F[kNlWaves][8][9];
    __shared__ double s_FF[kNlWaves][8][6];
    __shared__ double s_S[kNlWaves][8][6];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int e = blockIdx.x * kNlWaves + wv;
    const bool live = e < n_el;
    const int ec = live ? e : n_el - 1;
    const int gp = lane >> 3, t = lane & 7;
    const double g = 0.5773502691896257;

    const int node = conn_r[8 * (int64_t)ec + t];
    double X[3], U[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        X[c] = xyz_r[3 * (int64_t)node + c];
        U[c] = u[3 * (int64_t)node + c];
    }
    // natural derivatives of N_t at point gp
    const double sx = (((t ^ (t >> 1)) & 1) != 0) ? 1.0 : -1.0, sy = (t & 2) ? 1.0 : -1.0, sz = (t & 4) ? 1.0 : -1.0;
    const double p0 = (gp & 1) ? g : -g, p1 = (gp & 2) ? g : -g, p2 = (gp & 4) ? g : -g;
    const double fx = 1.0 + sx * p0, fy = 1.0 + sy * p1, fz = 1.0 + sz * p2;
    const double dn[3] = {0.125 * sx * fy * fz, 0.125 * fx * sy * fz, 0.125 * fx * fy * sz};
    double J[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) J[3 * i + c] = group8_sum(X[i] * dn[c]);
    const double c00 = J[4] * J[8] - J[5] * J[7];
    const double c01 = J[5] * J[6] - J[3] * J[8];
    const double c02 = J[3] * J[7] - J[4] * J[6];
    const double det = J[0] * c00 + J[1] * c01 + J[2] * c02;
    const double id = 1.0 / det;
    const double Ji[9] = {c00 * id,
                          (J[2] * J[7] - J[1] * J[8]) * id,
                          (J[1] * J[5] - J[2] * J[4]) * id,
                          c01 * id,
                          (J[0] * J[8] - J[2] * J[6]) * id,
                          (J[2] * J[3] - J[0] * J[5]) * id,
                          c02 * id,
                          (J[1] * J[6] - J[0] * J[7]) * id,
                          (J[0] * J[4] - J[1] * J[3]) * id};
    double gt[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) gt[c] = dn[0] * Ji[0 + c] + dn[1] * Ji[3 + c] + dn[2] * Ji[6 + c];
    // F = I + H, H_ij = sum_t u_t,i g_t,j
    double F[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[3 * i + j] = group8_sum(U[i] * gt[j]) + (i == j ? 1.0 : 0.0);
    // Green-Lagrange strain and second Piola stress
    double Eg[6];  // E00 E11 E22 E01 E02 E12
    {
        const int pj[6] = {0, 1, 2, 0, 0, 1}, pl[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            const int j = pj[m], l = pl[m];
            const double ftf = F[j] * F[l] + F[3 + j] * F[3 + l] + F[6 + j] * F[6 + l];
            Eg[m] = 0.5 * (ftf - (j == l ? 1.0 : 0.0));
        }
    }
    const double Ee = E[ec], ne = nu[ec];
    const double lam = Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne));
    const double mu = Ee / (2.0 * (1.0 + ne));
    const double tr = Eg[0] + Eg[1] + Eg[2];
    const double S[9] = {lam * tr + 2.0 * mu * Eg[0], 2.0 * mu * Eg[3], 2.0 * mu * Eg[4],
                         2.0 * mu * Eg[3], lam * tr + 2.0 * mu * Eg[1], 2.0 * mu * Eg[5],
                         2.0 * mu * Eg[4], 2.0 * mu * Eg[5], lam * tr + 2.0 * mu * Eg[2]};
    // residual of node t at point gp: det F S g_t, summed over the points
    double Sg[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) Sg[j] = S[3 * j] * gt[0] + S[3 * j + 1] * gt[1] + S[3 * j + 2] * gt[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double ri = stride8_sum(det * (F[3 * i] * Sg[0] + F[3 * i + 1] * Sg[1] + F[3 * i + 2] * Sg[2]));
        if (live && gp == 0) Re[24 * (int64_t)e + 3 * t + i] = ri;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) s_g[wv][gp][t][c] = gt[c];
    if (t == 0) {
        s_det[wv][gp] = det;
#pragma unroll
        for (int c = 0; c < 9; ++c) s_F[wv][gp][c] = F[c];
        const int pi[6] = {0, 1, 2, 0, 0, 1}, pk[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            const int i = pi[m], k = pk[m];
            s_FF[wv][gp][m] = F[3 * i] * F[3 * k] + F[3 * i + 1] * F[3 * k + 1] + F[3 * i + 2] * F[3 * k + 2];
        }
        s_S[wv][gp][0] = S[0];
        s_S[wv][gp][1] = S[4];
        s_S[wv][gp][2] = S[8];
        s_S[wv][gp][3] = S[1];
        s_S[wv][gp][4] = S[2];
        s_S[wv][gp][5] = S[5];
    }
    __syncthreads();

    // phase 2: block (a, b)
    const int a = lane >> 3, b = lane & 7;
    double kb[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) kb[c] = 0.0;
#pragma unroll 2
    for (int q = 0; q < 8; ++q) {
        const double ga[3] = {s_g[wv][q][a][0], s_g[wv][q][a][1], s_g[wv][q][a][2]};
        const double gb[3] = {s_g[wv][q][b][0], s_g[wv][q][b][1], s_g[wv][q][b][2]};
        const double* Fq = s_F[wv][q];
        const double* FFq = s_FF[wv][q];
        const double* Sq = s_S[wv][q];
        const double dq = s_det[wv][q];
        double Fa[3], Fb[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            Fa[i] = Fq[3 * i] * ga[0] + Fq[3 * i + 1] * ga[1] + Fq[3 * i + 2] * ga[2];
            Fb[i] = Fq[3 * i] * gb[0] + Fq[3 * i + 1] * gb[1] + Fq[3 * i + 2] * gb[2];
        }
        const double dgg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
        const double Sb[3] = {Sq[0] * gb[0] + Sq[3] * gb[1] + Sq[4] * gb[2],
                              Sq[3] * gb[0] + Sq[1] * gb[1] + Sq[5] * gb[2],
                              Sq[4] * gb[0] + Sq[5] * gb[1] + Sq[2] * gb[2]};
        const double gsg = ga[0] * Sb[0] + ga[1] * Sb[1] + ga[2] * Sb[2];
        const double FF[9] = {FFq[0], FFq[3], FFq[4], FFq[3], FFq[1], FFq[5], FFq[4], FFq[5], FFq[2]};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                kb[3 * i + k] += dq * (lam * Fa[i] * Fb[k] + mu * (FF[3 * i + k] * dgg + Fb[i] * Fa[k]) +
                                       (i == k ? gsg : 0.0));
    }
    if (live) {
        double* out = Ke + (int64_t)kRecStride * (8 * (int64_t)e + a) + 9 * b;
#pragma unroll
        for (int c = 0; c < 9; ++c) out[c] = kb[c];
    }
}

// Penalty part of the Newton load: rhs[d] -= mult sf u[d] on the penalty DOFs
// (build_elastic_const_load with q = 0, S/model_analysis.rs:275-287).
__global__ void k_nl_penalty_load(const int64_t* __restrict__ pen, const int32_t* __restrict__ mult, int n_pen,
                                  const Scalars* __restrict__ sc, const double* __restrict__ u,
                                  double* __restrict__ rhs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pen) return;
    const int64_t d = pen[i];
    rhs[d] -= mult[i] * sc->sf * u[d];
}

}  // namespace

void launch_hex8_nl_element(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                            const double* nu, const double* u, double* Ke, double* Re) {
    if (n_el <= 0) return;
    hipLaunchKernelGGL(k_hex8_nl_element, dim3((unsigned)((n_el + kNlWaves - 1) / kNlWaves)), dim3(64 * kNlWaves),
                       0, st, n_el, conn_r, xyz_r, E, nu, u, Ke, Re);
}

void launch_nl_penalty_load(hipStream_t st, const int64_t* pen, const int32_t* mult, int n_pen, const Scalars* sc,
                            const double* u, double* rhs) {
    if (n_pen <= 0) return;
    hipLaunchKernelGGL(k_nl_penalty_load, dim3((n_pen + 255) / 256), dim3(256), 0, st, pen, mult, n_pen, sc, u, rhs);
}

}  // namespace asend
