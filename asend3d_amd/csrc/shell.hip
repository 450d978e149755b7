// shell4 element integration and assembly on the device.
//
// Restates the reference's quad shell (type 41: 4 nodes x 6 DOF + 8 condensed
// internal DOF, S/element_meth.rs:253-297) for one isotropic layer of
// thickness t at the mid-plane:
//   shape functions: 4 bilinear + 6 bubbles (S/element_soln_fields.rs:194-243),
//     bubble derivatives with the centroid Jacobian (:419-443);
//   Jacobian from the nodal functions with J[2][2] = sign of the in-plane
//     determinant (:383-418);
//   section deformation [e11, e22, g12, k11, k22, k12, g13, g23, drill] of a
//     unit DOF (get_section_def_dfd0, :1160-1343);
//   ABD: A = Q t, B = 0, D = Q t^3/12, transverse-shear and drill terms = A66
//     (get_abd_dfd0, S/element_properties.rs:766-876);
//   K = sum_ip B^T C B detJ over 2x2 Gauss points, internal DOFs condensed
//     (condense_mat, S/element_equations.rs:19-57);
//   mass sum_ip b^T M6 b detJ with b the nodal interpolation of each DOF
//     (get_rum_dfd0, :514-581), M6 = diag(rho t x3, rho t^3/12 x3).
// The element frame is the identity (the reference's correct_orient leaves
// the section orientation untouched when the element normal is parallel to
// its z axis); the mesh upload rejects other elements.
//
// Each 6-DOF node is two 3x3 block rows of the operator (translations,
// rotations), so the element scatters 8 x 8 blocks exactly as the brick8
// kernel does and every solver kernel applies unchanged.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernels.hpp"

namespace asend {

namespace {

__device__ void shell_shape(double x, double y, double* N, double* dN /* 10 x 2 */) {
    N[0] = 0.25 * (1 - x) * (1 - y);
    N[1] = 0.25 * (1 + x) * (1 - y);
    N[2] = 0.25 * (1 + x) * (1 + y);
    N[3] = 0.25 * (1 - x) * (1 + y);
    N[4] = 1 - x * x;
    N[5] = 1 - y * y;
    N[6] = (1 - x * x) * (1 - y);
    N[7] = (1 - y * y) * (1 + x);
    N[8] = (1 - x * x) * (1 + y);
    N[9] = (1 - y * y) * (1 - x);
    const double d[20] = {-0.25 * (1 - y), -0.25 * (1 - x), 0.25 * (1 - y), -0.25 * (1 + x),
                          0.25 * (1 + y),  0.25 * (1 + x),  -0.25 * (1 + y), 0.25 * (1 - x),
                          -2 * x,          0.0,             0.0,            -2 * y,
                          -2 * x * (1 - y), -1 + x * x,     1 - y * y,      -2 * y * (1 + x),
                          -2 * x * (1 + y), 1 - x * x,      -1 + y * y,     -2 * y * (1 - x)};
    for (int k = 0; k < 20; ++k) dN[k] = d[k];
}

// inverse of the 3x3 Jacobian of the nodal functions at (x, y)
__device__ double shell_jinv(const double (*X)[3], double x, double y, double* Ji) {
    double N[10], dN[20];
    shell_shape(x, y, N, dN);
    double J[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 2; ++j) {
            double acc = 0.0;
            for (int a = 0; a < 4; ++a) acc += X[a][i] * dN[2 * a + j];
            J[3 * i + j] = acc;
        }
    J[2] = J[5] = 0.0;
    J[8] = (J[0] * J[4] - J[3] * J[1] > 0.0) ? 1.0 : -1.0;
    const double c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8], c02 = J[3] * J[7] - J[4] * J[6];
    const double det = J[0] * c00 + J[1] * c01 + J[2] * c02;
    const double id = 1.0 / det;
    Ji[0] = c00 * id;
    Ji[1] = (J[2] * J[7] - J[1] * J[8]) * id;
    Ji[2] = (J[1] * J[5] - J[2] * J[4]) * id;
    Ji[3] = c01 * id;
    Ji[4] = (J[0] * J[8] - J[2] * J[6]) * id;
    Ji[5] = (J[2] * J[3] - J[0] * J[5]) * id;
    Ji[6] = c02 * id;
    Ji[7] = (J[1] * J[6] - J[0] * J[7]) * id;
    Ji[8] = (J[0] * J[4] - J[1] * J[3]) * id;
    return det;
}

// One workgroup (64 lanes) per element of one colour.  Adds k_scale K_e +
// m_scale M_e into the operator.
__global__ __launch_bounds__(64) void k_shell4_assemble(const int32_t* __restrict__ elems, int n_elems,
                                                        const int32_t* __restrict__ conn_p,
                                                        const double* __restrict__ xyz_r, const double* __restrict__ E,
                                                        const double* __restrict__ nu, const double* __restrict__ thk,
                                                        const double* __restrict__ rho, double k_scale,
                                                        double m_scale, const int32_t* __restrict__ sptr,
                                                        const uint8_t* __restrict__ slot, double* __restrict__ K) {
    const int lane = threadIdx.x;
    const int ei = blockIdx.x;
    if (ei >= n_elems) return;
    const int e = elems[ei];
    __shared__ double X[4][3];
    __shared__ double dx[4][10][2];
    __shared__ double Nq[4][4];
    __shared__ double det[4];
    __shared__ double B[9][32];
    __shared__ double CB[9][32];
    __shared__ double Ks[32][33];
    __shared__ double Kii[8][16];
    __shared__ double Wm[8][24];
    if (lane < 12) {
        const int a = lane / 3, c = lane % 3;
        X[a][c] = xyz_r[3 * (int64_t)(conn_p[8 * (int64_t)e + 2 * a] >> 1) + c];
    }
    for (int t = lane; t < 32 * 33; t += 64) Ks[t / 33][t % 33] = 0.0;
    __syncthreads();
    const double g = 0.5773502691896257;
    if (lane < 4) {
        const double x = (lane & 1) ? g : -g, y = (lane & 2) ? g : -g;
        double N[10], dN[20], Ji[9], Jc[9];
        shell_shape(x, y, N, dN);
        det[lane] = shell_jinv(X, x, y, Ji);
        shell_jinv(X, 0.0, 0.0, Jc);
        for (int a = 0; a < 10; ++a) {
            const double* Jm = a < 4 ? Ji : Jc;
            dx[lane][a][0] = dN[2 * a] * Jm[0] + dN[2 * a + 1] * Jm[3];
            dx[lane][a][1] = dN[2 * a] * Jm[1] + dN[2 * a + 1] * Jm[4];
        }
        for (int a = 0; a < 4; ++a) Nq[lane][a] = N[a];
    }
    // material: Q of the isotropic layer, A = Q t, D = Q t^3 / 12, shear/drill A66
    const double Ee = E[e], ne = nu[e], te = thk[e];
    const double q11 = Ee / (1.0 - ne * ne), q12 = ne * q11, q66 = Ee / (2.0 * (1.0 + ne));
    const double a11 = q11 * te, a12 = q12 * te, a66 = q66 * te;
    const double t3 = te * te * te / 12.0;
    const double d11 = q11 * t3, d12 = q12 * t3, d66 = q66 * t3;
    __syncthreads();
    for (int q = 0; q < 4; ++q) {
        if (lane < 32) {
            double col[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            const int j = lane;
            if (j < 24) {
                const int a = j / 6, c = j % 6;
                const double Nx = dx[q][a][0], Ny = dx[q][a][1], N = Nq[q][a];
                if (c == 0) {
                    col[0] = Nx; col[2] = Ny; col[8] = Ny;
                } else if (c == 1) {
                    col[1] = Ny; col[2] = Nx; col[8] = -Nx;
                } else if (c == 2) {
                    col[6] = Nx; col[7] = Ny;
                } else if (c == 3) {
                    col[4] = -Ny; col[5] = -Nx; col[7] = -N;
                } else if (c == 4) {
                    col[3] = Nx; col[5] = Ny; col[6] = N;
                } else {
                    col[8] = 2.0 * N;
                }
            } else {
                // internal: (bubble, component) = (4,u) (5,u) (4,v) (5,v) (6,w) (7,w) (8,w) (9,w)
                const int k = j - 24;
                const int bb = (k < 4) ? 4 + (k & 1) : 2 + k, comp = (k < 4) ? (k >> 1) : 2;
                const double Nx = dx[q][bb][0], Ny = dx[q][bb][1];
                if (comp == 0) {
                    col[0] = Nx; col[2] = Ny; col[8] = Ny;
                } else if (comp == 1) {
                    col[1] = Ny; col[2] = Nx; col[8] = -Nx;
                } else {
                    col[6] = Nx; col[7] = Ny;
                }
            }
            for (int k = 0; k < 9; ++k) B[k][j] = col[k];
            const double w = det[q];
            CB[0][j] = w * (a11 * col[0] + a12 * col[1]);
            CB[1][j] = w * (a12 * col[0] + a11 * col[1]);
            CB[2][j] = w * a66 * col[2];
            CB[3][j] = w * (d11 * col[3] + d12 * col[4]);
            CB[4][j] = w * (d12 * col[3] + d11 * col[4]);
            CB[5][j] = w * d66 * col[5];
            CB[6][j] = w * a66 * col[6];
            CB[7][j] = w * a66 * col[7];
            CB[8][j] = w * a66 * col[8];
        }
        __syncthreads();
        for (int t = lane; t < 1024; t += 64) {
            const int i = t / 32, j = t % 32;
            double acc = 0.0;
            for (int k = 0; k < 9; ++k) acc += B[k][i] * CB[k][j];
            Ks[i][j] += acc;
        }
        __syncthreads();
    }
    // condensation of the 8 internal DOFs: Kc = Kee - Kei Kii^-1 Kie
    if (lane < 8)
        for (int j = 0; j < 8; ++j) {
            Kii[lane][j] = Ks[24 + lane][24 + j];
            Kii[lane][8 + j] = (lane == j) ? 1.0 : 0.0;
        }
    __syncthreads();
    if (lane == 0) {
        // Gauss-Jordan with partial pivoting on [Kii | I]
        for (int c = 0; c < 8; ++c) {
            int p = c;
            for (int r = c + 1; r < 8; ++r)
                if (fabs(Kii[r][c]) > fabs(Kii[p][c])) p = r;
            if (p != c)
                for (int k = 0; k < 16; ++k) {
                    const double tmp = Kii[c][k];
                    Kii[c][k] = Kii[p][k];
                    Kii[p][k] = tmp;
                }
            const double ip = 1.0 / Kii[c][c];
            for (int k = 0; k < 16; ++k) Kii[c][k] *= ip;
            for (int r = 0; r < 8; ++r) {
                if (r == c) continue;
                const double f = Kii[r][c];
                for (int k = 0; k < 16; ++k) Kii[r][k] -= f * Kii[c][k];
            }
        }
    }
    __syncthreads();
    for (int t = lane; t < 8 * 24; t += 64) {
        const int k = t / 24, j = t % 24;
        double acc = 0.0;
        for (int m = 0; m < 8; ++m) acc += Kii[k][8 + m] * Ks[24 + m][j];
        Wm[k][j] = acc;
    }
    __syncthreads();
    for (int t = lane; t < 24 * 24; t += 64) {
        const int i = t / 24, j = t % 24;
        double acc = 0.0;
        for (int k = 0; k < 8; ++k) acc += Ks[i][24 + k] * Wm[k][j];
        Ks[i][j] -= acc;
    }
    __syncthreads();
    // scatter: lane = (P, Q) pseudo-node pair, P = 2a + h
    const int P = lane / 8, Q = lane % 8;
    const int a = P >> 1, hp = P & 1, b = Q >> 1, hq = Q & 1;
    double sab = 0.0;
    for (int q = 0; q < 4; ++q) sab += Nq[q][a] * Nq[q][b] * det[q];
    const double re = rho ? rho[e] : 0.0;
    const double mt = re * te, mr = re * t3;
    const int rp = conn_p[8 * (int64_t)e + P];
    const int kk = slot[64 * (int64_t)e + lane];
    double* base = K + sell_index(sptr, rp, kk, 0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = k_scale * Ks[6 * a + 3 * hp + i][6 * b + 3 * hq + j];
            if (hp == hq && i == j) v += m_scale * (hp ? mr : mt) * sab;
            base[64 * (3 * i + j)] += v;
        }
}

}  // namespace

void launch_shell4_assemble(hipStream_t st, const int32_t* elems, int n_elems, const int32_t* conn_p,
                            const double* xyz_r, const double* E, const double* nu, const double* thk,
                            const double* rho, double k_scale, double m_scale, const int32_t* sptr,
                            const uint8_t* slot, double* K) {
    if (n_elems <= 0) return;
    hipLaunchKernelGGL(k_shell4_assemble, dim3(n_elems), dim3(64), 0, st, elems, n_elems, conn_p, xyz_r, E, nu, thk,
                       rho, k_scale, m_scale, sptr, slot, K);
}

}  // namespace asend
