// brickIM (type 81) element integration on the device: brick8 with three
// incompatible bubble modes N_{8+k} = 1 - s_k^2 per displacement component
// (S/element_soln_fields.rs:161-225), whose derivatives use the Jacobian at
// the element centroid (get_ip_data_dfd0, :383-463); the 9 internal DOFs are
// condensed out (condense_mat, S/element_equations.rs:19-57).
//
// One 64-lane workgroup per element.  The full 33 x 33 stiffness
//   K_AB,ij = sum_gp detJ [lam g_A,i g_B,j + mu g_A,j g_B,i + mu (g_A.g_B) d_ij]
// (A, B over the 8 nodes and 3 bubbles, the isotropic B^T C B of
// get_ruk_dfd0, :206-418) is built in LDS, condensed to 24 x 24 by a
// Gauss-Jordan inverse of the 9 x 9 internal block, and then, by MODE:
//   0: k_scale Kc + c_mass rho M added into the operator (one launch per
//      element colour, lane (a, b) owns block (a, b));
//   1: the E = 1, rho = 1 reference blocks Khat, Mhat of the element elems[0];
//   2: dJ/dE_e = -lam_e^T (Kc / E_e) u_e (Kc is linear in E).
// The consistent mass is the brick8 one on the 8 nodal functions.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernels.hpp"

namespace asend {

namespace {

__device__ __forceinline__ void im_sign(int t, double& sx, double& sy, double& sz) {
    sx = (((t ^ (t >> 1)) & 1) != 0) ? 1.0 : -1.0;
    sy = (t & 2) ? 1.0 : -1.0;
    sz = (t & 4) ? 1.0 : -1.0;
}

// 3x3 inverse and determinant of J (row-major, J[i][j] = dx_i/ds_j)
__device__ double inv3(const double* J, double* Ji) {
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

template <int MODE>
__global__ __launch_bounds__(64) void k_brickim(const int32_t* __restrict__ elems, int n_elems,
                                                const int32_t* __restrict__ conn_r, const double* __restrict__ xyz_r,
                                                const double* __restrict__ E, const double* __restrict__ nu,
                                                const double* __restrict__ rho, double k_scale, double c_mass,
                                                const int32_t* __restrict__ sptr, const uint8_t* __restrict__ slot,
                                                double* __restrict__ K, double* __restrict__ Khat,
                                                double* __restrict__ Mhat, const double* __restrict__ u,
                                                const double* __restrict__ lam, double* __restrict__ sens) {
    const int lane = threadIdx.x;
    const int ei = blockIdx.x;
    if (ei >= n_elems) return;
    const int e = elems ? elems[ei] : ei;
    __shared__ double X[8][3];
    __shared__ double Ji[8][9];
    __shared__ double W[8];
    __shared__ double Jci[9];
    __shared__ double Ns[8][8];
    __shared__ double G[8][11][3];
    __shared__ double Kf[33][34];
    __shared__ double Kii[9][18];
    __shared__ double Wm[9][24];
    __shared__ double red[64];
    if (lane < 24) X[lane / 3][lane % 3] = xyz_r[3 * (int64_t)conn_r[8 * (int64_t)e + lane / 3] + lane % 3];
    __syncthreads();
    const double g = 0.5773502691896257;
    if (lane < 9) {
        // lane < 8: Gauss point `lane`; lane 8: the centroid
        const double p0 = lane == 8 ? 0.0 : ((lane & 1) ? g : -g);
        const double p1 = lane == 8 ? 0.0 : ((lane & 2) ? g : -g);
        const double p2 = lane == 8 ? 0.0 : ((lane & 4) ? g : -g);
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int t = 0; t < 8; ++t) {
            double sx, sy, sz;
            im_sign(t, sx, sy, sz);
            const double fx = 1.0 + sx * p0, fy = 1.0 + sy * p1, fz = 1.0 + sz * p2;
            const double d[3] = {0.125 * sx * fy * fz, 0.125 * fx * sy * fz, 0.125 * fx * fy * sz};
            if (lane < 8) Ns[lane][t] = 0.125 * fx * fy * fz;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) J[3 * i + j] += X[t][i] * d[j];
        }
        double Jinv[9];
        const double det = inv3(J, Jinv);
        if (lane < 8) {
            for (int c = 0; c < 9; ++c) Ji[lane][c] = Jinv[c];
            W[lane] = det;  // Gauss weight 1
        } else {
            for (int c = 0; c < 9; ++c) Jci[c] = Jinv[c];
        }
    }
    __syncthreads();
    {
        // nodal functions: lane = (gp, a)
        const int gp = lane / 8, a = lane % 8;
        const double p0 = (gp & 1) ? g : -g, p1 = (gp & 2) ? g : -g, p2 = (gp & 4) ? g : -g;
        double sx, sy, sz;
        im_sign(a, sx, sy, sz);
        const double fx = 1.0 + sx * p0, fy = 1.0 + sy * p1, fz = 1.0 + sz * p2;
        const double d[3] = {0.125 * sx * fy * fz, 0.125 * fx * sy * fz, 0.125 * fx * fy * sz};
        for (int c = 0; c < 3; ++c) G[gp][a][c] = d[0] * Ji[gp][c] + d[1] * Ji[gp][3 + c] + d[2] * Ji[gp][6 + c];
        // bubbles: lane < 24 = (gp, k); dN_{8+k}/ds_j = -2 s_k d_jk, centroid Jacobian
        if (lane < 24) {
            const int gq = lane / 3, k = lane % 3;
            const double sk = ((gq >> k) & 1) ? g : -g;
            for (int c = 0; c < 3; ++c) G[gq][8 + k][c] = -2.0 * sk * Jci[3 * k + c];
        }
    }
    __syncthreads();
    const double Ee = (MODE == 1) ? 1.0 : E[e];
    const double ne = nu[e];
    const double lm = Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne));
    const double mu = Ee / (2.0 * (1.0 + ne));
    for (int t = lane; t < 121; t += 64) {
        const int A = t / 11, B = t % 11;
        double k[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int gp = 0; gp < 8; ++gp) {
            const double* ga = G[gp][A];
            const double* gb = G[gp][B];
            const double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    k[3 * i + j] += W[gp] * (lm * ga[i] * gb[j] + mu * ga[j] * gb[i] + (i == j ? mu * dot : 0.0));
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Kf[3 * A + i][3 * B + j] = k[3 * i + j];
    }
    __syncthreads();
    // condensation Kc = Kee - Kei Kii^-1 Kie
    for (int t = lane; t < 81; t += 64) {
        const int i = t / 9, j = t % 9;
        Kii[i][j] = Kf[24 + i][24 + j];
        Kii[i][9 + j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (lane == 0) {
        // SPD block: Gauss-Jordan without pivoting
        for (int c = 0; c < 9; ++c) {
            const double ip = 1.0 / Kii[c][c];
            for (int k = 0; k < 18; ++k) Kii[c][k] *= ip;
            for (int r = 0; r < 9; ++r) {
                if (r == c) continue;
                const double f = Kii[r][c];
                for (int k = 0; k < 18; ++k) Kii[r][k] -= f * Kii[c][k];
            }
        }
    }
    __syncthreads();
    for (int t = lane; t < 9 * 24; t += 64) {
        const int k = t / 24, j = t % 24;
        double acc = 0.0;
        for (int m = 0; m < 9; ++m) acc += Kii[k][9 + m] * Kf[24 + m][j];
        Wm[k][j] = acc;
    }
    __syncthreads();
    for (int t = lane; t < 24 * 24; t += 64) {
        const int i = t / 24, j = t % 24;
        double acc = 0.0;
        for (int k = 0; k < 9; ++k) acc += Kf[i][24 + k] * Wm[k][j];
        Kf[i][j] -= acc;
    }
    __syncthreads();
    const int a = lane / 8, b = lane % 8;
    if (MODE == 2) {
        const int ra = conn_r[8 * (int64_t)e + a], rb = conn_r[8 * (int64_t)e + b];
        double s = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) s += lam[3 * (int64_t)ra + i] * Kf[3 * a + i][3 * b + j] * u[3 * (int64_t)rb + j];
        red[lane] = s;
        __syncthreads();
        if (lane == 0) {
            double tot = 0.0;
            for (int l = 0; l < 64; ++l) tot += red[l];
            sens[e] = -tot / Ee;
        }
        return;
    }
    double m = 0.0;
    for (int gp = 0; gp < 8; ++gp) m += W[gp] * Ns[gp][a] * Ns[gp][b];
    if (MODE == 1) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Khat[9 * lane + 3 * i + j] = Kf[3 * a + i][3 * b + j];
        Mhat[lane] = m;
        return;
    }
    const double cm = (c_mass != 0.0) ? c_mass * rho[e] * m : 0.0;
    const int ra = conn_r[8 * (int64_t)e + a];
    const int kk = slot[64 * (int64_t)e + lane];
    double* base = K + sell_index(sptr, ra, kk, 0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) base[64 * (3 * i + j)] += k_scale * Kf[3 * a + i][3 * b + j] + (i == j ? cm : 0.0);
}

// Geometrically nonlinear brickIM (total Lagrangian, St. Venant-Kirchhoff):
// the brick8 kinematics of k_hex8_nl_element over the 8 nodes and the 3
// bubbles, whose 9 amplitudes alpha (per element, 3 k + component) enter
// the displacement gradient like nodal values (get_inst_disp_dfd0,
// S/element_soln_fields.rs:718-748).  One 64-lane workgroup per element:
// the 33-entry residual r = sum_gp detJ F S g_A and the 33 x 33 tangent in
// LDS, then the condensation of update_external / condense_mat
// (S/element_equations.rs:19-121): K_c = K_ee - K_ei K_ii^-1 K_ie and
// r_c = r_e - K_ei K_ii^-1 r_i, written as the brick8 row records and
// residuals of the shared gather, and the maps W = K_ii^-1 K_ie,
// w = K_ii^-1 r_i of the internal update (225 doubles per element).
__global__ __launch_bounds__(64) void k_brickim_nl(int n_el, const int32_t* __restrict__ conn_r,
                                                   const double* __restrict__ xyz_r, const double* __restrict__ E,
                                                   const double* __restrict__ nu, const double* __restrict__ u,
                                                   const double* __restrict__ alpha, double* __restrict__ Ke,
                                                   double* __restrict__ Re, double* __restrict__ Wmap) {
    const int lane = threadIdx.x;
    const int e = blockIdx.x;
    if (e >= n_el) return;
    __shared__ double X[8][3];
    __shared__ double U[11][3];
    __shared__ double Ji[8][9];
    __shared__ double W[8];
    __shared__ double Jci[9];
    __shared__ double G[8][11][3];
    __shared__ double sF[8][9], sS[8][9], sFF[8][9], sP[8][9];
    __shared__ double rv[33];
    __shared__ double Kf[33][34];
    __shared__ double Kii[9][18];
    __shared__ double Wm[9][25];
    if (lane < 24) {
        const int64_t row = conn_r[8 * (int64_t)e + lane / 3];
        X[lane / 3][lane % 3] = xyz_r[3 * row + lane % 3];
        U[lane / 3][lane % 3] = u[3 * row + lane % 3];
    } else if (lane < 33) {
        const int k = lane - 24;
        U[8 + k / 3][k % 3] = alpha[9 * (int64_t)e + k];
    }
    __syncthreads();
    const double g = 0.5773502691896257;
    if (lane < 9) {
        const double p0 = lane == 8 ? 0.0 : ((lane & 1) ? g : -g);
        const double p1 = lane == 8 ? 0.0 : ((lane & 2) ? g : -g);
        const double p2 = lane == 8 ? 0.0 : ((lane & 4) ? g : -g);
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int t = 0; t < 8; ++t) {
            double sx, sy, sz;
            im_sign(t, sx, sy, sz);
            const double fx = 1.0 + sx * p0, fy = 1.0 + sy * p1, fz = 1.0 + sz * p2;
            const double d[3] = {0.125 * sx * fy * fz, 0.125 * fx * sy * fz, 0.125 * fx * fy * sz};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) J[3 * i + j] += X[t][i] * d[j];
        }
        double Jinv[9];
        const double det = inv3(J, Jinv);
        if (lane < 8) {
            for (int c = 0; c < 9; ++c) Ji[lane][c] = Jinv[c];
            W[lane] = det;
        } else {
            for (int c = 0; c < 9; ++c) Jci[c] = Jinv[c];
        }
    }
    __syncthreads();
    {
        const int gp = lane / 8, a = lane % 8;
        const double p0 = (gp & 1) ? g : -g, p1 = (gp & 2) ? g : -g, p2 = (gp & 4) ? g : -g;
        double sx, sy, sz;
        im_sign(a, sx, sy, sz);
        const double fx = 1.0 + sx * p0, fy = 1.0 + sy * p1, fz = 1.0 + sz * p2;
        const double d[3] = {0.125 * sx * fy * fz, 0.125 * fx * sy * fz, 0.125 * fx * fy * sz};
        for (int c = 0; c < 3; ++c) G[gp][a][c] = d[0] * Ji[gp][c] + d[1] * Ji[gp][3 + c] + d[2] * Ji[gp][6 + c];
        if (lane < 24) {
            const int gq = lane / 3, k = lane % 3;
            const double sk = ((gq >> k) & 1) ? g : -g;
            for (int c = 0; c < 3; ++c) G[gq][8 + k][c] = -2.0 * sk * Jci[3 * k + c];
        }
    }
    __syncthreads();
    const double Ee = E[e], ne = nu[e];
    const double lm = Ee * ne / ((1.0 + ne) * (1.0 - 2.0 * ne));
    const double mu = Ee / (2.0 * (1.0 + ne));
    if (lane < 8) {
        // kinematics of Gauss point `lane`: F = I + sum_A U_A g_A^T, E, S, F S, F F^T
        const int q = lane;
        double F[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double h = 0.0;
                for (int A = 0; A < 11; ++A) h += U[A][i] * G[q][A][j];
                F[3 * i + j] = h + (i == j ? 1.0 : 0.0);
            }
        double Eg[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double c = 0.0;
                for (int k = 0; k < 3; ++k) c += F[3 * k + i] * F[3 * k + j];
                Eg[3 * i + j] = 0.5 * (c - (i == j ? 1.0 : 0.0));
            }
        const double tr = Eg[0] + Eg[4] + Eg[8];
        for (int c = 0; c < 9; ++c) {
            sS[q][c] = 2.0 * mu * Eg[c] + ((c % 4 == 0) ? lm * tr : 0.0);
            sF[q][c] = F[c];
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double p = 0.0, ff = 0.0;
                for (int k = 0; k < 3; ++k) {
                    p += F[3 * i + k] * sS[q][3 * k + j];
                    ff += F[3 * i + k] * F[3 * j + k];
                }
                sP[q][3 * i + j] = p;
                sFF[q][3 * i + j] = ff;
            }
    }
    __syncthreads();
    if (lane < 33) {
        const int A = lane / 3, i = lane % 3;
        double r = 0.0;
        for (int q = 0; q < 8; ++q)
            r += W[q] * (sP[q][3 * i] * G[q][A][0] + sP[q][3 * i + 1] * G[q][A][1] + sP[q][3 * i + 2] * G[q][A][2]);
        rv[lane] = r;
    }
    for (int t = lane; t < 121; t += 64) {
        const int A = t / 11, B = t % 11;
        double kb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 8; ++q) {
            const double* ga = G[q][A];
            const double* gb = G[q][B];
            const double* Fq = sF[q];
            const double* Sq = sS[q];
            const double* FF = sFF[q];
            double Fa[3], Fb[3];
            for (int i = 0; i < 3; ++i) {
                Fa[i] = Fq[3 * i] * ga[0] + Fq[3 * i + 1] * ga[1] + Fq[3 * i + 2] * ga[2];
                Fb[i] = Fq[3 * i] * gb[0] + Fq[3 * i + 1] * gb[1] + Fq[3 * i + 2] * gb[2];
            }
            const double dgg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
            double gsg = 0.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) gsg += ga[i] * Sq[3 * i + j] * gb[j];
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k)
                    kb[3 * i + k] += W[q] * (lm * Fa[i] * Fb[k] + mu * (FF[3 * i + k] * dgg + Fb[i] * Fa[k]) +
                                             (i == k ? gsg : 0.0));
        }
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) Kf[3 * A + i][3 * B + k] = kb[3 * i + k];
    }
    __syncthreads();
    for (int t = lane; t < 81; t += 64) {
        const int i = t / 9, j = t % 9;
        Kii[i][j] = Kf[24 + i][24 + j];
        Kii[i][9 + j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (lane == 0) {
        for (int c = 0; c < 9; ++c) {
            const double ip = 1.0 / Kii[c][c];
            for (int k = 0; k < 18; ++k) Kii[c][k] *= ip;
            for (int r = 0; r < 9; ++r) {
                if (r == c) continue;
                const double f = Kii[r][c];
                for (int k = 0; k < 18; ++k) Kii[r][k] -= f * Kii[c][k];
            }
        }
    }
    __syncthreads();
    // W = K_ii^-1 K_ie (columns 0..23), w = K_ii^-1 r_i (column 24)
    for (int t = lane; t < 9 * 25; t += 64) {
        const int k = t / 25, j = t % 25;
        double acc = 0.0;
        for (int m = 0; m < 9; ++m) acc += Kii[k][9 + m] * (j < 24 ? Kf[24 + m][j] : rv[24 + m]);
        Wm[k][j] = acc;
        Wmap[225 * (int64_t)e + t] = acc;
    }
    __syncthreads();
    for (int t = lane; t < 24 * 25; t += 64) {
        const int i = t / 25, j = t % 25;
        double acc = 0.0;
        for (int k = 0; k < 9; ++k) acc += Kf[i][24 + k] * Wm[k][j];
        if (j < 24) {
            const int a = i / 3, ii = i % 3, b = j / 3, kk = j % 3;
            Ke[(int64_t)kRecStride * (8 * (int64_t)e + a) + 9 * b + 3 * ii + kk] = Kf[i][j] - acc;
        } else {
            Re[24 * (int64_t)e + i] = rv[i] - acc;
        }
    }
}

// alpha_e -= w_e + W_e du_e (update_internal, S/element_equations.rs:123-194)
__global__ void k_brickim_nl_internal(int n_el, const int32_t* __restrict__ conn_r, const double* __restrict__ du,
                                      const double* __restrict__ Wmap, double* __restrict__ alpha) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 9 * (int64_t)n_el) return;
    const int64_t e = t / 9;
    const int k = (int)(t % 9);
    const double* Wk = Wmap + 225 * e + 25 * k;
    double acc = Wk[24];
    for (int j = 0; j < 24; ++j) acc += Wk[j] * du[3 * (int64_t)conn_r[8 * e + j / 3] + j % 3];
    alpha[t] -= acc;
}

}  // namespace

void launch_brickim_assemble(hipStream_t st, const int32_t* elems, int n_elems, const int32_t* conn_r,
                             const double* xyz_r, const double* E, const double* nu, const double* rho,
                             double k_scale, double c_mass, const int32_t* sptr, const uint8_t* slot, double* K) {
    if (n_elems <= 0) return;
    hipLaunchKernelGGL(k_brickim<0>, dim3(n_elems), dim3(64), 0, st, elems, n_elems, conn_r, xyz_r, E, nu, rho,
                       k_scale, c_mass, sptr, slot, K, nullptr, nullptr, nullptr, nullptr, nullptr);
}

void launch_brickim_reference(hipStream_t st, const int32_t* conn_r, const double* xyz_r, const double* nu,
                              double* Khat, double* Mhat) {
    hipLaunchKernelGGL(k_brickim<1>, dim3(1), dim3(64), 0, st, nullptr, 1, conn_r, xyz_r, nullptr, nu, nullptr, 1.0,
                       0.0, nullptr, nullptr, nullptr, Khat, Mhat, nullptr, nullptr, nullptr);
}

void launch_brickim_sens(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                         const double* nu, const double* u, const double* lam, double* sens) {
    if (n_el <= 0) return;
    hipLaunchKernelGGL(k_brickim<2>, dim3(n_el), dim3(64), 0, st, nullptr, n_el, conn_r, xyz_r, E, nu, nullptr, 1.0,
                       0.0, nullptr, nullptr, nullptr, nullptr, nullptr, u, lam, sens);
}

void launch_brickim_nl(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                       const double* nu, const double* u, const double* alpha, double* Ke, double* Re, double* Wmap) {
    if (n_el <= 0) return;
    hipLaunchKernelGGL(k_brickim_nl, dim3(n_el), dim3(64), 0, st, n_el, conn_r, xyz_r, E, nu, u, alpha, Ke, Re, Wmap);
}

void launch_brickim_nl_internal(hipStream_t st, int n_el, const int32_t* conn_r, const double* du,
                                const double* Wmap, double* alpha) {
    if (n_el <= 0) return;
    const int64_t n = 9 * (int64_t)n_el;
    hipLaunchKernelGGL(k_brickim_nl_internal, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n_el, conn_r, du,
                       Wmap, alpha);
}

}  // namespace asend
