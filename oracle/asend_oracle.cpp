/*
 * asend_oracle.cpp -- CPU FP64 restatement of the ASenD3D solve path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library; the product never does.
 *
 * Written from the mathematics of the reference routines cited at each
 * function (paths relative to the reference checkout, S/ =
 * solverSrc/asend_run_job/src/).  Serial FP64, as the reference runs.
 */
#include "asend_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

const double kTol = 1e-12;                    /* S/constants.rs:12 */
const double kR1ort3 = 0.5773502691896257;    /* S/constants.rs:2  */
const double kPi180 = 0.017453292519943295;   /* S/constants.rs:8  */

/* ---------------------------------------------------------------------------
 * Forward dual number: value + one tangent, the arithmetic of DiffDoub1
 * (S/diff_doub_meth.rs:82-185).  Plain double is the DiffDoub0 case.
 * ------------------------------------------------------------------------- */
struct Dual {
    double v, d;
    Dual() : v(0), d(0) {}
    Dual(double a) : v(a), d(0) {}
    Dual(double a, double b) : v(a), d(b) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.v * b.d + a.d * b.v); }
inline Dual operator/(Dual a, Dual b) {
    return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
inline Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
inline Dual& operator*=(Dual& a, Dual b) { a = a * b; return a; }
inline Dual datan(Dual a) { return Dual(std::atan(a.v), a.d / (1.0 + a.v * a.v)); }
inline Dual dsin(Dual a) { return Dual(std::sin(a.v), a.d * std::cos(a.v)); }
inline Dual dcos(Dual a) { return Dual(std::cos(a.v), -a.d * std::sin(a.v)); }
inline double dabs(Dual a) { return std::fabs(a.v); }
inline double datan(double a) { return std::atan(a); }
inline double dsin(double a) { return std::sin(a); }
inline double dcos(double a) { return std::cos(a); }
inline double dabs(double a) { return std::fabs(a); }
inline double val(double a) { return a; }
inline double val(Dual a) { return a.v; }

/* ---------------------------------------------------------------------------
 * Givens-rotation QR on a row range / column range of a row-major matrix
 * with `ncol` columns, and the matching solve.  Restates q_rfactor and
 * solveq_rx_eqb (S/matrix_functions.rs:167-283): the rotation angle
 * atan(a_rc / a_pc) is stored in the zeroed position; a pivot smaller than
 * 1e-12 in magnitude is replaced by 1e-12 before each rotation.
 * ------------------------------------------------------------------------- */
template <class T>
void qr_factor(T* m, int ncol, int r0, int r1, int c0, int c1) {
    for (int c = c0; c <= c1; ++c) {
        int p = r0 + (c - c0);
        for (int r = p + 1; r <= r1; ++r) {
            T& piv = m[p * ncol + c];
            if (dabs(piv) < kTol) piv = T(kTol);
            T th = datan(m[r * ncol + c] / piv);
            T s = dsin(th), co = dcos(th);
            for (int k = c; k <= c1; ++k) {
                T a = m[p * ncol + k], b = m[r * ncol + k];
                m[p * ncol + k] = co * a + s * b;
                m[r * ncol + k] = -s * a + co * b;
            }
            m[r * ncol + c] = th;
        }
    }
}

/* b is indexed by row (length >= r1+1) and is overwritten; x by column. */
template <class T>
void qr_solve(T* x, const T* m, T* b, int ncol, int r0, int r1, int c0, int c1) {
    for (int c = c0; c <= c1; ++c) {
        int p = r0 + (c - c0);
        for (int r = p + 1; r <= r1; ++r) {
            T th = m[r * ncol + c];
            T s = dsin(th), co = dcos(th);
            T a = b[p], bb = b[r];
            b[p] = co * a + s * bb;
            b[r] = -s * a + co * bb;
        }
        x[c] = T(0.0);
    }
    for (int c = c1; c >= c0; --c) {
        int p = r0 + (c - c0);
        T sum = T(0.0);
        for (int k = c + 1; k <= c1; ++k) sum += m[p * ncol + k] * x[k];
        x[c] = (b[p] - sum) / m[p * ncol + c];
    }
}

/* Determinant and inverse through the QR above: get_det_inv_ar_dfd0
 * (S/matrix_functions.rs:1707-1729).  `a` is destroyed. */
template <class T>
void det_inv(T& det, T* inv, T* a, int n) {
    qr_factor(a, n, 0, n - 1, 0, n - 1);
    T x[6], b[6];
    det = T(1.0);
    for (int c = 0; c < n; ++c) {
        for (int r = 0; r < n; ++r) b[r] = T(r == c ? 1.0 : 0.0);
        qr_solve(x, a, b, n, 0, n - 1, 0, n - 1);
        for (int r = 0; r < n; ++r) inv[r * n + c] = x[r];
        det *= a[c * n + c];
    }
}

/* ---------------------------------------------------------------------------
 * Orthotropic compliance -> stiffness (S/element_properties.rs:700-760).
 * Strain order [e11, e22, e33, g12, g13, g23].
 * ------------------------------------------------------------------------- */
template <class T>
void solid_cmat(const T* mod, const T* pr, const T* sm, T* C) {
    T s[36];
    for (int i = 0; i < 36; ++i) s[i] = T(0.0);
    s[0] = T(1.0) / mod[0];
    s[1] = -pr[0] / mod[0];
    s[2] = -pr[1] / mod[0];
    s[6] = s[1];
    s[7] = T(1.0) / mod[1];
    s[8] = -pr[2] / mod[1];
    s[12] = s[2];
    s[13] = s[8];
    s[14] = T(1.0) / mod[2];
    s[21] = T(1.0) / sm[0];
    s[28] = T(1.0) / sm[1];
    s[35] = T(1.0) / sm[2];
    T det;
    det_inv(det, C, s, 6);
}

/* ---------------------------------------------------------------------------
 * brick8 / brickIM shape functions (S/element_soln_fields.rs:161-225) and
 * Gauss points (S/element_meth.rs:157-227).
 * ------------------------------------------------------------------------- */
const double kHexNd[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                             {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

void hex_shape(int type, const double* s, double* N, double* dN /* n_dim x 3 */) {
    for (int a = 0; a < 8; ++a) {
        double fx = 1.0 + kHexNd[a][0] * s[0];
        double fy = 1.0 + kHexNd[a][1] * s[1];
        double fz = 1.0 + kHexNd[a][2] * s[2];
        N[a] = 0.125 * fx * fy * fz;
        dN[3 * a + 0] = 0.125 * kHexNd[a][0] * fy * fz;
        dN[3 * a + 1] = 0.125 * fx * kHexNd[a][1] * fz;
        dN[3 * a + 2] = 0.125 * fx * fy * kHexNd[a][2];
    }
    if (type == 81) {
        for (int k = 0; k < 3; ++k) {
            N[8 + k] = 1.0 - s[k] * s[k];
            for (int j = 0; j < 3; ++j) dN[3 * (8 + k) + j] = (j == k) ? -2.0 * s[k] : 0.0;
        }
    }
}

void hex_gauss(int gp, double* s) {
    /* x fastest, then y, then z (S/element_meth.rs:175-191) */
    s[0] = (gp & 1) ? kR1ort3 : -kR1ort3;
    s[1] = (gp & 2) ? kR1ort3 : -kR1ort3;
    s[2] = (gp & 4) ? kR1ort3 : -kR1ort3;
}

/* Integration-point data: Jacobian J = X dN, det/inverse by QR, physical
 * derivatives dNdx = dN J^-1, bubbles with the centroid Jacobian
 * (get_ip_data_dfd0, S/element_soln_fields.rs:383-463). */
void hex_ip_data(int type, const double* locX /* 3 x 8 SoA */, const double* s,
                 double* N, double* dNdx, double& detJ) {
    int ndim = (type == 81) ? 11 : 8;
    double dN[33];
    hex_shape(type, s, N, dN);
    double J[9], Ji[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int a = 0; a < 8; ++a) acc += locX[8 * i + a] * dN[3 * a + j];
            J[3 * i + j] = acc;
        }
    det_inv(detJ, Ji, J, 3);
    for (int a = 0; a < 8; ++a)
        for (int k = 0; k < 3; ++k) {
            double acc = 0.0;
            for (int j = 0; j < 3; ++j) acc += dN[3 * a + j] * Ji[3 * j + k];
            dNdx[3 * a + k] = acc;
        }
    if (ndim > 8) {
        double sc[3] = {0, 0, 0}, Nc[11], dNc[33], Jc[9], Jci[9], detc;
        hex_shape(type, sc, Nc, dNc);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double acc = 0.0;
                for (int a = 0; a < 8; ++a) acc += locX[8 * i + a] * dNc[3 * a + j];
                Jc[3 * i + j] = acc;
            }
        det_inv(detc, Jci, Jc, 3);
        for (int a = 8; a < ndim; ++a)
            for (int k = 0; k < 3; ++k) {
                double acc = 0.0;
                for (int j = 0; j < 3; ++j) acc += dN[3 * a + j] * Jci[3 * j + k];
                dNdx[3 * a + k] = acc;
            }
    }
}

/* Strain-displacement column for element DOF (nd, dof): get_solid_strain_dfd0
 * single-DOF branch (S/element_soln_fields.rs:1386-1419) with ori = local
 * orientation R (row-major), linear geometry. */
void strain_column(const double* R, const double* dNdx, int nd, int dof, double* b) {
    const double* g = dNdx + 3 * nd;
    b[0] = R[0 * 3 + dof] * g[0];
    b[1] = R[1 * 3 + dof] * g[1];
    b[2] = R[2 * 3 + dof] * g[2];
    b[3] = R[0 * 3 + dof] * g[1] + R[1 * 3 + dof] * g[0];
    b[4] = R[0 * 3 + dof] * g[2] + R[2 * 3 + dof] * g[0];
    b[5] = R[1 * 3 + dof] * g[2] + R[2 * 3 + dof] * g[1];
}

/* Full (uncondensed) element stiffness tot x tot and residual for brick8 /
 * brickIM (get_ruk_dfd0, S/element_equations.rs:206-418).  The material
 * matrix may carry a tangent (T = Dual); geometry is plain FP64.
 * u_full: tot DOFs, DOF k = (node k/3, comp k%3), bubbles after the 24 nodal. */
template <class T>
void hex_full(int type, const double* X, const double* ori, const T* C, const double* u_full,
              T* K, T* Rv) {
    int nn = 8, ndim = (type == 81) ? 11 : 8, tot = 3 * ndim;
    double Rm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (ori) std::memcpy(Rm, ori, sizeof(Rm));
    double locX[24];
    for (int i = 0; i < 3; ++i)
        for (int a = 0; a < nn; ++a) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += Rm[3 * i + k] * X[8 * k + a];
            locX[8 * i + a] = acc;
        }
    for (int i = 0; i < tot * tot; ++i) K[i] = T(0.0);
    for (int i = 0; i < tot; ++i) Rv[i] = T(0.0);
    std::vector<double> B(6 * tot);
    for (int gp = 0; gp < 8; ++gp) {
        double s[3], N[11], dNdx[33], detJ;
        hex_gauss(gp, s);
        hex_ip_data(type, locX, s, N, dNdx, detJ);
        double w = detJ * 1.0;
        for (int k = 0; k < tot; ++k) {
            double col[6];
            strain_column(Rm, dNdx, k / 3, k % 3, col);
            for (int r = 0; r < 6; ++r) B[r * tot + k] = col[r];
        }
        /* strain of the current displacement, stress = C strain */
        double eps[6] = {0, 0, 0, 0, 0, 0};
        if (u_full)
            for (int k = 0; k < tot; ++k)
                for (int r = 0; r < 6; ++r) eps[r] += B[r * tot + k] * u_full[k];
        T sig[6];
        for (int r = 0; r < 6; ++r) {
            T acc = T(0.0);
            for (int c = 0; c < 6; ++c) acc += C[6 * r + c] * T(eps[c]);
            sig[r] = acc;
        }
        for (int k = 0; k < tot; ++k) {
            T acc = T(0.0);
            for (int r = 0; r < 6; ++r) acc += sig[r] * T(B[r * tot + k]);
            Rv[k] += acc * T(w);
        }
        /* K += B^T (C B) w */
        for (int j = 0; j < tot; ++j) {
            T cb[6];
            for (int r = 0; r < 6; ++r) {
                T acc = T(0.0);
                for (int c = 0; c < 6; ++c) acc += C[6 * r + c] * T(B[c * tot + j]);
                cb[r] = acc * T(w);
            }
            for (int i = 0; i < tot; ++i) {
                T acc = T(0.0);
                for (int r = 0; r < 6; ++r) acc += T(B[r * tot + i]) * cb[r];
                K[i * tot + j] += acc;
            }
        }
    }
}

/* Static condensation of the internal block (condense_mat,
 * S/element_equations.rs:19-57): K_ee - K_ei K_ii^-1 K_ie, using the Givens QR
 * of K_ii stored in the (tot x nint) column block. */
template <class T>
void condense(const T* K, int tot, int next, T* Kc) {
    int nint = tot - next;
    std::vector<T> im(tot * nint);
    for (int i = 0; i < tot; ++i)
        for (int k = 0; k < nint; ++k) im[i * nint + k] = K[i * tot + next + k];
    qr_factor(im.data(), nint, next, tot - 1, 0, nint - 1);
    for (int i = 0; i < next; ++i)
        for (int j = 0; j < next; ++j) Kc[i * next + j] = K[i * tot + j];
    std::vector<T> b(tot), x(nint);
    for (int j = 0; j < next; ++j) {
        for (int i = 0; i < next; ++i) b[i] = T(0.0);
        for (int k = 0; k < nint; ++k) b[next + k] = im[j * nint + k];
        qr_solve(x.data(), im.data(), b.data(), nint, next, tot - 1, 0, nint - 1);
        for (int i = 0; i < next; ++i) {
            T acc = T(0.0);
            for (int k = 0; k < nint; ++k) acc += im[i * nint + k] * x[k];
            Kc[i * next + j] -= acc;
        }
    }
}

/* Internal DOFs that follow nodal values v_e with zero internal residual:
 * v_i = -K_ii^-1 K_ie v_e (update_internal, S/element_equations.rs:123-186). */
template <class T>
void internal_from_external(const T* K, int tot, int next, const double* ve, double* vi) {
    int nint = tot - next;
    std::vector<double> im(tot * nint);
    for (int i = 0; i < tot; ++i)
        for (int k = 0; k < nint; ++k) im[i * nint + k] = val(K[i * tot + next + k]);
    qr_factor(im.data(), nint, next, tot - 1, 0, nint - 1);
    std::vector<double> b(tot, 0.0), x(nint);
    for (int k = 0; k < nint; ++k) {
        double acc = 0.0;
        for (int i = 0; i < next; ++i) acc += im[i * nint + k] * ve[i];
        b[next + k] = -acc;
    }
    qr_solve(x.data(), im.data(), b.data(), nint, next, tot - 1, 0, nint - 1);
    for (int k = 0; k < nint; ++k) vi[k] = x[k];
}

void iso_props(double E, double nu, double* mod, double* pr, double* sm) {
    double G = E / (2.0 * (1.0 + nu));
    for (int i = 0; i < 3; ++i) { mod[i] = E; pr[i] = nu; sm[i] = G; }
}

void gather_X(const double* xyz, const int32_t* c, int nn, double* X) {
    for (int a = 0; a < nn; ++a)
        for (int k = 0; k < 3; ++k) X[nn * k + a] = xyz[3 * (int64_t)c[a] + k];
}

/* Element matrix (condensed for 81) for isotropic E, nu. */
void solid_ke_iso(int type, const double* X, double E, double nu, double* Ke) {
    double mod[3], pr[3], sm[3], C[36];
    iso_props(E, nu, mod, pr, sm);
    solid_cmat(mod, pr, sm, C);
    int tot = (type == 81) ? 33 : 24;
    std::vector<double> K(tot * tot), R(tot);
    hex_full<double>(type, X, nullptr, C, nullptr, K.data(), R.data());
    if (type == 81) condense(K.data(), tot, 24, Ke);
    else std::memcpy(Ke, K.data(), sizeof(double) * 576);
}


/* ---- shell4 (type 41) --------------------------------------------------
 * Element data (S/element_meth.rs:253-297): 4 nodes x 6 DOF + 8 internal DOF
 * (n_dim 10: 4 bilinear functions and 6 bubbles, S/element_soln_fields.rs:194-243),
 * 2x2 Gauss points (+-1/sqrt3, weight 1), section deformation
 * [e11, e22, g12, k11, k22, k12, g13, g23, drill]. */
void shell4_shape(const double* s, double* N, double* dN /* 10 x 3 */) {
    const double x = s[0], y = s[1];
    const double n[10] = {0.25 * (1 - x) * (1 - y), 0.25 * (1 + x) * (1 - y), 0.25 * (1 + x) * (1 + y),
                          0.25 * (1 - x) * (1 + y), 1 - x * x, 1 - y * y, (1 - x * x) * (1 - y),
                          (1 - y * y) * (1 + x), (1 - x * x) * (1 + y), (1 - y * y) * (1 - x)};
    const double d[10][2] = {{-0.25 * (1 - y), -0.25 * (1 - x)}, {0.25 * (1 - y), -0.25 * (1 + x)},
                             {0.25 * (1 + y), 0.25 * (1 + x)},   {-0.25 * (1 + y), 0.25 * (1 - x)},
                             {-2 * x, 0.0},                       {0.0, -2 * y},
                             {-2 * x * (1 - y), -1 + x * x},      {1 - y * y, -2 * y * (1 + x)},
                             {-2 * x * (1 + y), 1 - x * x},       {-1 + y * y, -2 * y * (1 - x)}};
    for (int a = 0; a < 10; ++a) {
        N[a] = n[a];
        dN[3 * a] = d[a][0];
        dN[3 * a + 1] = d[a][1];
        dN[3 * a + 2] = 0.0;
    }
}

/* Jacobian from the 4 nodal functions with J[2][2] = sign of the in-plane
 * determinant (get_ip_data_dfd0, S/element_soln_fields.rs:383-445). */
void shell4_jac(const double* loc /* 3 x 4 */, const double* dN, double* det, double* Ji) {
    double J[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int a = 0; a < 4; ++a) acc += loc[4 * i + a] * dN[3 * a + j];
            J[3 * i + j] = acc;
        }
    J[8] = (J[0] * J[4] - J[3] * J[1] > 0.0) ? 1.0 : -1.0;
    det_inv(*det, Ji, J, 3);
}

/* Element stiffness (condensed to the 24 nodal DOFs) and consistent mass of a
 * single isotropic layer of thickness t at the mid-plane (z offset 0).  The
 * element frame is the identity when the element normal is parallel to z
 * (correct_orient_dfd0, S/element_properties.rs:2052-2140, leaves the section
 * orientation untouched then); other elements are rejected (-2). */
int shell4_element(const double* X /* 3 x 4 SoA */, double E, double nu, double t, double rho, double* Ke,
                   double* Me) {
    double v1[3], v2[3], v3[3];
    for (int k = 0; k < 3; ++k) {
        v1[k] = X[4 * k + 2] - X[4 * k + 0];
        v2[k] = X[4 * k + 3] - X[4 * k + 1];
    }
    v3[0] = v1[1] * v2[2] - v1[2] * v2[1];
    v3[1] = v1[2] * v2[0] - v1[0] * v2[2];
    v3[2] = v1[0] * v2[1] - v1[1] * v2[0];
    if (v3[2] < 0.0)
        for (double& c : v3) c = -c;
    // cross of the section z axis (0,0,1) with v3
    if (std::sqrt(v3[0] * v3[0] + v3[1] * v3[1]) >= 1e-12) return -2;
    const double* loc = X;  // loc_ori = identity
    // ABD (get_abd_dfd0, S/element_properties.rs:766-876) of the one layer:
    // Q = S^-1 (get_layer_q_dfd0 :138-221), A = Q t, B = 0, D = Q t^3/12,
    // transverse shear and drill stiffness = A66
    double mod[3] = {E, E, E}, G = E / (2.0 * (1.0 + nu));
    double S[9] = {1.0 / mod[0], -nu / mod[0], 0.0, -nu / mod[0], 1.0 / mod[1], 0.0, 0.0, 0.0, 1.0 / G};
    double Q[9], dq;
    det_inv(dq, Q, S, 3);
    double C[81] = {0};
    const double zmin = -0.5 * t, zmax = 0.5 * t;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            C[9 * i + j] = (zmax - zmin) * Q[3 * i + j];
            C[9 * (i + 3) + j + 3] = (zmax * zmax * zmax - zmin * zmin * zmin) / 3.0 * Q[3 * i + j];
            C[9 * i + j + 3] = C[9 * (i + 3) + j] = 0.5 * (zmax * zmax - zmin * zmin) * Q[3 * i + j];
        }
    C[60] = C[70] = C[80] = C[20];
    // mass (get_shell_mass_dfd0, :1354-1408)
    double Mm[36] = {0};
    Mm[0] = Mm[7] = Mm[14] = rho * t;
    const double m1 = 0.5 * rho * (zmax * zmax - zmin * zmin), m2 = rho * (zmax * zmax * zmax - zmin * zmin * zmin) / 3.0;
    Mm[4] += m1;
    Mm[24] += m1;
    Mm[9] -= m1;
    Mm[19] -= m1;
    Mm[21] = Mm[28] = Mm[35] = m2;
    const int tot = 32;
    std::vector<double> K(tot * tot, 0.0);
    std::fill(Me, Me + 576, 0.0);
    const double g = 0.5773502691896257;
    const double ip[4][2] = {{-g, -g}, {g, -g}, {-g, g}, {g, g}};
    double N[10], dN[30], Nc[10], dNc[30], det, Ji[9], detc, Jic[9];
    const double zero[3] = {0, 0, 0};
    shell4_shape(zero, Nc, dNc);
    shell4_jac(loc, dNc, &detc, Jic);
    for (int q = 0; q < 4; ++q) {
        const double s[3] = {ip[q][0], ip[q][1], 0.0};
        shell4_shape(s, N, dN);
        shell4_jac(loc, dN, &det, Ji);
        double dx[10][3];
        for (int a = 0; a < 10; ++a) {
            const double* Jm = a < 4 ? Ji : Jic;
            for (int j = 0; j < 3; ++j) dx[a][j] = dN[3 * a] * Jm[j] + dN[3 * a + 1] * Jm[3 + j] + dN[3 * a + 2] * Jm[6 + j];
        }
        // section-deformation columns (get_section_def_dfd0 for one unit DOF,
        // S/element_soln_fields.rs:1160-1343)
        double B[9][32] = {{0}};
        for (int a = 0; a < 4; ++a) {
            const double Nx = dx[a][0], Ny = dx[a][1];
            for (int c = 0; c < 3; ++c) {  // translation along global c: local vector e_c
                double l[3] = {0, 0, 0};
                l[c] = 1.0;
                const int col = 6 * a + c;
                B[0][col] = l[0] * Nx;
                B[1][col] = l[1] * Ny;
                B[2][col] = l[0] * Ny + l[1] * Nx;
                B[6][col] = l[2] * Nx;
                B[7][col] = l[2] * Ny;
                B[8][col] = l[0] * Ny - l[1] * Nx;
            }
            for (int c = 0; c < 3; ++c) {  // rotation about global c
                double m[3] = {0, 0, 0};
                m[c] = 1.0;
                const int col = 6 * a + 3 + c;
                B[3][col] = m[1] * Nx;
                B[4][col] = -m[0] * Ny;
                B[5][col] = m[1] * Ny - m[0] * Nx;
                B[6][col] = m[1] * N[a];
                B[7][col] = -m[0] * N[a];
                B[8][col] = 2.0 * m[2] * N[a];
            }
        }
        // internal DOFs 24..31: (bubble, component) = (4,u) (5,u) (4,v) (5,v) (6,w) (7,w) (8,w) (9,w)
        const int ib[8] = {4, 5, 4, 5, 6, 7, 8, 9}, ic[8] = {0, 0, 1, 1, 2, 2, 2, 2};
        for (int k = 0; k < 8; ++k) {
            const double Nx = dx[ib[k]][0], Ny = dx[ib[k]][1];
            const int col = 24 + k;
            if (ic[k] == 0) {
                B[0][col] = Nx;
                B[2][col] = Ny;
                B[8][col] = Ny;
            } else if (ic[k] == 1) {
                B[1][col] = Ny;
                B[2][col] = Nx;
                B[8][col] = -Nx;
            } else {
                B[6][col] = Nx;
                B[7][col] = Ny;
            }
        }
        double CB[9][32];
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 32; ++j) {
                double acc = 0.0;
                for (int k = 0; k < 9; ++k) acc += C[9 * i + k] * B[k][j];
                CB[i][j] = acc * det;
            }
        for (int i = 0; i < 32; ++i)
            for (int j = 0; j < 32; ++j) {
                double acc = 0.0;
                for (int k = 0; k < 9; ++k) acc += B[k][i] * CB[k][j];
                K[i * tot + j] += acc;
            }
        // mass: b (6 x 24) = N_a times the local unit vector of each nodal DOF
        // (get_rum_dfd0, S/element_equations.rs:514-581)
        double b[6][24] = {{0}};
        for (int a = 0; a < 4; ++a)
            for (int c = 0; c < 6; ++c) b[c][6 * a + c] = N[a];
        for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j) {
                double acc = 0.0;
                for (int k = 0; k < 6; ++k)
                    for (int l = 0; l < 6; ++l) acc += b[k][i] * Mm[6 * k + l] * b[l][j];
                Me[24 * i + j] += acc * det;
            }
    }
    condense(K.data(), tot, 24, Ke);
    return 0;
}

struct Pattern {
    std::vector<int64_t> rowptr;
    std::vector<int32_t> cols;
};

/* Union of element DOF couplings, each row sorted ascending. */
Pattern build_pattern(int64_t n_el, int nn, const int32_t* conn, const int64_t* dof_index, int64_t n_dof,
                      int ndpn) {
    std::vector<std::vector<int32_t>> rows(n_dof);
    for (int64_t e = 0; e < n_el; ++e) {
        const int32_t* c = conn + nn * e;
        for (int a = 0; a < nn; ++a)
            for (int da = 0; da < ndpn; ++da) {
                int64_t r = dof_index[6 * (int64_t)c[a] + da];
                for (int b = 0; b < nn; ++b)
                    for (int db = 0; db < ndpn; ++db) rows[r].push_back((int32_t)dof_index[6 * (int64_t)c[b] + db]);
            }
    }
    Pattern p;
    p.rowptr.assign(n_dof + 1, 0);
    for (int64_t r = 0; r < n_dof; ++r) {
        auto& v = rows[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        p.rowptr[r + 1] = p.rowptr[r] + (int64_t)v.size();
    }
    p.cols.resize(p.rowptr[n_dof]);
    for (int64_t r = 0; r < n_dof; ++r) std::copy(rows[r].begin(), rows[r].end(), p.cols.begin() + p.rowptr[r]);
    return p;
}

inline int64_t find_col(const int64_t* rowptr, const int32_t* cols, int64_t r, int32_t c) {
    const int32_t* b = cols + rowptr[r];
    const int32_t* e = cols + rowptr[r + 1];
    const int32_t* it = std::lower_bound(b, e, c);
    if (it == e || *it != c) return -1;
    return it - cols;
}

// OpenMP threads of the Jacobi PCG and its SpMV (orc_set_threads; 1 = the
// serial restatement, whose sums run in index order)
int g_threads = 1;

void spmv_pen(int64_t n, const int64_t* rowptr, const int32_t* cols, const double* vals, int64_t n_pen,
              const int64_t* pen, double sf, const double* x, double* y) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t r = 0; r < n; ++r) {
        double acc = 0.0;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) acc += vals[k] * x[cols[k]];
        y[r] = acc;
    }
    for (int64_t i = 0; i < n_pen; ++i) y[pen[i]] += sf * x[pen[i]];
}

}  // namespace

/* ========================================================================= */

struct orc_ldl {
    int64_t dim = 0, size = 0, maxbw = 0;
    std::vector<int64_t> range, min_col;
    std::vector<double> mat, z;
};

extern "C" {

int orc_reorder_nodes(int64_t n_nodes, const double* xyz, int64_t n_el, int nn, int dof_per_nd,
                      const int32_t* conn, int64_t n_const, const int64_t* cterm_ptr,
                      const int64_t* term_set_ptr, const int32_t* set_nodes, int64_t block_dim,
                      int32_t* ordered, int64_t* dof_index, int64_t* n_dof_out) {
    if (n_nodes <= 0 || !ordered || !dof_index) return -1;
    /* nodal connectivity in element order, append-if-absent (S/model_prep.rs:76-101,
     * S/nd_el_set_meth.rs:5-14) */
    std::vector<std::vector<int32_t>> nc(n_nodes);
    auto add_if_absent = [&](int32_t a, int32_t b) {
        auto& v = nc[a];
        for (int32_t x : v)
            if (x == b) return;
        v.push_back(b);
    };
    std::vector<int> ndof(n_nodes, 3);
    for (int64_t e = 0; e < n_el; ++e) {
        const int32_t* c = conn + (int64_t)nn * e;
        for (int i = 0; i < nn; ++i) {
            if (dof_per_nd > 3) ndof[c[i]] = dof_per_nd;
            for (int j = i + 1; j < nn; ++j) {
                add_if_absent(c[i], c[j]);
                add_if_absent(c[j], c[i]);
            }
        }
    }
    /* constraint connectivity (S/model_prep.rs:20-54) */
    for (int64_t ci = 0; ci < n_const; ++ci) {
        for (int64_t t1 = cterm_ptr[ci]; t1 < cterm_ptr[ci + 1]; ++t1) {
            int64_t s1b = term_set_ptr[t1], s1e = term_set_ptr[t1 + 1];
            for (int64_t t2 = cterm_ptr[ci]; t2 < cterm_ptr[ci + 1]; ++t2) {
                int64_t s2b = term_set_ptr[t2], s2e = term_set_ptr[t2 + 1];
                int64_t l1 = s1e - s1b, l2 = s2e - s2b;
                if (l1 > 1 && l2 > 1) {
                    if (l1 > l2) return -2; /* mismatched set sizes */
                    for (int64_t k = 0; k < l1; ++k) {
                        int32_t a = set_nodes[s1b + k], b = set_nodes[s2b + k];
                        add_if_absent(a, b);
                        add_if_absent(b, a);
                    }
                } else {
                    for (int64_t i = s1b; i < s1e; ++i)
                        for (int64_t j = s2b; j < s2e; ++j) {
                            add_if_absent(set_nodes[i], set_nodes[j]);
                            add_if_absent(set_nodes[j], set_nodes[i]);
                        }
                }
            }
        }
    }
    /* least-connected start node, strict < (S/model_prep.rs:109-118) */
    int64_t min_ct = n_nodes, min_nd = 0;
    for (int64_t i = 0; i < n_nodes; ++i)
        if ((int64_t)nc[i].size() < min_ct) { min_ct = (int64_t)nc[i].size(); min_nd = i; }
    std::vector<char> ins(n_nodes, 0);
    ordered[0] = (int32_t)min_nd;
    ins[min_nd] = 1;
    int64_t on_len = 1, on_i = 0, since = 0, nd1 = min_nd;
    int64_t nd2 = min_nd;
    /* level-order insertion with restarts every block_dim nodes at the
     * nearest uninserted node (S/model_prep.rs:136-211, get_dist
     * S/matrix_functions.rs:33-41) */
    while (on_len < n_nodes) {
        size_t it = 0;
        const auto& nb = nc[nd1];
        bool end = false;
        while (!end && since < block_dim) {
            if (it < nb.size()) nd2 = nb[it++];
            else end = true;
            if (!ins[nd2]) {
                ordered[on_len++] = (int32_t)nd2;
                ins[nd2] = 1;
                ++since;
            }
        }
        if (on_len < n_nodes) {
            if (since >= block_dim || on_i == on_len - 1) {
                int64_t last = ordered[on_len - 1];
                double min_dist = 1.0e+100;
                const double* p1 = xyz + 3 * last;
                for (int64_t i = 0; i < n_nodes; ++i) {
                    if (ins[i]) continue;
                    const double* p2 = xyz + 3 * i;
                    double dx = p1[0] - p2[0], dy = p1[1] - p2[1], dz = p1[2] - p2[2];
                    double d = std::sqrt(dx * dx + dy * dy + dz * dz);
                    if (d < min_dist) { min_dist = d; min_nd = i; }
                }
                ordered[on_len++] = (int32_t)min_nd;
                nd1 = min_nd;
                ins[min_nd] = 1;
                ++since;
                on_i = on_len - 1;
                if (since >= block_dim) since = 0;
            } else {
                ++on_i;
                nd1 = ordered[on_i];
            }
        }
    }
    /* DOF numbering in insertion order (S/model_prep.rs:213-239) */
    for (int64_t i = 0; i < 6 * n_nodes; ++i) dof_index[i] = -1;
    int64_t d = 0;
    for (int64_t k = 0; k < n_nodes; ++k) {
        int64_t n = ordered[k];
        for (int c = 0; c < ndof[n]; ++c) dof_index[6 * n + c] = d++;
    }
    if (n_dof_out) *n_dof_out = d;
    return 0;
}

int orc_solid_cmat(const double* mod, const double* pr, const double* sm, double* C) {
    solid_cmat<double>(mod, pr, sm, C);
    return 0;
}

int orc_solid_element(int type, const double* X, const double* ori, const double* C, const double* ue,
                      double* K, double* R) {
    if (type != 8 && type != 81) return -1;
    int tot = (type == 81) ? 33 : 24;
    std::vector<double> Kf(tot * tot), Rf(tot), uf(tot, 0.0);
    if (ue) std::memcpy(uf.data(), ue, sizeof(double) * 24);
    if (type == 81 && ue) {
        /* internal DOFs consistent with the nodal state */
        std::vector<double> K0(tot * tot), R0(tot);
        hex_full<double>(type, X, ori, C, nullptr, K0.data(), R0.data());
        internal_from_external(K0.data(), tot, 24, ue, uf.data() + 24);
    }
    hex_full<double>(type, X, ori, C, uf.data(), Kf.data(), Rf.data());
    if (type == 81) condense(Kf.data(), tot, 24, K);
    else std::memcpy(K, Kf.data(), sizeof(double) * 576);
    if (R)
        for (int i = 0; i < 24; ++i) R[i] = Rf[i];
    return 0;
}

int orc_solid_element_dE(int type, const double* X, const double* ori, double E, double nu, double* K,
                         double* dK) {
    if (type != 8 && type != 81) return -1;
    int tot = (type == 81) ? 33 : 24;
    Dual mod[3], pr[3], sm[3], C[36];
    double G = E / (2.0 * (1.0 + nu)), dG = 1.0 / (2.0 * (1.0 + nu));
    for (int i = 0; i < 3; ++i) { mod[i] = Dual(E, 1.0); pr[i] = Dual(nu); sm[i] = Dual(G, dG); }
    solid_cmat<Dual>(mod, pr, sm, C);
    std::vector<Dual> Kf(tot * tot), Rf(tot);
    hex_full<Dual>(type, X, ori, C, nullptr, Kf.data(), Rf.data());
    std::vector<Dual> Kc(576);
    if (type == 81) condense(Kf.data(), tot, 24, Kc.data());
    else for (int i = 0; i < 576; ++i) Kc[i] = Kf[i];
    for (int i = 0; i < 576; ++i) { K[i] = Kc[i].v; dK[i] = Kc[i].d; }
    return 0;
}

/* dK_e/dp for one of the nine compliance parameters p = 0..8 (E1, E2, E3,
 * nu12, nu13, nu23, G12, G13, G23) by dual numbers through the pivot-clamped
 * QR inverse, as the reference's DiffDoub0 design-variable derivatives of
 * get_solid_stiff_dfd0 (S/element_properties.rs:700-759). */
int orc_solid_element_dmat(int type, const double* X, const double* modv, const double* prv, const double* smv,
                           int param, double* K, double* dK) {
    if (type != 8 && type != 81) return -1;
    if (param < 0 || param > 8) return -1;
    int tot = (type == 81) ? 33 : 24;
    Dual mod[3], pr[3], sm[3], C[36];
    for (int i = 0; i < 3; ++i) {
        mod[i] = Dual(modv[i], param == i ? 1.0 : 0.0);
        pr[i] = Dual(prv[i], param == 3 + i ? 1.0 : 0.0);
        sm[i] = Dual(smv[i], param == 6 + i ? 1.0 : 0.0);
    }
    solid_cmat<Dual>(mod, pr, sm, C);
    std::vector<Dual> Kf(tot * tot), Rf(tot);
    hex_full<Dual>(type, X, nullptr, C, nullptr, Kf.data(), Rf.data());
    std::vector<Dual> Kc(576);
    if (type == 81) condense(Kf.data(), tot, 24, Kc.data());
    else for (int i = 0; i < 576; ++i) Kc[i] = Kf[i];
    for (int i = 0; i < 576; ++i) { K[i] = Kc[i].v; dK[i] = Kc[i].d; }
    return 0;
}

int orc_solid_mass(int type, const double* X, double rho, double* M) {
    if (type != 8 && type != 81) return -1;
    std::memset(M, 0, sizeof(double) * 576);
    for (int gp = 0; gp < 8; ++gp) {
        double s[3], N[11], dNdx[33], detJ;
        hex_gauss(gp, s);
        hex_ip_data(type, X, s, N, dNdx, detJ);
        double w = detJ;
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) {
                double m = N[a] * N[b] * rho * w;
                for (int d = 0; d < 3; ++d) M[(3 * a + d) * 24 + 3 * b + d] += m;
            }
    }
    return 0;
}

int orc_assemble_solid(int type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                       const double* E, const double* nu, const int64_t* dof_index, int64_t n_dof,
                       int64_t* rowptr, int32_t* cols, double* vals, int64_t* nnz_out) {
    (void)n_nodes;
    if (type != 8 && type != 81) return -1;
    Pattern p = build_pattern(n_el, 8, conn, dof_index, n_dof, 3);
    if (nnz_out) *nnz_out = p.rowptr[n_dof];
    if (!rowptr) return 0;
    std::memcpy(rowptr, p.rowptr.data(), sizeof(int64_t) * (n_dof + 1));
    if (!cols || !vals) return 0;
    std::memcpy(cols, p.cols.data(), sizeof(int32_t) * p.cols.size());
    std::memset(vals, 0, sizeof(double) * p.cols.size());
    /* element matrices of a chunk are integrated in parallel (g_threads) and
     * scattered serially in element order, so the sums are those of the
     * serial element loop bit for bit */
    const int64_t chunk = 4096;
    std::vector<double> Kc(chunk * 576);
    int64_t gd[24];
    for (int64_t e0 = 0; e0 < n_el; e0 += chunk) {
        const int64_t ne = std::min(chunk, n_el - e0);
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t k = 0; k < ne; ++k) {
            double X[24];
            gather_X(xyz, conn + 8 * (e0 + k), 8, X);
            solid_ke_iso(type, X, E[e0 + k], nu[e0 + k], Kc.data() + 576 * k);
        }
        for (int64_t k = 0; k < ne; ++k) {
            const int32_t* c = conn + 8 * (e0 + k);
            const double* Ke = Kc.data() + 576 * k;
            for (int a = 0; a < 8; ++a)
                for (int d = 0; d < 3; ++d) gd[3 * a + d] = dof_index[6 * (int64_t)c[a] + d];
            for (int i = 0; i < 24; ++i)
                for (int j = 0; j < 24; ++j) {
                    int64_t kk = find_col(rowptr, cols, gd[i], (int32_t)gd[j]);
                    if (kk < 0) return -3;
                    vals[kk] += Ke[i * 24 + j];
                }
        }
    }
    return 0;
}

int orc_assemble_solid_mass(int type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                            const double* rho, const int64_t* dof_index, int64_t n_dof, const int64_t* rowptr,
                            const int32_t* cols, double* vals) {
    (void)n_nodes;
    std::memset(vals, 0, sizeof(double) * rowptr[n_dof]);
    double X[24], Me[576];
    int64_t gd[24];
    for (int64_t e = 0; e < n_el; ++e) {
        const int32_t* c = conn + 8 * e;
        gather_X(xyz, c, 8, X);
        orc_solid_mass(type, X, rho[e], Me);
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) gd[3 * a + d] = dof_index[6 * (int64_t)c[a] + d];
        for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j) {
                if (Me[i * 24 + j] == 0.0) continue;
                int64_t k = find_col(rowptr, cols, gd[i], (int32_t)gd[j]);
                if (k < 0) return -3;
                vals[k] += Me[i * 24 + j];
            }
    }
    return 0;
}

/* surfaceTraction (traction = load) and surfacePressure (traction = -p n of
 * the face, pressure non-null) of get_app_load_dfd0
 * (S/element_equations.rs:1970-2043); the load's normal direction is stored
 * normalised (Load::set_norm_dir, S/load_meth.rs:22-32). */
static int surface_load(int type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                        const int64_t* dof_index, int64_t n_set, const int32_t* el_set, const double* load,
                        const double* pressure, const double* norm_dir_in, double norm_tol_deg, double* f) {
    (void)n_nodes;
    if (type != 8 && type != 81) return -1;
    double norm_dir[3];
    {
        double m = 0.0;
        for (int k = 0; k < 3; ++k) m += norm_dir_in[k] * norm_dir_in[k];
        m = 1.0 / std::sqrt(m);
        for (int k = 0; k < 3; ++k) norm_dir[k] = m * norm_dir_in[k];
    }
    /* hex faces, local node order of Element::initialize_faces (S/element_meth.rs:537-590) */
    static const int fl[6][4] = {{3, 2, 1, 0}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                 {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
    /* surface faces: a face whose sorted node set appears on no other face
     * (find_surface_faces S/model_prep.rs:510-554, FacePtList::add_if_absent
     * S/face_meth.rs:1035-1066) */
    struct FK { int32_t n[4]; int64_t id; };
    std::vector<FK> keys(6 * n_el);
    for (int64_t e = 0; e < n_el; ++e)
        for (int f6 = 0; f6 < 6; ++f6) {
            FK k;
            for (int i = 0; i < 4; ++i) k.n[i] = conn[8 * e + fl[f6][i]];
            std::sort(k.n, k.n + 4);
            k.id = 6 * e + f6;
            keys[6 * e + f6] = k;
        }
    std::vector<char> on_surf(6 * n_el, 1);
    std::vector<int64_t> idx(6 * n_el);
    for (int64_t i = 0; i < 6 * n_el; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
        return std::lexicographical_compare(keys[a].n, keys[a].n + 4, keys[b].n, keys[b].n + 4);
    });
    for (int64_t i = 0; i + 1 < 6 * n_el; ++i) {
        const FK& a = keys[idx[i]];
        const FK& b = keys[idx[i + 1]];
        if (std::equal(a.n, a.n + 4, b.n)) { on_surf[idx[i]] = 0; on_surf[idx[i + 1]] = 0; }
    }
    double cos_tol = std::cos(kPi180 * norm_tol_deg);
    for (int64_t si = 0; si < n_set; ++si) {
        int64_t e = el_set[si];
        const int32_t* c = conn + 8 * e;
        double X[24];
        gather_X(xyz, c, 8, X);
        double acc[24];      /* glob_acc, index dof*8 + nd */
        double el_ld[24];    /* el_app_ld, index 3*nd + dof */
        double tot[3] = {0, 0, 0};
        bool any = false;
        for (int i = 0; i < 24; ++i) { acc[i] = 0.0; el_ld[i] = 0.0; }
        for (int f6 = 0; f6 < 6; ++f6) {
            if (!on_surf[6 * e + f6]) continue;
            const int* l = fl[f6];
            double v1[3], v2[3], n[3];
            for (int k = 0; k < 3; ++k) {
                v1[k] = X[8 * k + l[2]] - X[8 * k + l[0]];
                v2[k] = X[8 * k + l[3]] - X[8 * k + l[1]];
            }
            n[0] = v1[1] * v2[2] - v1[2] * v2[1];
            n[1] = v1[2] * v2[0] - v1[0] * v2[2];
            n[2] = v1[0] * v2[1] - v1[1] * v2[0];
            double area = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int k = 0; k < 3; ++k) n[k] *= 1.0 / area;
            area *= 0.5;
            double dp = norm_dir[0] * n[0] + norm_dir[1] * n[1] + norm_dir[2] * n[2];
            if (!(dp > cos_tol)) continue;
            any = true;
            double trac[3];
            for (int k = 0; k < 3; ++k) trac[k] = pressure ? -(*pressure) * n[k] : load[k];
            for (int i = 0; i < 4; ++i)
                for (int k = 0; k < 3; ++k) acc[k * 8 + l[i]] = trac[k];
            /* mass-like distribution with rho = 1 (get_rum_dfd0, actual_props = false) */
            for (int i = 0; i < 24; ++i) el_ld[i] = 0.0;
            for (int gp = 0; gp < 8; ++gp) {
                double s[3], N[11], dNdx[33], detJ;
                hex_gauss(gp, s);
                hex_ip_data(type, X, s, N, dNdx, detJ);
                double aip[3] = {0, 0, 0};
                for (int k = 0; k < 3; ++k)
                    for (int b = 0; b < 8; ++b) aip[k] += acc[k * 8 + b] * N[b];
                for (int a = 0; a < 8; ++a)
                    for (int k = 0; k < 3; ++k) el_ld[3 * a + k] += N[a] * 1.0 * aip[k] * detJ;
            }
            for (int a = 0; a < 8; ++a) {
                bool in = false;
                for (int i = 0; i < 4; ++i)
                    if (l[i] == a) in = true;
                for (int k = 0; k < 3; ++k) {
                    if (!in) el_ld[3 * a + k] = 0.0;
                    tot[k] += el_ld[3 * a + k];
                }
            }
            double inp = std::sqrt(trac[0] * trac[0] + trac[1] * trac[1] + trac[2] * trac[2]);
            double vm = std::sqrt(tot[0] * tot[0] + tot[1] * tot[1] + tot[2] * tot[2]);
            double sc = area * inp / vm;
            for (int i = 0; i < 24; ++i) el_ld[i] *= sc;
        }
        /* An element of the set with no matching face contributes nothing here
         * (the reference would scatter a stale scratch vector instead). */
        if (!any) continue;
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) f[dof_index[6 * (int64_t)c[a] + k]] += el_ld[3 * a + k];
    }
    return 0;
}

int orc_surface_traction(int type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                         const int64_t* dof_index, int64_t n_set, const int32_t* el_set, const double* load,
                         const double* norm_dir, double norm_tol_deg, double* f) {
    return surface_load(type, n_nodes, xyz, n_el, conn, dof_index, n_set, el_set, load, nullptr, norm_dir,
                        norm_tol_deg, f);
}

int orc_surface_pressure(int type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                         const int64_t* dof_index, int64_t n_set, const int32_t* el_set, double pressure,
                         const double* norm_dir, double norm_tol_deg, double* f) {
    return surface_load(type, n_nodes, xyz, n_el, conn, dof_index, n_set, el_set, nullptr, &pressure, norm_dir,
                        norm_tol_deg, f);
}

/* bodyForce (kind 0), gravitational (1), centrifugal (2) loads of
 * get_app_load_dfd0 (S/element_equations.rs:1876-1969): the element's nodal
 * acceleration field through get_rum_dfd0 (f_a = sum_gp detJ N_a rho
 * sum_b N_b acc_b, S/element_equations.rs:420-626) with rho = 1 for
 * bodyForce (then scaled by |b| V / |sum f|, V from get_volume_dfd0) and the
 * element density otherwise; the centrifugal field is omega^2 times the part
 * of (centroid - center) normal to the normalised axis (Load::set_axis). */
int orc_body_load(int type, int kind, int64_t n_el, const double* xyz, const int32_t* conn,
                  const int64_t* dof_index, int64_t n_set, const int32_t* el_set, const double* rho,
                  const double* load, const double* center, const double* axis_in, double ang_vel, double* f) {
    if (type != 8 && type != 81) return -1;
    double axis[3] = {0, 0, 0};
    if (kind == 2) {
        double m = 0.0;
        for (int k = 0; k < 3; ++k) m += axis_in[k] * axis_in[k];
        m = 1.0 / std::sqrt(m);
        for (int k = 0; k < 3; ++k) axis[k] = m * axis_in[k];
    }
    for (int64_t si = 0; si < n_set; ++si) {
        int64_t e = el_set[si];
        if (e < 0 || e >= n_el) return -1;
        const int32_t* c = conn + 8 * e;
        double X[24];
        gather_X(xyz, c, 8, X);
        double acc[24];  /* glob_acc, index dof*8 + nd */
        double den = 1.0;
        if (kind == 2) {
            double cent[3], dp = 0.0;
            for (int k = 0; k < 3; ++k) {
                cent[k] = 0.0;
                for (int a = 0; a < 8; ++a) cent[k] += X[8 * k + a];
                cent[k] *= 1.0 / 8.0;
                cent[k] -= center[k];
                dp += axis[k] * cent[k];
            }
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 8; ++a) acc[8 * k + a] = (cent[k] - axis[k] * dp) * ang_vel * ang_vel;
        } else {
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 8; ++a) acc[8 * k + a] = load[k];
        }
        if (kind != 0) den = rho[e];
        double el_ld[24], vol = 0.0;
        for (int i = 0; i < 24; ++i) el_ld[i] = 0.0;
        for (int gp = 0; gp < 8; ++gp) {
            double s[3], N[11], dNdx[33], detJ;
            hex_gauss(gp, s);
            hex_ip_data(type, X, s, N, dNdx, detJ);
            vol += detJ;
            double aip[3] = {0, 0, 0};
            for (int k = 0; k < 3; ++k)
                for (int b = 0; b < 8; ++b) aip[k] += acc[8 * k + b] * N[b];
            for (int a = 0; a < 8; ++a)
                for (int k = 0; k < 3; ++k) el_ld[3 * a + k] += N[a] * den * aip[k] * detJ;
        }
        if (kind == 0) {
            double tot[3] = {0, 0, 0}, inp = 0.0, vm = 0.0;
            for (int a = 0; a < 8; ++a)
                for (int k = 0; k < 3; ++k) tot[k] += el_ld[3 * a + k];
            for (int k = 0; k < 3; ++k) {
                inp += load[k] * load[k];
                vm += tot[k] * tot[k];
            }
            double sc = std::sqrt(inp) * vol / std::sqrt(vm);
            for (int i = 0; i < 24; ++i) el_ld[i] *= sc;
        }
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) f[dof_index[6 * (int64_t)c[a] + k]] += el_ld[3 * a + k];
    }
    return 0;
}

double orc_max_abs(int64_t nnz, const double* vals) {
    double m = 0.0;
    for (int64_t i = 0; i < nnz; ++i) m = std::max(m, std::fabs(vals[i]));
    return m;
}

int orc_spmv_penalty(int64_t n, const int64_t* rowptr, const int32_t* cols, const double* vals, int64_t n_pen,
                     const int64_t* pen_dofs, double sf, const double* x, double* y) {
    spmv_pen(n, rowptr, cols, vals, n_pen, pen_dofs, sf, x, y);
    return 0;
}

int orc_ldl_build(int64_t n, const int64_t* rowptr, const int32_t* cols, const double* vals, int64_t n_pen,
                  const int64_t* pen_dofs, double sf, int64_t block_dofs, orc_ldl** out) {
    orc_ldl* L = new orc_ldl();
    L->dim = n;
    L->range.assign(n + 1, 0);
    L->min_col.assign(n, 0);
    L->z.assign(n, 0.0);
    /* profile restricted to the diagonal block (allocate_from_sparse_mat,
     * S/lower_tri_mat_meth.rs:15-90); clamp rows add only their diagonal */
    for (int64_t r = 0; r < n; ++r) {
        int64_t rng = 1, blk = block_dofs * (r / block_dofs);
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            int64_t c = cols[k];
            if (c <= r && c >= blk) rng = std::max(rng, r - c + 1);
        }
        L->range[r] = rng;
    }
    L->size = 0;
    for (int64_t r = 0; r < n; ++r) {
        L->size += L->range[r];
        L->min_col[r] = r - (L->range[r] - 1);
        L->maxbw = std::max(L->maxbw, L->range[r]);
    }
    L->range[n] = L->size;
    for (int64_t r = n - 1; r >= 0; --r) L->range[r] = L->range[r + 1] - L->range[r];
    L->mat.assign(L->size, 0.0);
    /* populate_from_sparse_mat (S/lower_tri_mat_meth.rs:96-140) */
    for (int64_t r = 0; r < n; ++r)
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            int64_t c = cols[k];
            if (c <= r && c >= L->min_col[r]) L->mat[L->range[r] + (c - L->min_col[r])] += vals[k];
        }
    for (int64_t i = 0; i < n_pen; ++i) {
        int64_t d = pen_dofs[i];
        L->mat[L->range[d] + (d - L->min_col[d])] += sf;
    }
    /* ldl_factor (S/lower_tri_mat_meth.rs:142-192) */
    std::vector<double> ld(n, 0.0);
    auto& m = L->mat;
    for (int64_t i1 = 0; i1 < n; ++i1) {
        int64_t i3 = L->range[i1];
        for (int64_t i2 = L->min_col[i1]; i2 < i1; ++i2) {
            ld[i2] = m[i3] * m[L->range[i2 + 1] - 1];
            ++i3;
        }
        double sum = 0.0;
        int64_t i2 = L->range[i1];
        for (int64_t c = L->min_col[i1]; c < i1; ++c) sum += ld[c] * m[i2++];
        m[L->range[i1 + 1] - 1] -= sum;
        for (int64_t r = i1 + 1; r < i1 + L->maxbw && r < n; ++r) {
            if (L->min_col[r] > i1) continue;
            int64_t st = std::max(L->min_col[r], L->min_col[i1]);
            int64_t i4 = L->range[r] + (st - L->min_col[r]);
            double s2 = 0.0;
            for (int64_t c = st; c < i1; ++c) s2 += ld[c] * m[i4++];
            int64_t ii = L->range[r] + (i1 - L->min_col[r]);
            m[ii] = (m[ii] - s2) / m[L->range[i1 + 1] - 1];
        }
    }
    *out = L;
    return 0;
}

int orc_ldl_solve(orc_ldl* L, const double* rhs, double* x) {
    int64_t n = L->dim;
    auto& m = L->mat;
    auto& z = L->z;
    z[0] = rhs[0];
    for (int64_t i = 1; i < n; ++i) {
        double sum = 0.0;
        int64_t k = L->range[i];
        for (int64_t c = L->min_col[i]; c < i; ++c) sum += m[k++] * z[c];
        z[i] = rhs[i] - sum;
    }
    for (int64_t i = n - 1; i >= 0; --i) {
        double sum = 0.0;
        int64_t stop = std::min(i + L->maxbw + 1, n);
        for (int64_t r = i + 1; r < stop; ++r)
            if (i >= L->min_col[r]) sum += m[L->range[r] + (i - L->min_col[r])] * x[r];
        x[i] = z[i] / m[L->range[i + 1] - 1] - sum;
    }
    return 0;
}

int orc_ldl_pos_def(orc_ldl* L) {
    for (int64_t i = 0; i < L->dim; ++i)
        if (L->mat[L->range[i + 1] - 1] < 0.0) return 0;
    return 1;
}

int64_t orc_ldl_max_bandwidth(orc_ldl* L) { return L->maxbw; }

void orc_ldl_free(orc_ldl* L) { delete L; }

int orc_conj_grad_sparse(int64_t n, const int64_t* rowptr, const int32_t* cols, const double* vals,
                         int64_t n_pen, const int64_t* pen_dofs, double sf, orc_ldl* pc, const double* rhs,
                         double conv_tol, int64_t max_it, double* x, int64_t* it_out, double* res_out) {
    std::vector<double> h(n), g(n), z(n), w(n);
    for (int64_t i = 0; i < n; ++i) { x[i] = 0.0; g[i] = -rhs[i]; }
    orc_ldl_solve(pc, g.data(), w.data());
    double res = 0.0;
    for (int64_t i = 0; i < n; ++i) { h[i] = -w[i]; res += g[i] * w[i]; }
    int64_t it = 0;
    while (it < max_it && res > conv_tol) {
        spmv_pen(n, rowptr, cols, vals, n_pen, pen_dofs, sf, h.data(), z.data());
        double dp = 0.0;
        for (int64_t i = 0; i < n; ++i) dp += h[i] * z[i];
        double alpha = res / dp;
        for (int64_t i = 0; i < n; ++i) { x[i] += alpha * h[i]; g[i] += alpha * z[i]; }
        orc_ldl_solve(pc, g.data(), w.data());
        double rn = 0.0;
        for (int64_t i = 0; i < n; ++i) rn += g[i] * w[i];
        double beta = rn / res;
        for (int64_t i = 0; i < n; ++i) h[i] = -w[i] + beta * h[i];
        res = rn;
        ++it;
    }
    if (it_out) *it_out = it;
    if (res_out) *res_out = res;
    return 0;
}

int orc_set_threads(int n) {
    g_threads = n > 0 ? n : 1;
    return 0;
}

int orc_jacobi_pcg(int64_t n, const int64_t* rowptr, const int32_t* cols, const double* vals, int64_t n_pen,
                   const int64_t* pen_dofs, double sf, const double* rhs, double rel_tol, int64_t max_it,
                   double* x, int64_t* it_out, double* relres_out) {
    std::vector<double> dinv(n, 0.0), r(n), z(n), p(n), q(n);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
            if (cols[k] == i) dinv[i] += vals[k];
    for (int64_t i = 0; i < n_pen; ++i) dinv[pen_dofs[i]] += sf;
    for (int64_t i = 0; i < n; ++i) dinv[i] = 1.0 / dinv[i];
    double fn = 0.0;
    for (int64_t i = 0; i < n; ++i) { x[i] = 0.0; r[i] = rhs[i]; fn += rhs[i] * rhs[i]; }
    fn = std::sqrt(fn);
    double rz = 0.0, rr = 0.0;
    for (int64_t i = 0; i < n; ++i) { z[i] = dinv[i] * r[i]; p[i] = z[i]; rz += r[i] * z[i]; rr += r[i] * r[i]; }
    int64_t it = 0;
    while (it < max_it && std::sqrt(rr) > rel_tol * fn) {
        spmv_pen(n, rowptr, cols, vals, n_pen, pen_dofs, sf, p.data(), q.data());
        double pq = 0.0;
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(+ : pq)
        for (int64_t i = 0; i < n; ++i) pq += p[i] * q[i];
        double alpha = rz / pq;
        double rzn = 0.0;
        double rr_new = 0.0;
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(+ : rzn, rr_new)
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            z[i] = dinv[i] * r[i];
            rzn += r[i] * z[i];
            rr_new += r[i] * r[i];
        }
        rr = rr_new;
        double beta = rzn / rz;
        rz = rzn;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        ++it;
    }
    /* true residual */
    spmv_pen(n, rowptr, cols, vals, n_pen, pen_dofs, sf, x, q.data());
    double tr = 0.0;
    for (int64_t i = 0; i < n; ++i) tr += (rhs[i] - q[i]) * (rhs[i] - q[i]);
    if (it_out) *it_out = it;
    if (relres_out) *relres_out = std::sqrt(tr) / fn;
    return 0;
}

int orc_solid_sens(int type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                   const double* E, const double* nu, const int64_t* dof_index, const double* u, const double* lam,
                   double* dJdE) {
    (void)n_nodes;
    if (type != 8 && type != 81) return -1;
    const int tot = (type == 81) ? 33 : 24;
    /* elements are independent: parallel over g_threads */
#pragma omp parallel num_threads(g_threads)
    {
    std::vector<Dual> Kf(tot * tot), Rf(tot);
    std::vector<double> ue(tot), le(tot);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < n_el; ++e) {
        const int32_t* c = conn + 8 * e;
        double X[24];
        gather_X(xyz, c, 8, X);
        double G = E[e] / (2.0 * (1.0 + nu[e])), dG = 1.0 / (2.0 * (1.0 + nu[e]));
        Dual mod[3], pr[3], sm[3], C[36];
        for (int i = 0; i < 3; ++i) { mod[i] = Dual(E[e], 1.0); pr[i] = Dual(nu[e]); sm[i] = Dual(G, dG); }
        solid_cmat<Dual>(mod, pr, sm, C);
        for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) {
                int64_t g = dof_index[6 * (int64_t)c[a] + d];
                ue[3 * a + d] = u[g];
                le[3 * a + d] = lam[g];
            }
        if (type == 81) {
            /* condensed DOFs of the solution and of the adjoint (update_internal,
             * for_soln = 1 and 0; S/element_equations.rs:123-186) */
            hex_full<Dual>(type, X, nullptr, C, nullptr, Kf.data(), Rf.data());
            internal_from_external(Kf.data(), tot, 24, ue.data(), ue.data() + 24);
            internal_from_external(Kf.data(), tot, 24, le.data(), le.data() + 24);
        }
        /* dual-number element residual R(u; E + eps) -> dR/dE (get_ru_dfd1) */
        hex_full<Dual>(type, X, nullptr, C, ue.data(), Kf.data(), Rf.data());
        double s = 0.0;
        for (int i = 0; i < tot; ++i) s += le[i] * Rf[i].d;
        dJdE[e] = -s;
    }
    }
    return 0;
}


int orc_dv_loop(int type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn, const double* E,
                const double* nu, const int64_t* dof_index, int64_t n_dof, const double* u, const double* lam,
                int64_t n_load_el, const int32_t* load_el, int64_t n_dv, double* dJdE) {
    if (type != 8 && type != 81) return -1;
    if (n_dv > n_el) return -2;
    const int tot = (type == 81) ? 33 : 24;
    std::vector<Dual> crd(3 * n_nodes), drud(n_dof), Kf(tot * tot), Rf(tot);
    std::vector<char> el_in(n_el, 0);
    std::vector<double> ue(tot), le(tot);
    for (int64_t d = 0; d < n_dv; ++d) {
        /* coordinate duals of every node (no coordinate DVs: derivative 0) */
        for (int64_t i = 0; i < 3 * n_nodes; ++i) crd[i] = Dual(xyz[i], 0.0);
        for (int64_t i = 0; i < n_dof; ++i) drud[i] = Dual(0.0, 0.0);
        for (int64_t e = 0; e < n_el; ++e) el_in[e] = 0;
        el_in[d] = 1;
        /* applied loads on the DV's elements (a traction does not depend on
         * the modulus: the scan is the cost) */
        int64_t hits = 0;
        for (int64_t k = 0; k < n_load_el; ++k) hits += el_in[load_el[k]];
        (void)hits;
        for (int64_t i = 0; i < n_dof; ++i) drud[i] = Dual(-drud[i].v, -drud[i].d);
        /* solution-dependent residual of the DV's element with E + eps */
        const int32_t* c = conn + 8 * d;
        double X[24];
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) X[8 * k + a] = crd[3 * (int64_t)c[a] + k].v;
        const double G = E[d] / (2.0 * (1.0 + nu[d])), dG = 1.0 / (2.0 * (1.0 + nu[d]));
        Dual mod[3], pr[3], sm[3], C[36];
        for (int i = 0; i < 3; ++i) { mod[i] = Dual(E[d], 1.0); pr[i] = Dual(nu[d]); sm[i] = Dual(G, dG); }
        solid_cmat<Dual>(mod, pr, sm, C);
        int64_t gd[24];
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) {
                gd[3 * a + k] = dof_index[6 * (int64_t)c[a] + k];
                ue[3 * a + k] = u[gd[3 * a + k]];
                le[3 * a + k] = lam[gd[3 * a + k]];
            }
        if (type == 81) {
            hex_full<Dual>(type, X, nullptr, C, nullptr, Kf.data(), Rf.data());
            internal_from_external(Kf.data(), tot, 24, ue.data(), ue.data() + 24);
            internal_from_external(Kf.data(), tot, 24, le.data(), le.data() + 24);
        }
        hex_full<Dual>(type, X, nullptr, C, ue.data(), Kf.data(), Rf.data());
        for (int i = 0; i < 24; ++i) drud[gd[i]] = drud[gd[i]] + Rf[i];
        for (int64_t i = 0; i < 3 * n_nodes; ++i) crd[i] = Dual(xyz[i], 0.0);
        /* dL/dd -= lam . dR/dd over every DOF, then the internal-DOF terms */
        double s = 0.0;
        for (int64_t i = 0; i < n_dof; ++i) s += lam[i] * drud[i].d;
        double s_int = 0.0;
        for (int64_t e = 0; e < n_el; ++e)
            if (el_in[e] && type == 81)
                for (int i = 24; i < tot; ++i) s_int += le[i] * Rf[i].d;
        dJdE[d] = -s - s_int;
    }
    return 0;
}

int orc_shell4_element(const double* X, double E, double nu, double t, double rho, double* Ke, double* Me) {
    double Mtmp[576];
    return shell4_element(X, E, nu, t, rho, Ke, Me ? Me : Mtmp);
}

int orc_assemble_shell(int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn, const double* E,
                       const double* nu, const double* t, const double* rho, const int64_t* dof_index, int64_t n_dof,
                       int64_t* rowptr, int32_t* cols, double* vals, double* mvals, int64_t* nnz_out) {
    (void)n_nodes;
    Pattern p = build_pattern(n_el, 4, conn, dof_index, n_dof, 6);
    if (nnz_out) *nnz_out = p.rowptr[n_dof];
    if (!rowptr) return 0;
    std::memcpy(rowptr, p.rowptr.data(), sizeof(int64_t) * (n_dof + 1));
    if (!cols || !vals) return 0;
    std::memcpy(cols, p.cols.data(), sizeof(int32_t) * p.cols.size());
    std::memset(vals, 0, sizeof(double) * p.cols.size());
    if (mvals) std::memset(mvals, 0, sizeof(double) * p.cols.size());
    double X[12], Ke[576], Me[576];
    int64_t gd[24];
    for (int64_t e = 0; e < n_el; ++e) {
        const int32_t* c = conn + 4 * e;
        gather_X(xyz, c, 4, X);
        int rc = shell4_element(X, E[e], nu[e], t[e], rho ? rho[e] : 0.0, Ke, Me);
        if (rc) return rc;
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 6; ++d) gd[6 * a + d] = dof_index[6 * (int64_t)c[a] + d];
        for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j) {
                int64_t k = find_col(rowptr, cols, gd[i], (int32_t)gd[j]);
                if (k < 0) return -3;
                vals[k] += Ke[i * 24 + j];
                if (mvals) mvals[k] += Me[i * 24 + j];
            }
    }
    return 0;
}

}  // extern "C"
