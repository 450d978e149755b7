// Newmark time stepping on the device (FP64 state and effective load).
//
// The reference advances every node with update_vel_acc / advance_disp
// (S/node_meth.rs:127-199) around one linear solve per step in solve_step
// (S/model_analysis.rs:312-558, loop in Model::solve :560-615).  In its
// (beta - gamma) form, with c1 = 1/(dt^2 (beta - gamma)),
// c2 = dt^2 (0.5 + beta - gamma):
//   a(u) = c1 (u_n - u + dt v_n + c2 a_n)
//   solve (K + c_m M) du = f(t_n) - K u_n - M a(u_n),  c_m = -c1
//   u_{n+1} = u_n + du,  a_{n+1} = a(u_{n+1}),
//   v_{n+1} = v_n + dt ((1 - gamma) a_n + gamma a_{n+1}).
// With w = u_n + dt v_n + c2 a_n:  a(u_n) = c_m (u_n - w), a_{n+1} = c_m (u_{n+1} - w),
// and since only K_eff = K + c_m M (penalty included) and M are stored,
//   rhs = f(t_n) - K u_n - M a(u_n) = f(t_n) + c_m M w - K_eff u_n.
// The load of the step that produces u_{n+1} is the one at t_n (the reference
// calls solve_step(time) before advancing time).
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernels.hpp"

namespace asend {

namespace {

// w = u + dt v + c2 a
__global__ void k_nm_predict(int64_t n, const double* __restrict__ u, const double* __restrict__ v,
                             const double* __restrict__ a, double dt, double c2, double* __restrict__ w) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        w[i] = u[i] + dt * v[i] + c2 * a[i];
}

// rhs0 = fac[step] f_shape + c_m q  (q = M w); K_eff u_n is subtracted by the
// residual SpMV that follows
__global__ void k_nm_rhs(int64_t n, const int64_t* __restrict__ step, const double* __restrict__ fac,
                         const double* __restrict__ f_shape, double c_m, const double* __restrict__ q,
                         double* __restrict__ rhs) {
    const double s = fac[*step];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        rhs[i] = s * f_shape[i] + c_m * q[i];
}

// u += du ; a_new = c_m (u - w) ; v += dt ((1 - gamma) a + gamma a_new) ; a = a_new
__global__ void k_nm_update(int64_t n, const double* __restrict__ du, const double* __restrict__ w, double c_m,
                            double dt, double gamma, double* __restrict__ u, double* __restrict__ v,
                            double* __restrict__ a) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double un = u[i] + du[i];
        const double an = c_m * (un - w[i]);
        v[i] += dt * ((1.0 - gamma) * a[i] + gamma * an);
        u[i] = un;
        a[i] = an;
    }
}

// history of the watched DOFs, per-step iteration count and convergence flag;
// then advance the device step counter (one block: the increment follows
// every read of the counter)
__global__ __launch_bounds__(256) void k_nm_record(int n_watch, const int64_t* __restrict__ watch,
                                                   const double* __restrict__ u, double* __restrict__ hist,
                                                   int64_t* __restrict__ step, int64_t* __restrict__ its,
                                                   const Scalars* __restrict__ sc) {
    const int64_t s = *step;
    for (int k = threadIdx.x; k < n_watch; k += blockDim.x) hist[s * n_watch + k] = u[watch[k]];
    __syncthreads();
    if (threadIdx.x == 0) {
        const bool conv = sc->done && sc->rr <= sc->tol2 * sc->ff;
        its[s] = conv ? sc->it : -sc->it - 1;
        *step = s + 1;
    }
}

// uhist[step] = u, the displacement history kept for the adjoint sweep
// (launched before k_nm_record advances the step counter)
__global__ void k_nm_save(int64_t n, const int64_t* __restrict__ step, const double* __restrict__ u,
                          double* __restrict__ uhist) {
    double* dst = uhist + (*step) * n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = u[i];
}

// yhist[step] = u + ck v
__global__ void k_nm_save_y(int64_t n, const int64_t* __restrict__ step, const double* __restrict__ u,
                            const double* __restrict__ v, double ck, double* __restrict__ yhist) {
    double* dst = yhist + (*step) * n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = u[i] + ck * v[i];
}

// Rayleigh damping D = cm M + ck K enters the step's residual through the
// velocity v(u) = v + dt ((1 - gamma) a + gamma c_m (u - w)); with
// K_eff = K + c_m M + cd D (cd = c_m dt gamma) the residual at u_n is
// f + c_m M w - D z - K_eff u_n with z = v + dt (1 - gamma) a - cd w.
__global__ void k_nm_damp_z(int64_t n, const double* __restrict__ v, const double* __restrict__ a,
                            const double* __restrict__ w, double dt, double gamma, double cd, double* __restrict__ z) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        z[i] = v[i] + dt * (1.0 - gamma) * a[i] - cd * w[i];
}

// One backward step of the transient adjoint.  Objective partials at this
// step: dL/du = wn q, dL/dv = dL/da = 0.  Unless this is the last step, the
// coupling to step n+1 (augmentd_ld_u, S/model_analysis.rs:936-1063) with
// t = a_adj - M lam_{n+1} and d = D lam_{n+1} (Rayleigh damping, :1033-1056):
//   gu += c1 t - dvu d ; ga = c1 c2 t + dt (1 - gamma) v_adj - dva d ;
//   gv = c1 dt t + v_adj - dvv d
// then (solve_for_adjoint, :1089-1100):
//   v_adj = gv ; a_adj = ga + dt gamma v_adj ; rhs = gu - c1 a_adj.
__global__ void k_adj_prep(int64_t n, const double* __restrict__ q, double wn, double c1, double c2, double dt,
                           double gamma, int last, const double* __restrict__ mlam, const double* __restrict__ dlam,
                           double dvu, double dva, double dvv, double* __restrict__ aadj, double* __restrict__ vadj,
                           double* __restrict__ rhs) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double gu = wn * q[i], ga = 0.0, gv = 0.0;
        if (!last) {
            const double t = aadj[i] - mlam[i];
            const double d = dlam ? dlam[i] : 0.0;
            gu += c1 * t - dvu * d;
            ga = c1 * c2 * t + dt * (1.0 - gamma) * vadj[i] - dva * d;
            gv = c1 * dt * t + vadj[i] - dvv * d;
        }
        const double va = gv, aa = ga + dt * gamma * va;
        vadj[i] = va;
        aadj[i] = aa;
        rhs[i] = gu - c1 * aa;
    }
}

int nb(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 4096); }

}  // namespace

void launch_nm_predict(hipStream_t st, int64_t n, const double* u, const double* v, const double* a, double dt,
                       double c2, double* w) {
    hipLaunchKernelGGL(k_nm_predict, dim3(nb(n)), dim3(256), 0, st, n, u, v, a, dt, c2, w);
}

void launch_nm_rhs(hipStream_t st, int64_t n, const int64_t* step, const double* fac, const double* f_shape,
                   double c_m, const double* q, double* rhs) {
    hipLaunchKernelGGL(k_nm_rhs, dim3(nb(n)), dim3(256), 0, st, n, step, fac, f_shape, c_m, q, rhs);
}

void launch_nm_update(hipStream_t st, int64_t n, const double* du, const double* w, double c_m, double dt,
                      double gamma, double* u, double* v, double* a) {
    hipLaunchKernelGGL(k_nm_update, dim3(nb(n)), dim3(256), 0, st, n, du, w, c_m, dt, gamma, u, v, a);
}

void launch_nm_save(hipStream_t st, int64_t n, const int64_t* step, const double* u, double* uhist) {
    hipLaunchKernelGGL(k_nm_save, dim3(nb(n)), dim3(256), 0, st, n, step, u, uhist);
}

void launch_nm_save_y(hipStream_t st, int64_t n, const int64_t* step, const double* u, const double* v, double ck,
                      double* yhist) {
    hipLaunchKernelGGL(k_nm_save_y, dim3(nb(n)), dim3(256), 0, st, n, step, u, v, ck, yhist);
}

void launch_nm_damp_z(hipStream_t st, int64_t n, const double* v, const double* a, const double* w, double dt,
                      double gamma, double cd, double* z) {
    hipLaunchKernelGGL(k_nm_damp_z, dim3(nb(n)), dim3(256), 0, st, n, v, a, w, dt, gamma, cd, z);
}

void launch_adj_prep(hipStream_t st, int64_t n, const double* q, double wn, double c1, double c2, double dt,
                     double gamma, int last, const double* mlam, const double* dlam, double dvu, double dva,
                     double dvv, double* aadj, double* vadj, double* rhs) {
    hipLaunchKernelGGL(k_adj_prep, dim3(nb(n)), dim3(256), 0, st, n, q, wn, c1, c2, dt, gamma, last, mlam, dlam, dvu,
                       dva, dvv, aadj, vadj, rhs);
}

void launch_nm_record(hipStream_t st, int n_watch, const int64_t* watch, const double* u, double* hist,
                      int64_t* step, int64_t* its, const Scalars* sc) {
    hipLaunchKernelGGL(k_nm_record, dim3(1), dim3(256), 0, st, n_watch, watch, u, hist, step, its, sc);
}

}  // namespace asend
