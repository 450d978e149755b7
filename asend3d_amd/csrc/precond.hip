// Geometric multigrid preconditioner for structured brick8 grids (gfx950).
//
// Hierarchy: each level pairs the cells of the finer level along every axis
// (an odd count leaves its last cell single), until the coarsest level is a
// few hundred DOFs.  Prolongation is piecewise-linear interpolation in grid
// index space; restriction is its transpose.  Coarse operators are exactly
// Galerkin, A_{l+1} = P^T A_l P, computed element by element:
//   level 1: integrate each coarse element's trilinear basis over its fine
//            cells' Gauss points with the fine cells' moduli;
//   level l+1 >= 2: K_E = sum_children P_c^T K_c P_c (24x24 element matrices);
//   penalty rows: sf * P^T D P per component, formed on the host.
// Smoother: Chebyshev polynomial in the 3x3 block-Jacobi (or point-Jacobi, or
// on the finest level the learned factor) preconditioned operator, degree 2
// or 3 pre and post.  Coarsest level: explicit FP64 inverse (Gauss-Jordan,
// SPD, no pivoting: whole-matrix LDS for small levels, else blocked over
// 32-column pivot panels) applied as a dense GEMV.
// The Krylov wrapper is flexible CG with one truncation vector, since a
// V-cycle evaluated in FP32 is only approximately a fixed SPD operator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "asend_gpu.h"
#include "device_common.hpp"
#include "kernels.hpp"
#include "precond.hpp"
#include "learned.hpp"

namespace asend {

namespace {

constexpr int kChebDegreeDefault = 2;
constexpr int kChebMaxDegree = 8;
constexpr int kMaxCoarseNodes = 30;
constexpr int kMaxCoarseDenseNodes = 400;
constexpr int kPowerIts = 20;
constexpr int kMaxLevels = 32;  // bound of the hierarchy depth (coarsening stops at 17 levels)

// ---------------------------------------------------------------------------
// Host-side axis maps between a level and the next coarser level.
// ---------------------------------------------------------------------------
struct AxisMap {
    int nf = 0, nc = 0;          // fine / coarse cell counts
    std::vector<int> X;          // coarse node I -> fine node index (nc+1 entries)
    std::vector<int> pi;         // fine node i -> 2 coarse parents (-1 unused)
    std::vector<float> pw;       //   and weights
    std::vector<int> ri;         // coarse node I -> 3 fine children (-1 unused)
    std::vector<float> rw;
};

AxisMap make_axis(int nf) {
    AxisMap a;
    a.nf = nf;
    a.nc = (nf + 1) / 2;
    a.X.resize(a.nc + 1);
    for (int I = 0; I <= a.nc; ++I) a.X[I] = std::min(2 * I, nf);
    a.pi.assign(2 * (nf + 1), -1);
    a.pw.assign(2 * (nf + 1), 0.f);
    for (int i = 0; i <= nf; ++i) {
        int I = -1;
        if (i % 2 == 0) I = i / 2;
        else if (i == nf) I = a.nc;
        if (I >= 0) {
            a.pi[2 * i] = I;
            a.pw[2 * i] = 1.f;
        } else {
            a.pi[2 * i] = (i - 1) / 2;
            a.pw[2 * i] = 0.5f;
            a.pi[2 * i + 1] = (i - 1) / 2 + 1;
            a.pw[2 * i + 1] = 0.5f;
        }
    }
    a.ri.assign(3 * (a.nc + 1), -1);
    a.rw.assign(3 * (a.nc + 1), 0.f);
    std::vector<int> cnt(a.nc + 1, 0);
    for (int i = 0; i <= nf; ++i)
        for (int s = 0; s < 2; ++s) {
            int I = a.pi[2 * i + s];
            if (I < 0) continue;
            a.ri[3 * I + cnt[I]] = i;
            a.rw[3 * I + cnt[I]] = a.pw[2 * i + s];
            cnt[I]++;
        }
    return a;
}

// Device copy of the three axis maps for one level transition.
struct DevAxes {
    int* pi[3] = {nullptr, nullptr, nullptr};
    float* pw[3] = {nullptr, nullptr, nullptr};
    int* ri[3] = {nullptr, nullptr, nullptr};
    float* rw[3] = {nullptr, nullptr, nullptr};
    int* X[3] = {nullptr, nullptr, nullptr};
};

// Reduction scratch of one power iteration in flight: block partials, the
// step norms (ping-pong pair) and the result |M A x|^2 of the last step with
// x of unit norm.
struct PowerScratch {
    double* partials = nullptr;
    double* pp = nullptr;
    double* out = nullptr;
};

struct Level {
    int n[3] = {0, 0, 0};   // cells per axis
    int64_t n_nodes = 0, n_dof = 0, n_el = 0, nnzb = 0;
    bool owns_op = false;   // level 0 borrows the context's operator
    const int32_t* d_rp = nullptr;
    const int32_t* d_bc = nullptr;
    const float* d_val = nullptr;
    const float* d_dinv = nullptr;
    int32_t *o_rp = nullptr, *o_bc = nullptr, *d_kdiag = nullptr;
    float *o_val = nullptr, *o_dinv = nullptr;
    float* binv = nullptr;  // 3x3 block inverses, component-major (n_nodes x 9)
    std::vector<int32_t> h_rp, h_bc;
    // elements (levels >= 1)
    int32_t* d_conn = nullptr;
    uint8_t* d_slot = nullptr;
    float* d_Ke = nullptr;
    int32_t* d_elems = nullptr;
    std::vector<int> color_ptr;
    // transfer to the next finer level (levels >= 1)
    AxisMap ax[3];
    DevAxes dax;
    // penalty Galerkin additions
    int64_t n_padd = 0;
    int64_t* d_pidx = nullptr;
    float* d_pval = nullptr;
    // work vectors
    float *x = nullptr, *b = nullptr, *r = nullptr, *d = nullptr;
    double lmax = 0.0;
};

}  // namespace

struct Multigrid {
    std::string err;
    std::vector<Level> L;
    const int32_t* d_rank0 = nullptr;  // level-0 grid node -> fine row
    int32_t* o_rank0 = nullptr;
    std::vector<int32_t> h_rank0;
    const double* d_xyz_r = nullptr;
    const int32_t* d_conn_r = nullptr;
    int64_t n_el0 = 0;
    Sell3 fine;                        // level-0 operator layout (sliced ELL)
    // FP16 row-scaled level-0 values and the diagonal that restores them,
    // read by the cycle's residuals and the smoother bounds when cycle_half
    const __half* val16 = nullptr;
    const float* dinv16 = nullptr;
    bool cycle_half = false;
    // coarsest dense inverse
    int nco = 0;
    double* d_Ainv = nullptr;
    // FCG work vectors (fine level)
    float *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
    double* partials = nullptr;
    // the learned smoother's bound estimated on its own stream while the
    // hierarchy is set up on another (mg_learned_bound_begin/end): its own
    // reduction scratch and result
    double *partials_b = nullptr, *sv_b = nullptr, *h_sv_b = nullptr;
    // power iterations: the finest level's, beside a setup on another stream
    // (block-Jacobi or learned bound), and the coarse levels', which run on
    // two streams of their own while the hierarchy is built (level 1 on
    // pw_stream[0], deeper levels in turn on pw_stream[1]); pw_main also
    // serves host-read estimates after a setup
    PowerScratch pw_main, pw_side, pw_deep;
    double *partials_c = nullptr, *sv_c = nullptr;
    hipStream_t pw_stream[2] = {nullptr, nullptr};
    hipEvent_t ev_level[kMaxLevels] = {};
    hipEvent_t ev_pw_done[2] = {nullptr, nullptr};
    // that power iteration captured as a graph, with what it was captured for
    hipGraphExec_t pw_exec = nullptr;
    Factor pw_factor;
    const float* pw_val = nullptr;
    int pw_waves = 0;
    int pw_half = -1;
    // coarse-level smoother bounds: estimated and turned into Chebyshev
    // coefficients on the device during setup, read back once (d_lam ->
    // h_lam on lam_stream) when the host first needs them
    double *d_lam = nullptr, *h_lam = nullptr;
    bool lam_pending = false;
    // the finest level's block-Jacobi bound was enqueued by mg_fine_setup
    // (coefficients formed on the device, value in d_lam[0])
    bool fine_pending = false;
    hipStream_t lam_stream = nullptr;
    double* sv = nullptr;      // device scalars of the FCG loop
    double* h_sv = nullptr;    // pinned mirror
    // Chebyshev smoother: point (0) or block (1) Jacobi, degree and interval
    // [lo, hi] * lambda_max
    int smoother_kind = 1;
    int cheb_degree = kChebDegreeDefault;
    double cheb_lo = 0.1, cheb_hi = 1.15;
    // captured FCG iteration blocks (hipGraph), reused while their key matches
    struct FcgGraph {
        hipGraphExec_t exec = nullptr;
        int precond = -1, fine_smoother = -1, check = -1, degree = -1, kind = -1, waves = -1, half = -1;
        const float* val0 = nullptr;  // finest-level operator values
        const float* b = nullptr;
        float* x = nullptr;
        const float* ls = nullptr;
        const float* lst = nullptr;
        uint64_t use = 0;
    };
    FcgGraph graphs[4];
    uint64_t graph_clock = 0;
    bool use_graphs = true;
    // step coefficients (c1, c2) per level and step; slot L.size() is the
    // learned finest-level smoother
    float* d_coef = nullptr;
    std::vector<float> h_coef;
    // learned fine smoother: factor P = L L^T, its lambda_max(P A), scratch
    int fine_smoother = 0;
    bool has_factor = false;
    // penalty set the coarse penalty images (Level::d_pidx, d_pval) belong to
    std::vector<int64_t> pen_key;
    std::vector<int32_t> mult_key;
    bool pen_ready = false;
    Factor factor;
    double lmax_factor = 0.0;
    float *ft = nullptr, *fz = nullptr;
    bool keep_factor_bounds = false;  // diagnostics: reuse lambda_max(P A) of the previous factor
    // training of the learned smoother through one V-cycle: recorded vectors
    std::vector<float*> tv, tz;          // v_k (input of the k-th use of P) and dLoss/dz_k
    float *tx = nullptr, *td = nullptr, *txb = nullptr, *tdb = nullptr, *tb = nullptr, *tq = nullptr;
    bool ready = false;
};

namespace {

bool cycle_half(const Multigrid* mg) { return mg->cycle_half && mg->val16 != nullptr; }

int coef_slot(const Multigrid* mg, int level, int step) {
    (void)mg;
    return 2 * (level * kChebMaxDegree + step);
}

// Chebyshev coefficients of one smoothing call for bounds [lo, hi] * lmax:
// step 0: (0, 1/theta); step k: (rho_k rho_{k-1}, 2 rho_k / delta)
void cheb_coefficients(const Multigrid* mg, double lmax, float* out) {
    const double hi = lmax * mg->cheb_hi, lo = lmax * mg->cheb_lo;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho0 = 1.0 / sigma;
    out[0] = 0.f;
    out[1] = (float)(1.0 / theta);
    for (int k = 1; k < mg->cheb_degree; ++k) {
        const double rho1 = 1.0 / (2.0 * sigma - rho0);
        out[2 * k] = (float)(rho1 * rho0);
        out[2 * k + 1] = (float)(2.0 * rho1 / delta);
        rho0 = rho1;
    }
}

int mgfail(Multigrid* mg, int code, const char* msg) {
    if (mg) mg->err = msg;
    return code;
}

#define MGCHK(expr)                                                                                   \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) {                                                                       \
            mg->err = std::string(#expr) + ": " + hipGetErrorString(e_);                              \
            return ASEND_ERR_HIP;                                                                     \
        }                                                                                             \
    } while (0)

template <typename T>
hipError_t up(T** d, const T* h, size_t n) {
    hipError_t e = hipMalloc((void**)d, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    if (n) e = hipMemcpy(*d, h, n * sizeof(T), hipMemcpyHostToDevice);
    return e;
}

template <typename T>
void fr(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
__constant__ int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// Restriction: coarse node (I,J,K) <- sum of up to 27 fine nodes.
__global__ void k_restrict3(int ncx, int ncy, int ncz, int nfx, int nfy, const int* __restrict__ rix,
                            const float* __restrict__ rwx, const int* __restrict__ riy, const float* __restrict__ rwy,
                            const int* __restrict__ riz, const float* __restrict__ rwz,
                            const int32_t* __restrict__ frank, const float* __restrict__ rf, float* __restrict__ bc) {
    int64_t N = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t total = (int64_t)(ncx + 1) * (ncy + 1) * (ncz + 1);
    if (N >= total) return;
    int I = (int)(N % (ncx + 1));
    int J = (int)((N / (ncx + 1)) % (ncy + 1));
    int K = (int)(N / ((int64_t)(ncx + 1) * (ncy + 1)));
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    for (int c = 0; c < 3; ++c) {
        int k = riz[3 * K + c];
        if (k < 0) continue;
        float wz = rwz[3 * K + c];
        for (int b = 0; b < 3; ++b) {
            int j = riy[3 * J + b];
            if (j < 0) continue;
            float wyz = wz * rwy[3 * J + b];
            for (int a = 0; a < 3; ++a) {
                int i = rix[3 * I + a];
                if (i < 0) continue;
                float w = wyz * rwx[3 * I + a];
                int64_t g = i + (int64_t)(nfx + 1) * (j + (int64_t)(nfy + 1) * k);
                int64_t row = frank ? frank[g] : g;
                s0 += w * rf[3 * row];
                s1 += w * rf[3 * row + 1];
                s2 += w * rf[3 * row + 2];
            }
        }
    }
    bc[3 * N] = s0;
    bc[3 * N + 1] = s1;
    bc[3 * N + 2] = s2;
}

// Prolongation-add: fine node (i,j,k) += sum of up to 8 coarse parents.
__global__ void k_prolong_add3(int nfx, int nfy, int nfz, int ncx, int ncy, const int* __restrict__ pix,
                               const float* __restrict__ pwx, const int* __restrict__ piy,
                               const float* __restrict__ pwy, const int* __restrict__ piz,
                               const float* __restrict__ pwz, const int32_t* __restrict__ frank,
                               const float* __restrict__ xc, float* __restrict__ xf) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t total = (int64_t)(nfx + 1) * (nfy + 1) * (nfz + 1);
    if (g >= total) return;
    int i = (int)(g % (nfx + 1));
    int j = (int)((g / (nfx + 1)) % (nfy + 1));
    int k = (int)(g / ((int64_t)(nfx + 1) * (nfy + 1)));
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    for (int c = 0; c < 2; ++c) {
        int K = piz[2 * k + c];
        if (K < 0) continue;
        float wz = pwz[2 * k + c];
        for (int b = 0; b < 2; ++b) {
            int J = piy[2 * j + b];
            if (J < 0) continue;
            float wyz = wz * pwy[2 * j + b];
            for (int a = 0; a < 2; ++a) {
                int I = pix[2 * i + a];
                if (I < 0) continue;
                float w = wyz * pwx[2 * i + a];
                int64_t N = I + (int64_t)(ncx + 1) * (J + (int64_t)(ncy + 1) * K);
                s0 += w * xc[3 * N];
                s1 += w * xc[3 * N + 1];
                s2 += w * xc[3 * N + 2];
            }
        }
    }
    int64_t row = frank ? frank[g] : g;
    xf[3 * row] += s0;
    xf[3 * row + 1] += s1;
    xf[3 * row + 2] += s2;
}

// Level-1 element matrices from the fine mesh.  One wave per coarse element,
// lane (A, B) owns the 3x3 block; each of the (up to 8) fine cells in the
// coarse element's index box is integrated with its own 2x2x2 Gauss rule and
// moduli, using the coarse trilinear basis restricted to that cell.
__global__ __launch_bounds__(64) void k_level1_ke(int ncx, int ncy, int ncz, int nfx, int nfy,
                                                  const int* __restrict__ Xx, const int* __restrict__ Xy,
                                                  const int* __restrict__ Xz, const int32_t* __restrict__ conn_r,
                                                  const double* __restrict__ xyz_r, const double* __restrict__ E,
                                                  const double* __restrict__ nu, float* __restrict__ Ke) {
    const int lane = threadIdx.x;
    const int64_t ce = blockIdx.x;
    const int I = (int)(ce % ncx), J = (int)((ce / ncx) % ncy), K = (int)(ce / ((int64_t)ncx * ncy));
    const int x0 = Xx[I], x1 = Xx[I + 1], y0 = Xy[J], y1 = Xy[J + 1], z0 = Xz[K], z1 = Xz[K + 1];
    __shared__ float Xs[8][3];
    __shared__ float Ji[8][9];
    __shared__ float W[8];
    __shared__ float Gf[8][8][3];   // fine gradients [gp][a]
    __shared__ float Gc[8][8][3];   // coarse gradients [gp][A]
    __shared__ float P[8][8];       // fine node a -> coarse node A weights
    const int A = lane / 8, B = lane % 8;
    float k[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const float g = 0.5773502691896257f;
    for (int fz = z0; fz < z1; ++fz)
        for (int fy = y0; fy < y1; ++fy)
            for (int fx = x0; fx < x1; ++fx) {
                const int64_t e = fx + (int64_t)nfx * (fy + (int64_t)nfy * fz);
                __syncthreads();
                if (lane < 24) {
                    int a = lane / 3, c = lane % 3;
                    Xs[a][c] = (float)xyz_r[3 * (int64_t)conn_r[8 * e + a] + c];
                }
                {
                    // P[a][A] for this fine cell: per-axis linear weights in index space
                    int a = lane / 8, Ac = lane % 8;
                    float tx = (float)(fx + kCorner[a][0] - x0) / (float)(x1 - x0);
                    float ty = (float)(fy + kCorner[a][1] - y0) / (float)(y1 - y0);
                    float tz = (float)(fz + kCorner[a][2] - z0) / (float)(z1 - z0);
                    float w = (kCorner[Ac][0] ? tx : 1.f - tx) * (kCorner[Ac][1] ? ty : 1.f - ty) *
                              (kCorner[Ac][2] ? tz : 1.f - tz);
                    P[a][Ac] = w;
                }
                __syncthreads();
                if (lane < 8) {
                    const int gp = lane;
                    float s0 = (gp & 1) ? g : -g, s1 = (gp & 2) ? g : -g, s2 = (gp & 4) ? g : -g;
                    float Jm[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                    for (int a = 0; a < 8; ++a) {
                        float sx = kCorner[a][0] ? 1.f : -1.f, sy = kCorner[a][1] ? 1.f : -1.f,
                              sz = kCorner[a][2] ? 1.f : -1.f;
                        float fxx = 1.f + sx * s0, fyy = 1.f + sy * s1, fzz = 1.f + sz * s2;
                        float d0 = 0.125f * sx * fyy * fzz, d1 = 0.125f * fxx * sy * fzz, d2 = 0.125f * fxx * fyy * sz;
                        for (int i = 0; i < 3; ++i) {
                            Jm[i][0] += Xs[a][i] * d0;
                            Jm[i][1] += Xs[a][i] * d1;
                            Jm[i][2] += Xs[a][i] * d2;
                        }
                    }
                    float c00 = Jm[1][1] * Jm[2][2] - Jm[1][2] * Jm[2][1];
                    float c01 = Jm[1][2] * Jm[2][0] - Jm[1][0] * Jm[2][2];
                    float c02 = Jm[1][0] * Jm[2][1] - Jm[1][1] * Jm[2][0];
                    float det = Jm[0][0] * c00 + Jm[0][1] * c01 + Jm[0][2] * c02;
                    float id = 1.f / det;
                    Ji[gp][0] = c00 * id;
                    Ji[gp][1] = (Jm[0][2] * Jm[2][1] - Jm[0][1] * Jm[2][2]) * id;
                    Ji[gp][2] = (Jm[0][1] * Jm[1][2] - Jm[0][2] * Jm[1][1]) * id;
                    Ji[gp][3] = c01 * id;
                    Ji[gp][4] = (Jm[0][0] * Jm[2][2] - Jm[0][2] * Jm[2][0]) * id;
                    Ji[gp][5] = (Jm[0][2] * Jm[1][0] - Jm[0][0] * Jm[1][2]) * id;
                    Ji[gp][6] = c02 * id;
                    Ji[gp][7] = (Jm[0][1] * Jm[2][0] - Jm[0][0] * Jm[2][1]) * id;
                    Ji[gp][8] = (Jm[0][0] * Jm[1][1] - Jm[0][1] * Jm[1][0]) * id;
                    W[gp] = det;
                }
                __syncthreads();
                {
                    const int gp = lane / 8, a = lane % 8;
                    float s0 = (gp & 1) ? g : -g, s1 = (gp & 2) ? g : -g, s2 = (gp & 4) ? g : -g;
                    float sx = kCorner[a][0] ? 1.f : -1.f, sy = kCorner[a][1] ? 1.f : -1.f,
                          sz = kCorner[a][2] ? 1.f : -1.f;
                    float fxx = 1.f + sx * s0, fyy = 1.f + sy * s1, fzz = 1.f + sz * s2;
                    float d0 = 0.125f * sx * fyy * fzz, d1 = 0.125f * fxx * sy * fzz, d2 = 0.125f * fxx * fyy * sz;
                    for (int c = 0; c < 3; ++c) Gf[gp][a][c] = d0 * Ji[gp][c] + d1 * Ji[gp][3 + c] + d2 * Ji[gp][6 + c];
                }
                __syncthreads();
                {
                    const int gp = lane / 8, Ac = lane % 8;
                    float s[3] = {0.f, 0.f, 0.f};
                    for (int a = 0; a < 8; ++a) {
                        float w = P[a][Ac];
                        s[0] += w * Gf[gp][a][0];
                        s[1] += w * Gf[gp][a][1];
                        s[2] += w * Gf[gp][a][2];
                    }
                    Gc[gp][Ac][0] = s[0];
                    Gc[gp][Ac][1] = s[1];
                    Gc[gp][Ac][2] = s[2];
                }
                __syncthreads();
                const float Ee = (float)E[e], ne = (float)nu[e];
                const float lam = Ee * ne / ((1.f + ne) * (1.f - 2.f * ne));
                const float mu = Ee / (2.f * (1.f + ne));
                for (int gp = 0; gp < 8; ++gp) {
                    const float w = W[gp];
                    const float* ga = Gc[gp][A];
                    const float* gb = Gc[gp][B];
                    float dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j)
                            k[i][j] += w * (lam * ga[i] * gb[j] + mu * ga[j] * gb[i] + (i == j ? mu * dot : 0.f));
                }
            }
    float* out = Ke + 576 * ce;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[(3 * A + i) * 24 + 3 * B + j] = k[i][j];
}

// Galerkin element matrices of level l+1 from level l: K_E = sum_c P_c^T K_c P_c.
__global__ __launch_bounds__(64) void k_galerkin_ke(int ncx, int ncy, int ncz, int nfx, int nfy,
                                                    const int* __restrict__ Xx, const int* __restrict__ Xy,
                                                    const int* __restrict__ Xz, const float* __restrict__ Kf,
                                                    float* __restrict__ Kc) {
    const int lane = threadIdx.x;
    const int64_t ce = blockIdx.x;
    const int I = (int)(ce % ncx), J = (int)((ce / ncx) % ncy), K = (int)(ce / ((int64_t)ncx * ncy));
    const int x0 = Xx[I], x1 = Xx[I + 1], y0 = Xy[J], y1 = Xy[J + 1], z0 = Xz[K], z1 = Xz[K + 1];
    __shared__ float Ks[576];
    __shared__ float P[8][8];
    const int A = lane / 8, B = lane % 8;
    float k[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int fz = z0; fz < z1; ++fz)
        for (int fy = y0; fy < y1; ++fy)
            for (int fx = x0; fx < x1; ++fx) {
                const int64_t e = fx + (int64_t)nfx * (fy + (int64_t)nfy * fz);
                __syncthreads();
                for (int t = lane; t < 576; t += 64) Ks[t] = Kf[576 * e + t];
                {
                    int a = lane / 8, Ac = lane % 8;
                    float tx = (float)(fx + kCorner[a][0] - x0) / (float)(x1 - x0);
                    float ty = (float)(fy + kCorner[a][1] - y0) / (float)(y1 - y0);
                    float tz = (float)(fz + kCorner[a][2] - z0) / (float)(z1 - z0);
                    P[a][Ac] = (kCorner[Ac][0] ? tx : 1.f - tx) * (kCorner[Ac][1] ? ty : 1.f - ty) *
                               (kCorner[Ac][2] ? tz : 1.f - tz);
                }
                __syncthreads();
                for (int a = 0; a < 8; ++a) {
                    float wa = P[a][A];
                    if (wa == 0.f) continue;
                    for (int b = 0; b < 8; ++b) {
                        float w = wa * P[b][B];
                        if (w == 0.f) continue;
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j) k[i][j] += w * Ks[(3 * a + i) * 24 + 3 * b + j];
                    }
                }
            }
    float* out = Kc + 576 * ce;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[(3 * A + i) * 24 + 3 * B + j] = k[i][j];
}

// Coloured scatter of element matrices into the level's component-major BSR.
__global__ __launch_bounds__(64) void k_asm_ke(const int32_t* __restrict__ elems, const int32_t* __restrict__ conn,
                                               const float* __restrict__ Ke, const int32_t* __restrict__ rp,
                                               const uint8_t* __restrict__ slot, float* __restrict__ K) {
    const int lane = threadIdx.x;
    const int e = elems[blockIdx.x];
    const int A = lane / 8, B = lane % 8;
    const int ra = conn[8 * (int64_t)e + A];
    const int s = rp[ra], nb = rp[ra + 1] - s;
    const int kk = slot[64 * (int64_t)e + lane];
    const float* src = Ke + 576 * (int64_t)e;
    float* base = K + 9 * (int64_t)s + kk;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) base[(int64_t)(3 * i + j) * nb] += src[(3 * A + i) * 24 + 3 * B + j];
}

__global__ void k_add_values(int64_t n, const int64_t* __restrict__ idx, const float* __restrict__ v, float scale,
                             float* __restrict__ K) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) K[idx[i]] += scale * v[i];
}

__global__ void k_dinv32(int n_rows, const int32_t* __restrict__ rp, const int32_t* __restrict__ kdiag,
                         const float* __restrict__ K, float* __restrict__ dinv) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    int s = rp[r], nb = rp[r + 1] - s;
    for (int c = 0; c < 3; ++c) dinv[3 * r + c] = 1.f / K[9 * (int64_t)s + (4 * c) * nb + kdiag[r]];
}

// Inverse of the symmetric part of each 3x3 diagonal block (FP64 arithmetic),
// stored component-major: binv[c * n + r].  sptr != null selects the sliced-ELL
// layout of the finest level, else the component-major BSR of coarse levels.
__global__ void k_binv32(int n_rows, const int32_t* __restrict__ sptr, const int32_t* __restrict__ rp,
                         const int32_t* __restrict__ kdiag, const float* __restrict__ K, float* __restrict__ binv) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    double B[9];
    if (sptr) {
        for (int c = 0; c < 9; ++c) B[c] = K[sell_index(sptr, r, kdiag[r], c)];
    } else {
        const int s = rp[r], nb = rp[r + 1] - s;
        for (int c = 0; c < 9; ++c) B[c] = K[9 * (int64_t)s + c * nb + kdiag[r]];
    }
    const double a = B[0], b = 0.5 * (B[1] + B[3]), c = 0.5 * (B[2] + B[6]);
    const double d = B[4], e = 0.5 * (B[5] + B[7]), f = B[8];
    const double A00 = d * f - e * e, A01 = c * e - b * f, A02 = b * e - c * d;
    const double A11 = a * f - c * c, A12 = b * c - a * e, A22 = a * d - b * b;
    const double id = 1.0 / (a * A00 + b * A01 + c * A02);
    const double inv[9] = {A00 * id, A01 * id, A02 * id, A01 * id, A11 * id, A12 * id, A02 * id, A12 * id, A22 * id};
    for (int k = 0; k < 9; ++k) binv[(int64_t)k * n_rows + r] = (float)inv[k];
}

// Block-Jacobi Chebyshev steps (z_r = Binv_r r_r per node), coefficients
// (c1, c2) read from device memory so that captured graphs stay valid when
// the bounds change:   first: d = c2 z, x = d;   step: d = c1 d + c2 z, x += d
template <bool FIRST>
__global__ void k_cheb_block(int n_nodes, const float* __restrict__ binv, const float* __restrict__ r,
                             const float* __restrict__ coef, float* __restrict__ d, float* __restrict__ x) {
    const float c1 = coef[0], c2 = coef[1];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_nodes; i += gridDim.x * blockDim.x) {
        const float r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
        float m[9];
        for (int k = 0; k < 9; ++k) m[k] = binv[(int64_t)k * n_nodes + i];
        const float z[3] = {m[0] * r0 + m[1] * r1 + m[2] * r2, m[3] * r0 + m[4] * r1 + m[5] * r2,
                            m[6] * r0 + m[7] * r1 + m[8] * r2};
        for (int c = 0; c < 3; ++c) {
            if (FIRST) {
                const float v = c2 * z[c];
                d[3 * i + c] = v;
                x[3 * i + c] = v;
            } else {
                const float v = c1 * d[3 * i + c] + c2 * z[c];
                d[3 * i + c] = v;
                x[3 * i + c] += v;
            }
        }
    }
}

// The same per DOF, z = dinv r (point Jacobi) or z = r (already preconditioned)
template <bool FIRST>
__global__ void k_cheb_dof(int64_t n, const float* __restrict__ dinv, const float* __restrict__ r,
                           const float* __restrict__ coef, float* __restrict__ d, float* __restrict__ x) {
    const float c1 = coef[0], c2 = coef[1];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float z = dinv ? dinv[i] * r[i] : r[i];
        const float v = FIRST ? c2 * z : c1 * d[i] + c2 * z;
        d[i] = v;
        if (FIRST) x[i] = v;
        else x[i] += v;
    }
}

// out = a X + b Y
__global__ void k_lincomb(int64_t n, float a, const float* __restrict__ X, float b, const float* __restrict__ Y,
                          float* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = a * X[i] + b * Y[i];
}

// Reverse step of a Chebyshev update d = c1 d_prev + c2 z, x = x_prev + d:
// with t = dbar + xbar:  zbar = c2 t, dbar_prev = c1 t  (xbar passes through)
__global__ void k_cheb_adjoint(int64_t n, float c1, float c2, const float* __restrict__ xbar, float* __restrict__ dbar,
                               float* __restrict__ zbar) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float t = dbar[i] + xbar[i];
        zbar[i] = c2 * t;
        dbar[i] = c1 * t;
    }
}

__global__ __launch_bounds__(kBlock) void k_dot2_partials(int64_t n, const float* __restrict__ a,
                                                          const float* __restrict__ b, const float* __restrict__ c,
                                                          const float* __restrict__ d, double* partials) {
    double t[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        t[0] += (double)a[i] * b[i];
        t[1] += (double)c[i] * d[i];
    }
    block_partials<2>(t, partials);
}

__global__ void k_fill(int64_t n, float v, float* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

// Deterministic pseudo-random start vector in [-1, 1] (integer hash of the index).
__global__ void k_fill_hash(int64_t n, float* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t h = (uint32_t)(i * 2654435761ull) ^ 0x9E3779B9u;
        h ^= h >> 16;
        h *= 0x7feb352du;
        h ^= h >> 15;
        h *= 0x846ca68bu;
        h ^= h >> 16;
        x[i] = (float)h * (2.0f / 4294967296.0f) - 1.0f;
    }
}

__global__ void k_copy(int64_t n, const float* __restrict__ a, float* __restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void k_scal(int64_t n, float a, float* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= a;
}

// Chebyshev step: d = c1 d + c2 z; x += d   (z = D^-1 r or the preconditioned r)
__global__ void k_cheb_step(int64_t n, const float* __restrict__ dinv, const float* __restrict__ r, float c1,
                            float c2, float* __restrict__ d, float* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float z = dinv ? dinv[i] * r[i] : r[i];
        const float e = c1 * d[i] + c2 * z;
        d[i] = e;
        x[i] += e;
    }
}

// Chebyshev coefficients of one level from its power-iteration estimate
// lambda = sqrt(nrm2[0]), the device twin of cheb_coefficients; lambda is
// kept in lam_out.  A broken-down estimate (not finite or not positive)
// writes zero coefficients (no smoothing) and is reported when the host
// reads the bounds back.
// Power iteration helpers.  x is kept at the scale of the previous step:
// step k computes w = M A (x / |x|), |w|^2 (the estimate of lambda_max^2 of
// M A), and stores x = w.  M is the 3x3 block-Jacobi (binv), point-Jacobi
// (dinv) or identity (both null: y already preconditioned) smoother.

// The second half of a step after y = A x: w = s M y with s = 1 / sqrt(*nrm_prev)
// (1 on the first step, whose x is the raw start), x = w, partial |w|^2 per
// workgroup (k_power_fin adds them).
__global__ __launch_bounds__(kBlock) void k_power_finish(int n_nodes, const float* __restrict__ binv,
                                                         const float* __restrict__ dinv, const float* __restrict__ y,
                                                         float* __restrict__ x, const double* __restrict__ nrm_prev,
                                                         double* partials) {
    const float s = nrm_prev ? (float)(1.0 / sqrt(*nrm_prev)) : 1.f;
    double t[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_nodes; i += gridDim.x * blockDim.x) {
        const float r0 = y[3 * i], r1 = y[3 * i + 1], r2 = y[3 * i + 2];
        float z0 = r0, z1 = r1, z2 = r2;
        if (binv) {
            float m[9];
            for (int q = 0; q < 9; ++q) m[q] = binv[(int64_t)q * n_nodes + i];
            z0 = m[0] * r0 + m[1] * r1 + m[2] * r2;
            z1 = m[3] * r0 + m[4] * r1 + m[5] * r2;
            z2 = m[6] * r0 + m[7] * r1 + m[8] * r2;
        } else if (dinv) {
            z0 *= dinv[3 * i];
            z1 *= dinv[3 * i + 1];
            z2 *= dinv[3 * i + 2];
        }
        z0 *= s;
        z1 *= s;
        z2 *= s;
        x[3 * i] = z0;
        x[3 * i + 1] = z1;
        x[3 * i + 2] = z2;
        t[0] += (double)z0 * z0 + (double)z1 * z1 + (double)z2 * z2;
    }
    block_partials<1>(t, partials);
}

// *out = sum of the step's partials, in workgroup order
__global__ __launch_bounds__(kBlock) void k_power_fin(const double* __restrict__ partials, int nparts,
                                                      double* __restrict__ out) {
    double t[1] = {0.0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) t[0] += partials[b];
    block_sum<1>(t);
    if (threadIdx.x == 0) *out = t[0];
}

__global__ void k_cheb_coef(const double* __restrict__ nrm2, double lo_frac, double hi_frac, int degree,
                            float* __restrict__ out, double* __restrict__ lam_out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double lam = sqrt(nrm2[0]);
    *lam_out = lam;
    if (!(lam > 0.0) || !isfinite(lam)) {
        for (int k = 0; k < 2 * degree; ++k) out[k] = 0.f;
        return;
    }
    const double hi = lam * hi_frac, lo = lam * lo_frac;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho0 = 1.0 / sigma;
    out[0] = 0.f;
    out[1] = (float)(1.0 / theta);
    for (int k = 1; k < degree; ++k) {
        const double rho1 = 1.0 / (2.0 * sigma - rho0);
        out[2 * k] = (float)(rho1 * rho0);
        out[2 * k + 1] = (float)(2.0 * rho1 / delta);
        rho0 = rho1;
    }
}

__global__ void k_scale(int64_t n, const float* __restrict__ a, float s, float* __restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = s * a[i];
}

// Sum the partials of `nv` interleaved values into out[0..nv).
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double* __restrict__ partials, int nparts, int nv,
                                                         double* __restrict__ out) {
    double t[4] = {0, 0, 0, 0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x)
        for (int i = 0; i < nv; ++i) t[i] += partials[(size_t)nv * b + i];
    block_sum<4>(t);
    if (threadIdx.x == 0)
        for (int i = 0; i < nv; ++i) out[i] = t[i];
}

// Coarsest level: BSR (FP32) -> dense FP64 row-major.
__global__ void k_bsr_to_dense(int n_rows, const int32_t* __restrict__ rp, const int32_t* __restrict__ bc,
                               const float* __restrict__ K, double* __restrict__ D) {
    int r = blockIdx.x;
    if (r >= n_rows) return;
    int s = rp[r], nb = rp[r + 1] - s;
    int n = 3 * n_rows;
    for (int t = threadIdx.x; t < 9 * nb; t += blockDim.x) {
        int c = t / nb, k = t % nb;
        int i = c / 3, j = c % 3;
        D[(int64_t)(3 * r + i) * n + 3 * bc[s + k] + j] = (double)K[9 * (int64_t)s + t];
    }
}

// Same elimination with the whole matrix resident in LDS (n*n*8 <= 64 KiB).
__global__ __launch_bounds__(1024) void k_gauss_jordan_lds(int n, double* __restrict__ Ag) {
    extern __shared__ double sm[];
    double* A = sm;
    double* colk = sm + (size_t)n * n;
    for (int64_t t = threadIdx.x; t < (int64_t)n * n; t += blockDim.x) A[t] = Ag[t];
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) colk[i] = A[i * n + k];
        __syncthreads();
        const double ip = 1.0 / colk[k];
        for (int j = threadIdx.x; j < n; j += blockDim.x) A[k * n + j] = (j == k ? 1.0 : A[k * n + j]) * ip;
        __syncthreads();
        for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
            int i = t / n, j = t % n;
            if (i == k) continue;
            double aij = (j == k) ? 0.0 : A[t];
            A[t] = aij - colk[i] * A[k * n + j];
        }
        __syncthreads();
    }
    for (int64_t t = threadIdx.x; t < (int64_t)n * n; t += blockDim.x) Ag[t] = A[t];
}

// Blocked in-place Gauss-Jordan inverse of an SPD matrix (no pivoting),
// kGjBs pivots per step K (columns k0 .. k0 + kb - 1), two launches per step:
//   k_gj_pivot: P = A_KK^-1, then the pivot row panel A_{K,:} <- P A_{K,:}
//               (A_KK <- P);
//   k_gj_update: every other row i, A_ij <- (j in K ? 0 : A_ij) - sum_l A_il(old) (new A_K)_lj.
// This is the scalar elimination applied kb pivots at a time; A is streamed
// once per step by many workgroups instead of once per pivot by one.
constexpr int kGjBs = 32;

// one workgroup of 256 threads
__global__ __launch_bounds__(256) void k_gj_pivot(int n, int k0, double* __restrict__ A) {
    __shared__ double P[kGjBs * kGjBs];
    const int kb = min(kGjBs, n - k0);
    const int tid = threadIdx.x;
    for (int t = tid; t < kb * kb; t += blockDim.x) P[t] = A[(int64_t)(k0 + t / kb) * n + k0 + t % kb];
    __syncthreads();
    // P <- P^-1 in one wavefront without barriers: lane j < 32 holds column j,
    // padded with the identity to 32 x 32; the pivot column of each step is
    // read from lane kk by shuffles
    if (tid < 64) {
        const int j = tid;
        double a[kGjBs];
#pragma unroll
        for (int i = 0; i < kGjBs; ++i) a[i] = (j < kb && i < kb) ? P[i * kb + j] : (i == j ? 1.0 : 0.0);
        for (int kk = 0; kk < kb; ++kk) {
            double akk = 0.0;
#pragma unroll
            for (int i = 0; i < kGjBs; ++i) akk = (i == kk) ? a[i] : akk;
            const double ip = 1.0 / __shfl(akk, kk);
            const double rk = (j == kk ? 1.0 : akk) * ip;
#pragma unroll
            for (int i = 0; i < kGjBs; ++i) {
                const double ci = __shfl(a[i], kk);
                a[i] = (i == kk) ? rk : (j == kk ? 0.0 : a[i]) - ci * rk;
            }
        }
#pragma unroll
        for (int i = 0; i < kGjBs; ++i)
            if (j < kb && i < kb) P[i * kb + j] = a[i];
    }
    __syncthreads();
    // row panel, one column per thread (each thread reads its column before
    // writing it)
    for (int j = tid; j < n; j += blockDim.x) {
        if (j >= k0 && j < k0 + kb) {
            for (int m = 0; m < kb; ++m) A[(int64_t)(k0 + m) * n + j] = P[m * kb + (j - k0)];
        } else {
            double col[kGjBs];
#pragma unroll
            for (int l = 0; l < kGjBs; ++l) col[l] = l < kb ? A[(int64_t)(k0 + l) * n + j] : 0.0;
            for (int m = 0; m < kb; ++m) {
                double acc = 0.0;
#pragma unroll
                for (int l = 0; l < kGjBs; ++l)
                    if (l < kb) acc += P[m * kb + l] * col[l];
                A[(int64_t)(k0 + m) * n + j] = acc;
            }
        }
    }
}

// one workgroup (256 threads) per row outside the pivot block
__global__ __launch_bounds__(256) void k_gj_update(int n, int k0, double* __restrict__ A) {
    __shared__ double c[kGjBs];
    const int kb = min(kGjBs, n - k0);
    const int i = (int)blockIdx.x < k0 ? (int)blockIdx.x : (int)blockIdx.x + kb;
    if ((int)threadIdx.x < kb) c[threadIdx.x] = A[(int64_t)i * n + k0 + threadIdx.x];
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const bool in_k = (j >= k0 && j < k0 + kb);
        double acc = in_k ? 0.0 : A[(int64_t)i * n + j];
#pragma unroll
        for (int l = 0; l < kGjBs; ++l)
            if (l < kb) acc -= c[l] * A[(int64_t)(k0 + l) * n + j];
        A[(int64_t)i * n + j] = acc;
    }
}

// x = Ainv b (FP64 accumulation), one wave per row.
__global__ __launch_bounds__(256) void k_dense_gemv(int n, const double* __restrict__ Ainv, const float* __restrict__ b,
                                                    float* __restrict__ x) {
    int row = blockIdx.x * 4 + threadIdx.x / 64;
    int lane = threadIdx.x % 64;
    if (row >= n) return;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += Ainv[(int64_t)row * n + j] * (double)b[j];
    s = wave_sum(s);
    if (lane == 0) x[row] = (float)s;
}

// FCG helpers --------------------------------------------------------------
// sv layout: 0 pq, 1 pr, 2 rr, 3 rr0, 4 alpha, 5 beta, 6 zq, 7 tol2, 8 it, 9 max_it, 10 done
enum { SV_PQ = 0, SV_PR, SV_RR, SV_RR0, SV_ALPHA, SV_BETA, SV_ZQ, SV_TOL2, SV_IT, SV_MAXIT, SV_DONE, SV_N };

__global__ __launch_bounds__(kBlock) void k_dot2(int64_t n, const float* __restrict__ a, const float* __restrict__ b,
                                                 const float* __restrict__ c, const float* __restrict__ d,
                                                 double* partials, const double* __restrict__ sv) {
    if (sv[SV_DONE] != 0.0) return;
    double t[2] = {0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        t[0] += (double)a[i] * b[i];
        t[1] += (double)c[i] * d[i];
    }
    block_partials<2>(t, partials);
}

// op 0: pq, pr -> alpha ; op 1: rr -> done test ; op 2: zq -> beta ; op 3: rr0 init
__global__ __launch_bounds__(kBlock) void k_fcg_fin(const double* __restrict__ partials, int nparts, int op,
                                                    double* __restrict__ sv) {
    if (op != 3 && sv[SV_DONE] != 0.0) return;
    const int nv = (op == 0 || op == 2) ? 2 : 1;
    double t[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        t[0] += partials[(size_t)nv * b];
        if (nv == 2) t[1] += partials[(size_t)nv * b + 1];
    }
    block_sum<2>(t);
    if (threadIdx.x != 0) return;
    if (op == 0) {
        sv[SV_PQ] = t[0];
        sv[SV_PR] = t[1];
        sv[SV_ALPHA] = t[1] / t[0];
    } else if (op == 1) {
        sv[SV_RR] = t[0];
        sv[SV_IT] += 1.0;
        if (t[0] <= sv[SV_TOL2] * sv[SV_RR0] || sv[SV_IT] >= sv[SV_MAXIT] || !(t[0] == t[0])) sv[SV_DONE] = 1.0;
    } else if (op == 2) {
        sv[SV_ZQ] = t[0];
        sv[SV_BETA] = -t[0] / sv[SV_PQ];
    } else {
        sv[SV_RR0] = t[0];
        sv[SV_RR] = t[0];
        sv[SV_IT] = 0.0;
        sv[SV_DONE] = (t[0] == 0.0) ? 1.0 : 0.0;
    }
}

// x += alpha p; r -= alpha q; partial r.r
__global__ __launch_bounds__(kBlock) void k_fcg_update(int64_t n, float* __restrict__ x, float* __restrict__ r,
                                                       const float* __restrict__ p, const float* __restrict__ q,
                                                       double* partials, const double* __restrict__ sv) {
    if (sv[SV_DONE] != 0.0) return;
    const float alpha = (float)sv[SV_ALPHA];
    double t[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        float ri = r[i] - alpha * q[i];
        r[i] = ri;
        t[0] += (double)ri * ri;
    }
    block_partials<1>(t, partials);
}

// p = z + beta p
__global__ void k_fcg_dir(int64_t n, const float* __restrict__ z, float* __restrict__ p,
                          const double* __restrict__ sv) {
    if (sv[SV_DONE] != 0.0) return;
    const float beta = (float)sv[SV_BETA];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
}

// x = 0, r = b, partial b.b
__global__ __launch_bounds__(kBlock) void k_fcg_init(int64_t n, const float* __restrict__ b, float* __restrict__ x,
                                                     float* __restrict__ r, double* partials) {
    double t[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float v = b[i];
        x[i] = 0.f;
        r[i] = v;
        t[0] += (double)v * v;
    }
    block_partials<1>(t, partials);
}

inline int nb_for(int64_t n, int bs = 256) { return (int)std::min<int64_t>((n + bs - 1) / bs, 4096); }

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------

// 27-point node-block pattern of a structured level in natural order, the
// element slot map and the parity colouring.
void build_structured_pattern(Level& lv, std::vector<int32_t>& kdiag, std::vector<int32_t>& conn,
                              std::vector<uint8_t>& slot, std::vector<int32_t>& elems) {
    const int nx = lv.n[0], ny = lv.n[1], nz = lv.n[2];
    const int mx = nx + 1, my = ny + 1, mz = nz + 1;
    lv.n_nodes = (int64_t)mx * my * mz;
    lv.n_dof = 3 * lv.n_nodes;
    lv.n_el = (int64_t)nx * ny * nz;
    lv.h_rp.assign(lv.n_nodes + 1, 0);
    lv.h_bc.clear();
    kdiag.assign(lv.n_nodes, 0);
    for (int k = 0; k < mz; ++k)
        for (int j = 0; j < my; ++j)
            for (int i = 0; i < mx; ++i) {
                int64_t r = i + (int64_t)mx * (j + (int64_t)my * k);
                for (int dk = -1; dk <= 1; ++dk)
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int di = -1; di <= 1; ++di) {
                            int ii = i + di, jj = j + dj, kk = k + dk;
                            if (ii < 0 || jj < 0 || kk < 0 || ii >= mx || jj >= my || kk >= mz) continue;
                            int64_t c = ii + (int64_t)mx * (jj + (int64_t)my * kk);
                            if (c == r) kdiag[r] = (int32_t)(lv.h_bc.size() - lv.h_rp[r]);
                            lv.h_bc.push_back((int32_t)c);
                        }
                lv.h_rp[r + 1] = (int32_t)lv.h_bc.size();
            }
    lv.nnzb = (int64_t)lv.h_bc.size();
    conn.assign(8 * lv.n_el, 0);
    slot.assign(64 * lv.n_el, 0);
    static const int cor[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int64_t e = i + (int64_t)nx * (j + (int64_t)ny * k);
                for (int a = 0; a < 8; ++a)
                    conn[8 * e + a] = (int32_t)((i + cor[a][0]) + (int64_t)mx * ((j + cor[a][1]) + (int64_t)my * (k + cor[a][2])));
                for (int a = 0; a < 8; ++a) {
                    int64_t r = conn[8 * e + a];
                    const int32_t* row = lv.h_bc.data() + lv.h_rp[r];
                    int nbk = lv.h_rp[r + 1] - lv.h_rp[r];
                    for (int b = 0; b < 8; ++b)
                        slot[64 * e + 8 * a + b] = (uint8_t)(std::lower_bound(row, row + nbk, conn[8 * e + b]) - row);
                }
            }
    // 8-colour parity colouring
    lv.color_ptr.assign(9, 0);
    elems.resize(lv.n_el);
    std::vector<int> cnt(8, 0);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) cnt[(i & 1) + 2 * (j & 1) + 4 * (k & 1)]++;
    for (int c = 0; c < 8; ++c) lv.color_ptr[c + 1] = lv.color_ptr[c] + cnt[c];
    std::vector<int> fill(lv.color_ptr.begin(), lv.color_ptr.end() - 1);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                elems[fill[(i & 1) + 2 * (j & 1) + 4 * (k & 1)]++] = (int32_t)(i + (int64_t)nx * (j + (int64_t)ny * k));
}

int apply_level(Multigrid* mg, hipStream_t st, int l, const float* x, float* y) {
    Level& lv = mg->L[l];
    if (l == 0) launch_sell3<float>(st, 0, mg->fine, lv.d_val, x, y, nullptr, nullptr, nullptr);
    else launch_spmv3<float>(st, 0, (int)lv.n_nodes, lv.d_rp, lv.d_bc, lv.d_val, x, y, nullptr, nullptr, nullptr);
    return ASEND_OK;
}

// The operator the cycle itself applies on level l: on the finest level the
// FP16 copy when it is selected (smoothing and restriction residuals, and
// the power iterations that bound the smoother on it).
void apply_cycle_op(Multigrid* mg, hipStream_t st, int l, const float* x, float* y) {
    if (l == 0 && cycle_half(mg)) launch_sell3_h(st, 0, mg->fine, mg->val16, mg->dinv16, x, y, nullptr);
    else apply_level(mg, st, l, x, y);
}

void residual_level(Multigrid* mg, hipStream_t st, int l, const float* b, const float* x, float* r) {
    Level& lv = mg->L[l];
    if (l == 0 && cycle_half(mg)) launch_sell3_h(st, 3, mg->fine, mg->val16, mg->dinv16, x, r, b);
    else if (l == 0) launch_sell3<float>(st, 3, mg->fine, lv.d_val, x, r, b, nullptr, nullptr);
    else launch_spmv3<float>(st, 3, (int)lv.n_nodes, lv.d_rp, lv.d_bc, lv.d_val, x, r, b, nullptr, nullptr);
}

// Chebyshev smoothing of degree mg->cheb_degree on level l with the
// preconditioner M = 3x3 block Jacobi (default) or point Jacobi; the finest
// level uses the learned factor P = L L^T instead when it is selected.  The
// step coefficients live in device memory (coef_slot).
void smooth(Multigrid* mg, hipStream_t st, int l, const float* b, float* x, bool zero_start) {
    Level& lv = mg->L[l];
    const bool learned = (l == 0 && mg->fine_smoother == 1);
    const bool block = !learned && mg->smoother_kind == 1;
    const float* dinv = (learned || block) ? nullptr : lv.d_dinv;
    const float* coef = mg->d_coef + coef_slot(mg, learned ? (int)mg->L.size() : l, 0);
    const int nblk = nb_for(lv.n_nodes);
    for (int k = 0; k < mg->cheb_degree; ++k) {
        const bool first = (k == 0 && zero_start);
        const float* r = b;
        if (!first) {
            residual_level(mg, st, l, b, x, lv.r);
            r = lv.r;
        }
        if (block) {
            if (first)
                hipLaunchKernelGGL(k_cheb_block<true>, dim3(nblk), dim3(256), 0, st, (int)lv.n_nodes, lv.binv, r,
                                   coef + 2 * k, lv.d, x);
            else
                hipLaunchKernelGGL(k_cheb_block<false>, dim3(nblk), dim3(256), 0, st, (int)lv.n_nodes, lv.binv, r,
                                   coef + 2 * k, lv.d, x);
            continue;
        }
        if (learned && cycle_half(mg) && mg->factor.Ls16) {
            // FP16 factor, Chebyshev step fused into its second product
            launch_factor_cheb16(st, mg->factor, r, mg->ft, coef + 2 * k, lv.d, x, first);
            continue;
        }
        if (learned) {
            launch_factor_apply(st, mg->factor, r, mg->ft, mg->fz);
            r = mg->fz;
        }
        if (first)
            hipLaunchKernelGGL(k_cheb_dof<true>, dim3(nblk), dim3(256), 0, st, lv.n_dof, dinv, r, coef + 2 * k, lv.d,
                               x);
        else
            hipLaunchKernelGGL(k_cheb_dof<false>, dim3(nblk), dim3(256), 0, st, lv.n_dof, dinv, r, coef + 2 * k, lv.d,
                               x);
    }
}

void restrict_to(Multigrid* mg, hipStream_t st, int l, const float* rf, float* bc) {
    Level& c = mg->L[l + 1];
    Level& f = mg->L[l];
    const int32_t* frank = (l == 0) ? mg->d_rank0 : nullptr;
    int64_t nnodes = c.n_nodes;
    hipLaunchKernelGGL(k_restrict3, dim3((unsigned)((nnodes + 255) / 256)), dim3(256), 0, st, c.n[0], c.n[1], c.n[2],
                       f.n[0], f.n[1], c.dax.ri[0], c.dax.rw[0], c.dax.ri[1], c.dax.rw[1], c.dax.ri[2], c.dax.rw[2],
                       frank, rf, bc);
}

void prolong_add(Multigrid* mg, hipStream_t st, int l, const float* xc, float* xf) {
    Level& c = mg->L[l + 1];
    Level& f = mg->L[l];
    const int32_t* frank = (l == 0) ? mg->d_rank0 : nullptr;
    int64_t nnodes = f.n_nodes;
    hipLaunchKernelGGL(k_prolong_add3, dim3((unsigned)((nnodes + 255) / 256)), dim3(256), 0, st, f.n[0], f.n[1],
                       f.n[2], c.n[0], c.n[1], c.dax.pi[0], c.dax.pw[0], c.dax.pi[1], c.dax.pw[1], c.dax.pi[2],
                       c.dax.pw[2], frank, xc, xf);
}

// One V-cycle: x = M b on level l (x overwritten).
void vcycle(Multigrid* mg, hipStream_t st, int l, const float* b, float* x) {
    const int nl = (int)mg->L.size();
    if (l == nl - 1) {
        hipLaunchKernelGGL(k_dense_gemv, dim3((mg->nco + 3) / 4), dim3(256), 0, st, mg->nco, mg->d_Ainv, b, x);
        return;
    }
    Level& lv = mg->L[l];
    Level& cv = mg->L[l + 1];
    smooth(mg, st, l, b, x, true);
    residual_level(mg, st, l, b, x, lv.r);
    restrict_to(mg, st, l, lv.r, cv.b);
    vcycle(mg, st, l + 1, cv.b, cv.x);
    prolong_add(mg, st, l, cv.x, x);
    smooth(mg, st, l, b, x, false);
}

// kPowerIts steps of the power iteration for lambda_max(M A) on level l,
// enqueued on st with the given reduction scratch; ||M A x||^2 of the last
// step lands in nrm2[0].
void enqueue_power(Multigrid* mg, hipStream_t st, int l, bool learned, const PowerScratch& ps) {
    Level& lv = mg->L[l];
    const int64_t n = lv.n_dof;
    const bool block = !learned && mg->smoother_kind == 1;
    const float* binv = block ? lv.binv : nullptr;
    const float* dinv = (learned || block) ? nullptr : lv.d_dinv;
    hipLaunchKernelGGL(k_fill_hash, dim3(nb_for(n)), dim3(256), 0, st, n, lv.x);
    const int nblk = std::min<int>(kVecBlocks, (int)((lv.n_nodes + kBlock - 1) / kBlock));
    for (int it = 0; it < kPowerIts; ++it) {
        apply_cycle_op(mg, st, l, lv.x, lv.r);
        const float* y = lv.r;
        if (learned) {
            if (cycle_half(mg) && mg->factor.Ls16) launch_factor_apply16(st, mg->factor, lv.r, mg->ft, mg->fz);
            else launch_factor_apply(st, mg->factor, lv.r, mg->ft, mg->fz);
            y = mg->fz;
        }
        const double* prev = it == 0 ? nullptr : ps.pp + (it - 1) % 2;
        double* out = it == kPowerIts - 1 ? ps.out : ps.pp + it % 2;
        hipLaunchKernelGGL(k_power_finish, dim3(nblk), dim3(kBlock), 0, st, (int)lv.n_nodes, binv, dinv, y, lv.x, prev,
                           ps.partials);
        hipLaunchKernelGGL(k_power_fin, dim3(1), dim3(kBlock), 0, st, ps.partials, nblk, out);
    }
}

// lambda_max of M A by power iteration (M = the level's Jacobi smoother, or
// the learned factor on the finest level when `learned`), read back here.
int estimate_lmax(Multigrid* mg, hipStream_t st, int l, bool learned, double* out) {
    // deterministic pseudo-random start vector (a smooth start such as the
    // constant vector lies near the rigid-body modes and converges slowly);
    // kPowerIts steps without host round trips, ||M A x|| with ||x|| = 1 of
    // the last step is the estimate
    enqueue_power(mg, st, l, learned, mg->pw_main);
    MGCHK(hipGetLastError());
    MGCHK(hipMemcpyAsync(mg->h_sv, mg->sv + SV_N, sizeof(double), hipMemcpyDeviceToHost, st));
    MGCHK(hipStreamSynchronize(st));
    const double lam = std::sqrt(mg->h_sv[0]);
    if (!(lam > 0.0) || !std::isfinite(lam)) return mgfail(mg, ASEND_ERR_NUMERIC, "power iteration broke down");
    *out = lam;
    return ASEND_OK;
}

void free_level(Level& lv) {
    fr(lv.o_rp);
    fr(lv.o_bc);
    fr(lv.o_val);
    fr(lv.o_dinv);
    fr(lv.binv);
    fr(lv.d_kdiag);
    fr(lv.d_conn);
    fr(lv.d_slot);
    fr(lv.d_Ke);
    fr(lv.d_elems);
    fr(lv.d_pidx);
    fr(lv.d_pval);
    fr(lv.x);
    fr(lv.b);
    fr(lv.r);
    fr(lv.d);
    for (int a = 0; a < 3; ++a) {
        fr(lv.dax.pi[a]);
        fr(lv.dax.pw[a]);
        fr(lv.dax.ri[a]);
        fr(lv.dax.rw[a]);
        fr(lv.dax.X[a]);
    }
}

}  // namespace

// ===========================================================================

int mg_create(Multigrid** out, int nx, int ny, int nz, const int32_t* rank_of_grid, const double* d_xyz_r,
              const int32_t* d_conn_r, const Sell3& fine, int64_t n_nodes, int64_t n_el) {
    Multigrid* mg = new Multigrid();
    mg->fine = fine;
    *out = mg;
    mg->d_xyz_r = d_xyz_r;
    mg->d_conn_r = d_conn_r;
    mg->n_el0 = n_el;
    mg->h_rank0.assign(rank_of_grid, rank_of_grid + n_nodes);
    MGCHK(up(&mg->o_rank0, rank_of_grid, (size_t)n_nodes));
    mg->d_rank0 = mg->o_rank0;
    // level 0 = fine grid (operator borrowed at setup)
    Level l0;
    l0.n[0] = nx;
    l0.n[1] = ny;
    l0.n[2] = nz;
    l0.n_nodes = n_nodes;
    l0.n_dof = 3 * n_nodes;
    l0.n_el = n_el;
    mg->L.push_back(l0);
    // coarse levels
    while (true) {
        Level& f = mg->L.back();
        if (f.n_nodes <= kMaxCoarseNodes || (f.n[0] == 1 && f.n[1] == 1 && f.n[2] == 1)) break;
        // Trilinear coarse elements lock in bending once a direction that is
        // resolved by two or more elements drops to one; the Galerkin operator
        // of such a level is far too stiff and the coarse correction stalls
        // (a 200x40x40 cantilever needs 79 instead of 43 iterations).  Stop
        // above it while the dense coarsest inverse stays small.
        bool locks = false;
        for (int a = 0; a < 3; ++a) locks |= (f.n[a] >= 2 && (f.n[a] + 1) / 2 < 2);
        if (locks && f.n_nodes <= kMaxCoarseDenseNodes) break;
        Level c;
        for (int a = 0; a < 3; ++a) {
            c.ax[a] = make_axis(f.n[a]);
            c.n[a] = c.ax[a].nc;
        }
        std::vector<int32_t> kdiag, conn, elems;
        std::vector<uint8_t> slot;
        build_structured_pattern(c, kdiag, conn, slot, elems);
        c.owns_op = true;
        MGCHK(up(&c.o_rp, c.h_rp.data(), c.h_rp.size()));
        MGCHK(up(&c.o_bc, c.h_bc.data(), c.h_bc.size()));
        MGCHK(up(&c.d_kdiag, kdiag.data(), kdiag.size()));
        MGCHK(up(&c.d_conn, conn.data(), conn.size()));
        MGCHK(up(&c.d_slot, slot.data(), slot.size()));
        MGCHK(up(&c.d_elems, elems.data(), elems.size()));
        MGCHK(hipMalloc((void**)&c.o_val, 9 * c.nnzb * sizeof(float)));
        MGCHK(hipMalloc((void**)&c.o_dinv, c.n_dof * sizeof(float)));
        MGCHK(hipMalloc((void**)&c.d_Ke, 576 * c.n_el * sizeof(float)));
        c.d_rp = c.o_rp;
        c.d_bc = c.o_bc;
        c.d_val = c.o_val;
        c.d_dinv = c.o_dinv;
        for (int a = 0; a < 3; ++a) {
            MGCHK(up(&c.dax.pi[a], c.ax[a].pi.data(), c.ax[a].pi.size()));
            MGCHK(up(&c.dax.pw[a], c.ax[a].pw.data(), c.ax[a].pw.size()));
            MGCHK(up(&c.dax.ri[a], c.ax[a].ri.data(), c.ax[a].ri.size()));
            MGCHK(up(&c.dax.rw[a], c.ax[a].rw.data(), c.ax[a].rw.size()));
            MGCHK(up(&c.dax.X[a], c.ax[a].X.data(), c.ax[a].X.size()));
        }
        mg->L.push_back(c);
        if (mg->L.size() > 16) break;
    }
    if (mg->L.size() < 2) return mgfail(mg, ASEND_ERR_UNSUPPORTED, "grid too small for a multigrid hierarchy");
    for (auto& lv : mg->L) {
        MGCHK(hipMalloc((void**)&lv.x, lv.n_dof * sizeof(float)));
        MGCHK(hipMalloc((void**)&lv.b, lv.n_dof * sizeof(float)));
        MGCHK(hipMalloc((void**)&lv.r, lv.n_dof * sizeof(float)));
        MGCHK(hipMalloc((void**)&lv.d, lv.n_dof * sizeof(float)));
        MGCHK(hipMalloc((void**)&lv.binv, 9 * lv.n_nodes * sizeof(float)));
    }
    mg->nco = (int)mg->L.back().n_dof;
    mg->h_coef.assign(2 * kChebMaxDegree * (mg->L.size() + 1), 0.f);
    MGCHK(hipMalloc((void**)&mg->d_coef, mg->h_coef.size() * sizeof(float)));
    MGCHK(hipMalloc((void**)&mg->d_Ainv, (size_t)mg->nco * mg->nco * sizeof(double)));
    const int64_t n0 = mg->L[0].n_dof;
    MGCHK(hipMalloc((void**)&mg->r, n0 * sizeof(float)));
    MGCHK(hipMalloc((void**)&mg->z, n0 * sizeof(float)));
    MGCHK(hipMalloc((void**)&mg->p, n0 * sizeof(float)));
    MGCHK(hipMalloc((void**)&mg->q, n0 * sizeof(float)));
    MGCHK(hipMalloc((void**)&mg->partials, 2 * 4096 * sizeof(double)));
    MGCHK(hipMalloc((void**)&mg->sv, 32 * sizeof(double)));
    MGCHK(hipMalloc((void**)&mg->d_lam, kMaxLevels * sizeof(double)));
    MGCHK(hipMalloc((void**)&mg->partials_b, kVecBlocks * sizeof(double)));
    MGCHK(hipMalloc((void**)&mg->sv_b, 4 * sizeof(double)));
    MGCHK(hipHostMalloc((void**)&mg->h_sv_b, sizeof(double), hipHostMallocDefault));
    MGCHK(hipMalloc((void**)&mg->partials_c, kVecBlocks * sizeof(double)));
    MGCHK(hipMalloc((void**)&mg->sv_c, 4 * sizeof(double)));
    mg->pw_main.partials = mg->partials;
    mg->pw_main.out = mg->sv + SV_N;
    mg->pw_main.pp = mg->sv + SV_N + 1;
    mg->pw_side.partials = mg->partials_b;
    mg->pw_side.out = mg->sv_b;
    mg->pw_side.pp = mg->sv_b + 1;
    mg->pw_deep.partials = mg->partials_c;
    mg->pw_deep.out = mg->sv_c;
    mg->pw_deep.pp = mg->sv_c + 1;
    for (auto& sp : mg->pw_stream) MGCHK(hipStreamCreateWithFlags(&sp, hipStreamNonBlocking));
    for (size_t l = 0; l < mg->L.size(); ++l)
        MGCHK(hipEventCreateWithFlags(&mg->ev_level[l], hipEventDisableTiming));
    for (auto& e : mg->ev_pw_done) MGCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    MGCHK(hipHostMalloc((void**)&mg->h_lam, kMaxLevels * sizeof(double), hipHostMallocDefault));
    MGCHK(hipHostMalloc((void**)&mg->h_sv, 32 * sizeof(double), hipHostMallocDefault));
    return ASEND_OK;
}

void mg_destroy(Multigrid* mg) {
    if (!mg) return;
    for (auto& lv : mg->L) free_level(lv);
    fr(mg->o_rank0);
    fr(mg->d_Ainv);
    fr(mg->r);
    fr(mg->z);
    fr(mg->p);
    fr(mg->q);
    fr(mg->partials);
    fr(mg->sv);
    fr(mg->ft);
    fr(mg->fz);
    fr(mg->partials_b);
    fr(mg->sv_b);
    fr(mg->partials_c);
    fr(mg->sv_c);
    for (auto& sp : mg->pw_stream) {
        if (sp) (void)hipStreamSynchronize(sp);
        if (sp) (void)hipStreamDestroy(sp);
        sp = nullptr;
    }
    for (auto& e : mg->ev_level)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : mg->ev_pw_done)
        if (e) (void)hipEventDestroy(e);
    if (mg->pw_exec) (void)hipGraphExecDestroy(mg->pw_exec);
    fr(mg->d_lam);
    if (mg->h_lam) (void)hipHostFree(mg->h_lam);
    if (mg->h_sv_b) (void)hipHostFree(mg->h_sv_b);
    fr(mg->d_coef);
    for (auto& g : mg->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (auto& q : mg->tv) fr(q);
    for (auto& q : mg->tz) fr(q);
    fr(mg->tx);
    fr(mg->td);
    fr(mg->txb);
    fr(mg->tdb);
    fr(mg->tb);
    fr(mg->tq);
    if (mg->h_sv) (void)hipHostFree(mg->h_sv);
    delete mg;
}

const char* mg_last_error(const Multigrid* mg) { return mg ? mg->err.c_str() : ""; }

int mg_set_fine_smoother(Multigrid* mg, int kind) {
    if (kind != 0 && kind != 1) return mgfail(mg, ASEND_ERR_ARG, "unknown smoother");
    if (kind == 1 && !mg_factor_has_bounds(mg)) return mgfail(mg, ASEND_ERR_STATE, "no learned factor set");
    mg->fine_smoother = kind;
    return ASEND_OK;
}

int mg_set_smoother(Multigrid* mg, int kind, int degree, double lo_frac, double hi_frac) {
    if ((kind != 0 && kind != 1) || degree < 1 || degree > kChebMaxDegree || !(lo_frac > 0.0) ||
        !(hi_frac > lo_frac))
        return mgfail(mg, ASEND_ERR_ARG, "bad Chebyshev smoother parameters");
    const bool rebound = (kind != mg->smoother_kind);
    mg->smoother_kind = kind;
    mg->cheb_degree = degree;
    mg->cheb_lo = lo_frac;
    mg->cheb_hi = hi_frac;
    // new Jacobi kind: the bounds of the current operator are stale
    if (rebound) mg->ready = false;
    return ASEND_OK;
}

// Block-Jacobi Chebyshev bound of the finest level, estimated on first use
// after a setup (host synchronisation: call outside graph capture).
int fine_bound(Multigrid* mg, hipStream_t st) {
    if (mg->L.size() < 2 || mg->L[0].lmax > 0.0) return ASEND_OK;
    // enqueued with the setup: its coefficients are already on the device
    if (mg->fine_pending) return ASEND_OK;
    int rc = estimate_lmax(mg, st, 0, false, &mg->L[0].lmax);
    if (rc) return rc;
    const int slot = coef_slot(mg, 0, 0);
    cheb_coefficients(mg, mg->L[0].lmax, mg->h_coef.data() + slot);
    MGCHK(hipMemcpyAsync(mg->d_coef + slot, mg->h_coef.data() + slot, 2 * kChebMaxDegree * sizeof(float),
                         hipMemcpyHostToDevice, st));
    MGCHK(hipStreamSynchronize(st));
    return ASEND_OK;
}

// Recompute the device Chebyshev coefficients from the stored bounds (after a
// change of degree or interval that keeps the Jacobi kind).
int mg_check_bounds(Multigrid* mg) {
    if (!mg->lam_pending) return ASEND_OK;
    MGCHK(hipStreamSynchronize(mg->lam_stream));
    mg->lam_pending = false;
    const size_t l_first = mg->fine_pending ? 0 : 1;
    mg->fine_pending = false;
    for (size_t l = l_first; l + 1 < mg->L.size(); ++l) {
        const double lam = mg->h_lam[l];
        if (!(lam > 0.0) || !std::isfinite(lam)) {
            mg->ready = false;
            return mgfail(mg, ASEND_ERR_NUMERIC, "power iteration broke down");
        }
        mg->L[l].lmax = lam;
    }
    return ASEND_OK;
}

int mg_refresh_coefficients(Multigrid* mg, hipStream_t st) {
    if (!mg->ready) return ASEND_OK;
    int rc = mg_check_bounds(mg);
    if (rc) return rc;
    for (size_t l = 0; l + 1 < mg->L.size(); ++l)
        if (mg->L[l].lmax > 0.0)  // the finest level's may still be pending
            cheb_coefficients(mg, mg->L[l].lmax, mg->h_coef.data() + coef_slot(mg, (int)l, 0));
    if (mg->lmax_factor > 0.0)
        cheb_coefficients(mg, mg->lmax_factor, mg->h_coef.data() + coef_slot(mg, (int)mg->L.size(), 0));
    MGCHK(hipMemcpyAsync(mg->d_coef, mg->h_coef.data(), mg->h_coef.size() * sizeof(float), hipMemcpyHostToDevice,
                         st));
    MGCHK(hipStreamSynchronize(st));
    return ASEND_OK;
}

int mg_refresh_fine(Multigrid* mg, hipStream_t st, const float* d_K32, const float* d_dinv32,
                    const int32_t* d_kdiag0) {
    if (!mg->ready) return ASEND_ERR_STATE;
    Level& l0 = mg->L[0];
    l0.d_val = d_K32;
    l0.d_dinv = d_dinv32;
    hipLaunchKernelGGL(k_binv32, dim3((unsigned)((l0.n_nodes + 255) / 256)), dim3(256), 0, st, (int)l0.n_nodes,
                       mg->fine.sptr, (const int32_t*)nullptr, d_kdiag0, d_K32, l0.binv);
    MGCHK(hipGetLastError());
    l0.lmax = 0.0;
    mg->fine_pending = false;  // a pending bound belongs to the old values
    return ASEND_OK;
}

void mg_set_fine_geometry(Multigrid* mg, int waves_per_block) { mg->fine.waves_per_block = waves_per_block; }

void mg_set_fine_half(Multigrid* mg, const __half* val16, const float* dinv32) {
    mg->val16 = val16;
    mg->dinv16 = dinv32;
}

// The finest level's smoother bound is estimated on the operator the cycle
// applies: switching it re-estimates that bound on first use.
void mg_set_cycle_half(Multigrid* mg, bool on) {
    if (mg->cycle_half != on && !mg->L.empty()) {
        mg->L[0].lmax = 0.0;
        mg->lmax_factor = 0.0;
        mg->fine_pending = false;
    }
    mg->cycle_half = on;
}

int mg_smoother_degree(const Multigrid* mg) { return mg->cheb_degree;
}

int upload_factor_coefficients(Multigrid* mg, hipStream_t st) {
    const int slot = coef_slot(mg, (int)mg->L.size(), 0);
    cheb_coefficients(mg, mg->lmax_factor, mg->h_coef.data() + slot);
    MGCHK(hipMemcpyAsync(mg->d_coef + slot, mg->h_coef.data() + slot, 2 * kChebMaxDegree * sizeof(float),
                         hipMemcpyHostToDevice, st));
    // the host copy is read by the copy engine: keep it unchanged until done
    MGCHK(hipStreamSynchronize(st));
    return ASEND_OK;
}

int mg_set_learned_factor(Multigrid* mg, hipStream_t st, const Factor& F, bool smoother_bounds) {
    if (!mg->ready) return mgfail(mg, ASEND_ERR_STATE, "multigrid not set up");
    Level& l0 = mg->L[0];
    if (F.n != l0.n_nodes) return mgfail(mg, ASEND_ERR_ARG, "factor size does not match the finest level");
    if (!mg->ft) {
        MGCHK(hipMalloc((void**)&mg->ft, l0.n_dof * sizeof(float)));
        MGCHK(hipMalloc((void**)&mg->fz, l0.n_dof * sizeof(float)));
    }
    const double keep = mg->lmax_factor;
    mg->factor = F;
    mg->has_factor = true;
    mg->lmax_factor = 0.0;
    if (smoother_bounds && mg->keep_factor_bounds && keep > 0.0) {
        mg->lmax_factor = keep;
        return upload_factor_coefficients(mg, st);
    }
    if (!smoother_bounds) return ASEND_OK;
    double lam = 0.0;
    int rc = estimate_lmax(mg, st, 0, true, &lam);
    if (rc) {
        mg->has_factor = false;
        return rc;
    }
    mg->lmax_factor = lam;
    return upload_factor_coefficients(mg, st);
}

int mg_learned_bound_begin(Multigrid* mg, hipStream_t st, const Factor& F, const float* d_K32,
                           const float* d_dinv32) {
    Level& l0 = mg->L[0];
    if (F.n != l0.n_nodes) return mgfail(mg, ASEND_ERR_ARG, "factor size does not match the finest level");
    if (!mg->ft) {
        MGCHK(hipMalloc((void**)&mg->ft, l0.n_dof * sizeof(float)));
        MGCHK(hipMalloc((void**)&mg->fz, l0.n_dof * sizeof(float)));
    }
    l0.d_val = d_K32;
    l0.d_dinv = d_dinv32;
    mg->factor = F;
    mg->has_factor = true;
    mg->lmax_factor = 0.0;
    // the finest level's vectors x, r and the factor's ft, fz are not touched
    // by the setup of the coarser levels.  The 20 steps (about 100 launches)
    // replay as one graph while the factor storage, the operator and the
    // SpMV geometry stay where they were captured.
    if (mg->use_graphs) {
        const Factor& P = mg->pw_factor;
        const bool same = mg->pw_exec && P.n == F.n && P.n_edge == F.n_edge && P.Ls == F.Ls && P.LsT == F.LsT &&
                          P.enb == F.enb && P.erev == F.erev && mg->pw_val == d_K32 &&
                          mg->pw_waves == mg->fine.waves_per_block && mg->pw_half == (int)cycle_half(mg);
        if (!same) {
            if (mg->pw_exec) (void)hipGraphExecDestroy(mg->pw_exec);
            mg->pw_exec = nullptr;
            hipGraph_t graph = nullptr;
            MGCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            enqueue_power(mg, st, 0, true, mg->pw_side);
            hipError_t e1 = hipGetLastError();
            hipError_t e2 = hipStreamEndCapture(st, &graph);
            if (e1 != hipSuccess || e2 != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return mgfail(mg, ASEND_ERR_HIP, "capturing the bound's power iteration failed");
            }
            hipError_t e3 = hipGraphInstantiate(&mg->pw_exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e3 != hipSuccess) {
                mg->pw_exec = nullptr;
                return mgfail(mg, ASEND_ERR_HIP, "instantiating the bound's power iteration failed");
            }
            mg->pw_factor = F;
            mg->pw_val = d_K32;
            mg->pw_waves = mg->fine.waves_per_block;
            mg->pw_half = (int)cycle_half(mg);
        }
        MGCHK(hipGraphLaunch(mg->pw_exec, st));
    } else {
        enqueue_power(mg, st, 0, true, mg->pw_side);
    }
    MGCHK(hipGetLastError());
    MGCHK(hipMemcpyAsync(mg->h_sv_b, mg->sv_b, sizeof(double), hipMemcpyDeviceToHost, st));
    return ASEND_OK;
}

int mg_learned_bound_end(Multigrid* mg, hipStream_t st) {
    MGCHK(hipStreamSynchronize(st));
    const double lam = std::sqrt(mg->h_sv_b[0]);
    if (!(lam > 0.0) || !std::isfinite(lam)) {
        mg->has_factor = false;
        return mgfail(mg, ASEND_ERR_NUMERIC, "power iteration broke down");
    }
    mg->lmax_factor = lam;
    return upload_factor_coefficients(mg, st);
}

bool mg_ready(const Multigrid* mg) { return mg->ready; }

bool mg_factor_has_bounds(const Multigrid* mg) { return mg->has_factor && mg->lmax_factor > 0.0; }

void mg_clear_learned_factor(Multigrid* mg) {
    mg->has_factor = false;
    mg->fine_smoother = 0;
}

void mg_keep_factor_bounds(Multigrid* mg, bool keep) { mg->keep_factor_bounds = keep; }

void mg_use_graphs(Multigrid* mg, bool on) { mg->use_graphs = on; }

// Galerkin images of the penalty on every coarse level, S_l = P^T S_{l-1} P
// per component (S_0 = the penalised fine DOFs with their multiplicities), as
// lists of (value index, weight); each setup adds sf * weight.  Built on the
// host once per penalty set, since the penalty set does not change between
// the assemblies of a design loop.
static int prepare_penalty(Multigrid* mg, int n_pen, const int64_t* h_pen, const int32_t* h_mult) {
    std::vector<int64_t> pk(h_pen, h_pen + n_pen);
    std::vector<int32_t> mk(h_mult, h_mult + n_pen);
    if (mg->pen_ready && pk == mg->pen_key && mk == mg->mult_key) return ASEND_OK;
    mg->pen_ready = false;
    Level& l0 = mg->L[0];
    std::vector<int32_t> grid_of_rank(l0.n_nodes);
    for (int64_t g = 0; g < l0.n_nodes; ++g) grid_of_rank[mg->h_rank0[g]] = (int32_t)g;
    std::vector<std::map<std::pair<int64_t, int64_t>, double>> S(3);
    for (int i = 0; i < n_pen; ++i) {
        int64_t d = h_pen[i];
        int64_t g = grid_of_rank[d / 3];
        S[d % 3][{g, g}] += (double)h_mult[i];
    }
    for (size_t l = 1; l < mg->L.size(); ++l) {
        Level& c = mg->L[l];
        Level& f = mg->L[l - 1];
        const int fmx = f.n[0] + 1, fmy = f.n[1] + 1;
        const int cmx = c.n[0] + 1, cmy = c.n[1] + 1;
        // Galerkin penalty: S_c = P^T S_f P per component
        std::vector<std::map<std::pair<int64_t, int64_t>, double>> Sc(3);
        for (int comp = 0; comp < 3; ++comp)
            for (auto& kv : S[comp]) {
                int64_t g1 = kv.first.first, g2 = kv.first.second;
                int i1 = (int)(g1 % fmx), j1 = (int)((g1 / fmx) % fmy), k1 = (int)(g1 / ((int64_t)fmx * fmy));
                int i2 = (int)(g2 % fmx), j2 = (int)((g2 / fmx) % fmy), k2 = (int)(g2 / ((int64_t)fmx * fmy));
                for (int a1 = 0; a1 < 8; ++a1) {
                    int I1 = c.ax[0].pi[2 * i1 + (a1 & 1)], J1 = c.ax[1].pi[2 * j1 + ((a1 >> 1) & 1)],
                        K1 = c.ax[2].pi[2 * k1 + ((a1 >> 2) & 1)];
                    if (I1 < 0 || J1 < 0 || K1 < 0) continue;
                    double w1 = (double)c.ax[0].pw[2 * i1 + (a1 & 1)] * c.ax[1].pw[2 * j1 + ((a1 >> 1) & 1)] *
                                c.ax[2].pw[2 * k1 + ((a1 >> 2) & 1)];
                    int64_t N1 = I1 + (int64_t)cmx * (J1 + (int64_t)cmy * K1);
                    for (int a2 = 0; a2 < 8; ++a2) {
                        int I2 = c.ax[0].pi[2 * i2 + (a2 & 1)], J2 = c.ax[1].pi[2 * j2 + ((a2 >> 1) & 1)],
                            K2 = c.ax[2].pi[2 * k2 + ((a2 >> 2) & 1)];
                        if (I2 < 0 || J2 < 0 || K2 < 0) continue;
                        double w2 = (double)c.ax[0].pw[2 * i2 + (a2 & 1)] * c.ax[1].pw[2 * j2 + ((a2 >> 1) & 1)] *
                                    c.ax[2].pw[2 * k2 + ((a2 >> 2) & 1)];
                        int64_t N2 = I2 + (int64_t)cmx * (J2 + (int64_t)cmy * K2);
                        Sc[comp][{N1, N2}] += w1 * kv.second * w2;
                    }
                }
            }
        S.swap(Sc);
        std::vector<int64_t> pidx;
        std::vector<float> pval;
        for (int comp = 0; comp < 3; ++comp)
            for (auto& kv : S[comp]) {
                int64_t N1 = kv.first.first, N2 = kv.first.second;
                const int32_t* row = c.h_bc.data() + c.h_rp[N1];
                int nbk = c.h_rp[N1 + 1] - c.h_rp[N1];
                const int32_t* it = std::lower_bound(row, row + nbk, (int32_t)N2);
                if (it == row + nbk || *it != N2) return mgfail(mg, ASEND_ERR_STATE, "penalty coupling outside pattern");
                int64_t slotk = it - row;
                pidx.push_back(9 * (int64_t)c.h_rp[N1] + (4 * comp) * nbk + slotk);
                pval.push_back((float)kv.second);
            }
        fr(c.d_pidx);
        fr(c.d_pval);
        c.n_padd = (int64_t)pidx.size();
        MGCHK(up(&c.d_pidx, pidx.data(), pidx.size()));
        MGCHK(up(&c.d_pval, pval.data(), pval.size()));
    }
    mg->pen_key.swap(pk);
    mg->mult_key.swap(mk);
    mg->pen_ready = true;
    return ASEND_OK;
}

int mg_fine_setup(Multigrid* mg, hipStream_t st, const float* d_K32, const float* d_dinv32,
                  const int32_t* d_kdiag0) {
    Level& l0 = mg->L[0];
    l0.d_val = d_K32;
    l0.d_dinv = d_dinv32;
    hipLaunchKernelGGL(k_binv32, dim3((unsigned)((l0.n_nodes + 255) / 256)), dim3(256), 0, st, (int)l0.n_nodes,
                       mg->fine.sptr, (const int32_t*)nullptr, d_kdiag0, d_K32, l0.binv);
    l0.lmax = 0.0;
    mg->fine_pending = false;
    if (mg->L.size() >= 2) {
        // the bound's own reduction scratch: the coarse levels' run alongside
        enqueue_power(mg, st, 0, false, mg->pw_side);
        hipLaunchKernelGGL(k_cheb_coef, dim3(1), dim3(64), 0, st, mg->sv_b, mg->cheb_lo, mg->cheb_hi, mg->cheb_degree,
                           mg->d_coef + coef_slot(mg, 0, 0), mg->d_lam + 0);
        mg->fine_pending = true;
    }
    MGCHK(hipGetLastError());
    return ASEND_OK;
}

int mg_setup_finish(Multigrid* mg, hipStream_t st) {
    MGCHK(hipMemcpyAsync(mg->h_lam, mg->d_lam, mg->L.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    mg->lam_pending = true;
    mg->lam_stream = st;
    return ASEND_OK;
}

int mg_setup(Multigrid* mg, hipStream_t st, const double* d_E, const double* d_nu, const float* d_K32,
             const float* d_dinv32, const int32_t* d_kdiag0, double sf, int n_pen, const int64_t* h_pen,
             const int32_t* h_mult, bool fine_elsewhere) {
    mg->ready = false;
    Level& l0 = mg->L[0];
    l0.d_val = d_K32;
    l0.d_dinv = d_dinv32;
    if (!fine_elsewhere) {
        hipLaunchKernelGGL(k_binv32, dim3((unsigned)((l0.n_nodes + 255) / 256)), dim3(256), 0, st, (int)l0.n_nodes,
                           mg->fine.sptr, (const int32_t*)nullptr, d_kdiag0, d_K32, l0.binv);
        mg->L[0].lmax = 0.0;
        mg->fine_pending = false;
    }
    {
        int rc = prepare_penalty(mg, n_pen, h_pen, h_mult);
        if (rc) return rc;
    }
    for (size_t l = 1; l < mg->L.size(); ++l) {
        Level& c = mg->L[l];
        Level& f = mg->L[l - 1];
        // element matrices
        if (l == 1) {
            hipLaunchKernelGGL(k_level1_ke, dim3((unsigned)c.n_el), dim3(64), 0, st, c.n[0], c.n[1], c.n[2], f.n[0],
                               f.n[1], c.dax.X[0], c.dax.X[1], c.dax.X[2], mg->d_conn_r, mg->d_xyz_r, d_E, d_nu,
                               c.d_Ke);
        } else {
            hipLaunchKernelGGL(k_galerkin_ke, dim3((unsigned)c.n_el), dim3(64), 0, st, c.n[0], c.n[1], c.n[2], f.n[0],
                               f.n[1], c.dax.X[0], c.dax.X[1], c.dax.X[2], f.d_Ke, c.d_Ke);
        }
        // assemble
        hipLaunchKernelGGL(k_fill, dim3(nb_for(9 * c.nnzb)), dim3(256), 0, st, 9 * c.nnzb, 0.f, c.o_val);
        for (int col = 0; col < 8; ++col) {
            int b = c.color_ptr[col], e = c.color_ptr[col + 1];
            if (e > b)
                hipLaunchKernelGGL(k_asm_ke, dim3(e - b), dim3(64), 0, st, c.d_elems + b, c.d_conn, c.d_Ke, c.o_rp,
                                   c.d_slot, c.o_val);
        }
        if (c.n_padd)
            hipLaunchKernelGGL(k_add_values, dim3((unsigned)((c.n_padd + 255) / 256)), dim3(256), 0, st, c.n_padd,
                               c.d_pidx, c.d_pval, (float)sf, c.o_val);
        hipLaunchKernelGGL(k_dinv32, dim3((unsigned)((c.n_nodes + 255) / 256)), dim3(256), 0, st, (int)c.n_nodes,
                           c.o_rp, c.d_kdiag, c.o_val, c.o_dinv);
        hipLaunchKernelGGL(k_binv32, dim3((unsigned)((c.n_nodes + 255) / 256)), dim3(256), 0, st, (int)c.n_nodes,
                           (const int32_t*)nullptr, c.o_rp, c.d_kdiag, c.o_val, c.binv);
        // smoother bound of this level (all but the coarsest), started as
        // soon as the level is assembled on a power stream of its own: level
        // 1 on the first, deeper levels in turn on the second.  The bounds
        // are turned into Chebyshev coefficients on the device, without host
        // round trips (mg_setup_finish copies them back, mg_check_bounds
        // reads them).  The finest level's is enqueued by mg_fine_setup
        // (fine_elsewhere), or else estimated on first use (fine_bound): when
        // the learned factor smooths that level, its block-Jacobi bound is
        // never needed.
        if (l + 1 < mg->L.size()) {
            const int k = (l == 1) ? 0 : 1;
            const PowerScratch& ps = (l == 1) ? mg->pw_main : mg->pw_deep;
            hipStream_t sp = mg->pw_stream[k];
            MGCHK(hipEventRecord(mg->ev_level[l], st));
            MGCHK(hipStreamWaitEvent(sp, mg->ev_level[l], 0));
            enqueue_power(mg, sp, (int)l, false, ps);
            hipLaunchKernelGGL(k_cheb_coef, dim3(1), dim3(64), 0, sp, ps.out, mg->cheb_lo, mg->cheb_hi,
                               mg->cheb_degree, mg->d_coef + coef_slot(mg, (int)l, 0), mg->d_lam + l);
        }
    }
    MGCHK(hipGetLastError());
    // coarsest: dense FP64 inverse (on the setup stream, beside the bounds)
    Level& lc = mg->L.back();
    hipLaunchKernelGGL(k_fill, dim3(nb_for((int64_t)mg->nco * mg->nco / 2)), dim3(256), 0, st,
                       (int64_t)mg->nco * mg->nco * 2, 0.f, (float*)mg->d_Ainv);
    hipLaunchKernelGGL(k_bsr_to_dense, dim3((unsigned)lc.n_nodes), dim3(64), 0, st, (int)lc.n_nodes, lc.d_rp,
                       lc.d_bc, lc.d_val, mg->d_Ainv);
    if ((size_t)mg->nco * (mg->nco + 1) * sizeof(double) <= 65536)
        hipLaunchKernelGGL(k_gauss_jordan_lds, dim3(1), dim3(1024), (size_t)mg->nco * (mg->nco + 1) * sizeof(double), st,
                           mg->nco, mg->d_Ainv);
    else
        for (int k0 = 0; k0 < mg->nco; k0 += kGjBs) {
            const int kb = std::min(kGjBs, mg->nco - k0);
            hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(256), 0, st, mg->nco, k0, mg->d_Ainv);
            if (mg->nco > kb)
                hipLaunchKernelGGL(k_gj_update, dim3((unsigned)(mg->nco - kb)), dim3(256), 0, st, mg->nco, k0,
                                   mg->d_Ainv);
        }
    MGCHK(hipGetLastError());
    // join the power streams: everything that uses the hierarchy is ordered
    // on st, with no host wait
    for (int k = 0; k < 2; ++k) {
        MGCHK(hipEventRecord(mg->ev_pw_done[k], mg->pw_stream[k]));
        MGCHK(hipStreamWaitEvent(st, mg->ev_pw_done[k], 0));
    }
    mg->ready = true;
    return ASEND_OK;
}

int mg_fcg_solve(Multigrid* mg, hipStream_t st, const float* b, float* x, double rel_tol, int64_t max_it,
                 int check_every, int64_t* it_out, int* fallback_used, int precond, FcgRefresh* ru) {
    if (!mg->ready) return mgfail(mg, ASEND_ERR_STATE, "multigrid not set up");
    if (precond == 1 && !mg->has_factor) return mgfail(mg, ASEND_ERR_STATE, "no learned factor set");
    if (precond == 0 && mg->fine_smoother == 0) {
        int rc = fine_bound(mg, st);
        if (rc) return rc;
    }
    // z = B r: one V-cycle, or the learned factor alone
    auto apply_b = [&](const float* r, float* z) {
        if (precond == 1) launch_factor_apply(st, mg->factor, r, mg->ft, z);
        else vcycle(mg, st, 0, r, z);
    };
    Level& l0 = mg->L[0];
    const int64_t n = l0.n_dof;
    const int nvb = kVecBlocks;
    if (fallback_used) *fallback_used = 0;
    // scalars
    double init[SV_N];
    std::memset(init, 0, sizeof(init));
    // with reliable updates the FP64 residual decides convergence: the
    // device-side test stops only at max_it
    init[SV_TOL2] = ru ? 0.0 : rel_tol * rel_tol;
    init[SV_MAXIT] = (double)max_it;
    std::memcpy(mg->h_sv, init, sizeof(init));
    MGCHK(hipMemcpyAsync(mg->sv, mg->h_sv, sizeof(init), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fcg_init, dim3(nvb), dim3(kBlock), 0, st, n, b, x, mg->r, mg->partials);
    hipLaunchKernelGGL(k_fcg_fin, dim3(1), dim3(kBlock), 0, st, mg->partials, nvb, 3, mg->sv);
    // z = M r ; p = z
    apply_b(mg->r, mg->z);
    hipLaunchKernelGGL(k_copy, dim3(nb_for(n)), dim3(256), 0, st, n, mg->z, mg->p);
    // one block of check_every iterations; kernels only (scalars stay on the
    // device), so the block can be captured once and replayed as a graph
    auto iterations = [&]() {
        for (int k = 0; k < check_every; ++k) {
            // q = A p ; alpha = (p.r)/(p.q)
            apply_level(mg, st, 0, mg->p, mg->q);
            hipLaunchKernelGGL(k_dot2, dim3(nvb), dim3(kBlock), 0, st, n, mg->p, mg->q, mg->p, mg->r, mg->partials,
                               mg->sv);
            hipLaunchKernelGGL(k_fcg_fin, dim3(1), dim3(kBlock), 0, st, mg->partials, nvb, 0, mg->sv);
            hipLaunchKernelGGL(k_fcg_update, dim3(nvb), dim3(kBlock), 0, st, n, x, mg->r, mg->p, mg->q, mg->partials,
                               mg->sv);
            hipLaunchKernelGGL(k_fcg_fin, dim3(1), dim3(kBlock), 0, st, mg->partials, nvb, 1, mg->sv);
            // z = M r ; beta = -(z.q)/(p.q) ; p = z + beta p
            apply_b(mg->r, mg->z);
            hipLaunchKernelGGL(k_dot2, dim3(nvb), dim3(kBlock), 0, st, n, mg->z, mg->q, mg->z, mg->q, mg->partials,
                               mg->sv);
            hipLaunchKernelGGL(k_fcg_fin, dim3(1), dim3(kBlock), 0, st, mg->partials, nvb, 2, mg->sv);
            hipLaunchKernelGGL(k_fcg_dir, dim3(nvb), dim3(kBlock), 0, st, n, mg->z, mg->p, mg->sv);
        }
    };
    Multigrid::FcgGraph* gr = nullptr;
    if (mg->use_graphs) {
        const float* ls = mg->has_factor ? mg->factor.Ls : nullptr;
        const float* lst = mg->has_factor ? mg->factor.LsT : nullptr;
        for (auto& g : mg->graphs)
            if (g.exec && g.precond == precond && g.fine_smoother == mg->fine_smoother && g.check == check_every &&
                g.degree == mg->cheb_degree && g.kind == mg->smoother_kind && g.b == b && g.x == x && g.ls == ls &&
                g.lst == lst && g.waves == mg->fine.waves_per_block && g.val0 == l0.d_val &&
                g.half == (int)cycle_half(mg))
                gr = &g;
        if (!gr) {
            // capture into the least recently used slot
            gr = &mg->graphs[0];
            for (auto& g : mg->graphs)
                if (g.use < gr->use) gr = &g;
            if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
            gr->exec = nullptr;
            hipGraph_t graph = nullptr;
            MGCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            iterations();
            hipError_t e1 = hipGetLastError();
            hipError_t e2 = hipStreamEndCapture(st, &graph);
            if (e1 != hipSuccess || e2 != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return mgfail(mg, ASEND_ERR_HIP, "capturing the FCG iteration graph failed");
            }
            hipError_t e3 = hipGraphInstantiate(&gr->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e3 != hipSuccess) {
                gr->exec = nullptr;
                return mgfail(mg, ASEND_ERR_HIP, "instantiating the FCG iteration graph failed");
            }
            gr->precond = precond;
            gr->fine_smoother = mg->fine_smoother;
            gr->check = check_every;
            gr->degree = mg->cheb_degree;
            gr->kind = mg->smoother_kind;
            gr->b = b;
            gr->x = x;
            gr->ls = ls;
            gr->lst = lst;
            gr->waves = mg->fine.waves_per_block;
            gr->val0 = l0.d_val;
            gr->half = (int)cycle_half(mg);
        }
        gr->use = ++mg->graph_clock;
    }
    int64_t it = 0;
    for (;;) {
        if (gr) MGCHK(hipGraphLaunch(gr->exec, st));
        else iterations();
        MGCHK(hipGetLastError());
        MGCHK(hipMemcpyAsync(mg->h_sv, mg->sv, SV_N * sizeof(double), hipMemcpyDeviceToHost, st));
        MGCHK(hipStreamSynchronize(st));
        it = (int64_t)mg->h_sv[SV_IT];
        if (mg->h_sv[SV_DONE] != 0.0) break;
        if (ru && mg->h_sv[SV_RR] <= ru->next) {
            double ratio = 1.0, next = ru->next;
            bool conv = false;
            const int rc = ru->run(&ratio, &next, &conv);
            if (rc) return rc;
            ru->refreshes += 1;
            if (conv) {
                ru->converged = true;
                break;
            }
            // r = b (the FP64 residual, normalised); p keeps its direction in
            // the new normalisation
            hipLaunchKernelGGL(k_copy, dim3(nb_for(n)), dim3(256), 0, st, n, b, mg->r);
            hipLaunchKernelGGL(k_scal, dim3(nb_for(n)), dim3(256), 0, st, n, (float)ratio, mg->p);
            MGCHK(hipGetLastError());
            ru->next = next;
        }
    }
    if (it_out) *it_out = it;
    if (!(mg->h_sv[SV_RR] == mg->h_sv[SV_RR]))
        return mgfail(mg, ASEND_ERR_NUMERIC, "multigrid FCG produced a non-finite residual");
    return ASEND_OK;
}

// One V-cycle x = B b on b = A e with the learned smoother, its error energy
// ratio loss = (e - x)^T A (e - x) / e^T A e, and the reverse pass that gives
// dLoss/dz_k for every use z_k = P v_k of the factor (2 x Chebyshev degree
// uses on the finest level; the coarse correction Q = Pr V_1 Pr^T is a fixed
// symmetric operator).  The Chebyshev coefficients are held fixed.
int mg_vcycle_uses(Multigrid* mg, hipStream_t st, const float* e, FactorUses* uses, double* loss) {
    if (!mg->ready || !mg_factor_has_bounds(mg)) return mgfail(mg, ASEND_ERR_STATE, "learned smoother not set up");
    const int D = mg->cheb_degree;
    if (2 * D > kMaxUses) return mgfail(mg, ASEND_ERR_UNSUPPORTED, "Chebyshev degree too high for training");
    Level& lv = mg->L[0];
    Level& cv = mg->L[1];
    const int64_t n = lv.n_dof;
    if ((int)mg->tv.size() < 2 * D) {
        for (int k = (int)mg->tv.size(); k < 2 * D; ++k) {
            float *a = nullptr, *b = nullptr;
            MGCHK(hipMalloc((void**)&a, n * sizeof(float)));
            MGCHK(hipMalloc((void**)&b, n * sizeof(float)));
            mg->tv.push_back(a);
            mg->tz.push_back(b);
        }
    }
    if (!mg->tx) {
        for (float** q : {&mg->tx, &mg->td, &mg->txb, &mg->tdb, &mg->tb, &mg->tq})
            MGCHK(hipMalloc((void**)q, n * sizeof(float)));
    }
    const int nv = nb_for(n);
    // Chebyshev coefficients of one smoothing call (the same for pre and post)
    const double lmax = mg->lmax_factor;
    const double hi = lmax * mg->cheb_hi, lo = lmax * mg->cheb_lo;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    std::vector<float> c1(D), c2(D);
    double rho0 = 1.0 / sigma;
    c1[0] = 0.f;
    c2[0] = (float)(1.0 / theta);
    for (int k = 1; k < D; ++k) {
        const double rho1 = 1.0 / (2.0 * sigma - rho0);
        c1[k] = (float)(rho1 * rho0);
        c2[k] = (float)(2.0 * rho1 / delta);
        rho0 = rho1;
    }
    // forward: b = A e, x = B b, recording the inputs of P
    apply_level(mg, st, 0, e, mg->tb);
    hipLaunchKernelGGL(k_fill, dim3(nv), dim3(256), 0, st, n, 0.f, mg->tx);
    hipLaunchKernelGGL(k_fill, dim3(nv), dim3(256), 0, st, n, 0.f, mg->td);
    for (int j = 0; j < 2 * D; ++j) {
        const int k = j % D;
        if (j == D) {
            // coarse correction x += Q (b - A x)
            residual_level(mg, st, 0, mg->tb, mg->tx, lv.r);
            restrict_to(mg, st, 0, lv.r, cv.b);
            vcycle(mg, st, 1, cv.b, cv.x);
            prolong_add(mg, st, 0, cv.x, mg->tx);
        }
        if (j == 0) hipLaunchKernelGGL(k_copy, dim3(nv), dim3(256), 0, st, n, mg->tb, mg->tv[0]);
        else residual_level(mg, st, 0, mg->tb, mg->tx, mg->tv[j]);
        launch_factor_apply(st, mg->factor, mg->tv[j], mg->ft, mg->fz);
        hipLaunchKernelGGL(k_cheb_step, dim3(nv), dim3(256), 0, st, n, (const float*)nullptr, mg->fz, c1[k], c2[k],
                           mg->td, mg->tx);
    }
    // loss: e' = e - x, w = A e'
    hipLaunchKernelGGL(k_lincomb, dim3(nv), dim3(256), 0, st, n, 1.f, e, -1.f, mg->tx, mg->tq);
    apply_level(mg, st, 0, mg->tq, mg->txb);
    hipLaunchKernelGGL(k_dot2_partials, dim3(kVecBlocks), dim3(kBlock), 0, st, n, mg->tq, mg->txb, e, mg->tb,
                       mg->partials);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, st, mg->partials, kVecBlocks, 2, mg->sv + SV_N);
    MGCHK(hipMemcpyAsync(mg->h_sv, mg->sv + SV_N, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    MGCHK(hipStreamSynchronize(st));
    const double num = mg->h_sv[0], den = mg->h_sv[1];
    if (!(den > 0.0) || !std::isfinite(num)) return mgfail(mg, ASEND_ERR_NUMERIC, "V-cycle training loss not finite");
    *loss = num / den;
    // reverse pass: xbar = dLoss/dx = -2 A e' / e^T A e, dbar = 0
    hipLaunchKernelGGL(k_scale, dim3(nv), dim3(256), 0, st, n, mg->txb, (float)(-2.0 / den), mg->txb);
    hipLaunchKernelGGL(k_fill, dim3(nv), dim3(256), 0, st, n, 0.f, mg->tdb);
    for (int j = 2 * D - 1; j >= 0; --j) {
        const int k = j % D;
        hipLaunchKernelGGL(k_cheb_adjoint, dim3(nv), dim3(256), 0, st, n, c1[k], c2[k], mg->txb, mg->tdb, mg->tz[j]);
        if (j > 0) {
            // v_j = b - A x_prev:  xbar -= A P zbar_j (A: the cycle's operator)
            launch_factor_apply(st, mg->factor, mg->tz[j], mg->ft, mg->fz);
            apply_cycle_op(mg, st, 0, mg->fz, mg->tq);
            hipLaunchKernelGGL(k_lincomb, dim3(nv), dim3(256), 0, st, n, 1.f, mg->txb, -1.f, mg->tq, mg->txb);
        }
        if (j == D) {
            // through x += Q (b - A x):  xbar -= A Q xbar
            restrict_to(mg, st, 0, mg->txb, cv.b);
            vcycle(mg, st, 1, cv.b, cv.x);
            hipLaunchKernelGGL(k_fill, dim3(nv), dim3(256), 0, st, n, 0.f, mg->tq);
            prolong_add(mg, st, 0, cv.x, mg->tq);
            apply_cycle_op(mg, st, 0, mg->tq, lv.r);
            hipLaunchKernelGGL(k_lincomb, dim3(nv), dim3(256), 0, st, n, 1.f, mg->txb, -1.f, lv.r, mg->txb);
        }
    }
    MGCHK(hipGetLastError());
    uses->m = 2 * D;
    for (int j = 0; j < 2 * D; ++j) {
        uses->v[j] = mg->tv[j];
        uses->zbar[j] = mg->tz[j];
    }
    return ASEND_OK;
}

int mg_levels(const Multigrid* mg) { return (int)mg->L.size(); }
int64_t mg_level_ndof(const Multigrid* mg, int l) { return mg->L[l].n_dof; }

int mg_apply_level(Multigrid* mg, hipStream_t st, int l, const float* d_x, float* d_y) {
    if (l < 0 || l >= (int)mg->L.size()) return mgfail(mg, ASEND_ERR_ARG, "bad level");
    return apply_level(mg, st, l, d_x, d_y);
}

int mg_restrict(Multigrid* mg, hipStream_t st, int l, const float* d_fine, float* d_coarse) {
    if (l < 0 || l + 1 >= (int)mg->L.size()) return mgfail(mg, ASEND_ERR_ARG, "bad level");
    restrict_to(mg, st, l, d_fine, d_coarse);
    return ASEND_OK;
}

int mg_prolong(Multigrid* mg, hipStream_t st, int l, const float* d_coarse, float* d_fine) {
    if (l < 0 || l + 1 >= (int)mg->L.size()) return mgfail(mg, ASEND_ERR_ARG, "bad level");
    hipLaunchKernelGGL(k_fill, dim3(nb_for(mg->L[l].n_dof)), dim3(256), 0, st, mg->L[l].n_dof, 0.f, d_fine);
    prolong_add(mg, st, l, d_coarse, d_fine);
    return ASEND_OK;
}

int mg_vcycle(Multigrid* mg, hipStream_t st, const float* d_b, float* d_x) {
    if (!mg->ready) return mgfail(mg, ASEND_ERR_STATE, "multigrid not set up");
    if (mg->fine_smoother == 0) {
        int rc = fine_bound(mg, st);
        if (rc) return rc;
    }
    vcycle(mg, st, 0, d_b, d_x);
    return ASEND_OK;
}

int mg_coarse_solve(Multigrid* mg, hipStream_t st, const float* d_b, float* d_x) {
    if (!mg->ready) return mgfail(mg, ASEND_ERR_STATE, "multigrid not set up");
    vcycle(mg, st, (int)mg->L.size() - 1, d_b, d_x);
    return ASEND_OK;
}

}  // namespace asend
